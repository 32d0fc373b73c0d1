"""What the trained surrogate gains: the same sweep points run from noise
(two independent seeds), from the surrogate's warm start, and from the mean
training target (the same seed for every point: the ablation that separates
what the network adds from what any plausible seed at the right scale adds),
FP64, batched, to the steady-state test.  Reports round trips per point,
finish states and the steady-state power of each warm start against the cold
start, next to the cold-vs-cold spread that bounds any warm-start accuracy
claim.

    python -m cuopo_amd.ab_surrogate --weights cuopo_amd/data/surrogate.npz
"""
import argparse
import json
import time

import numpy as np

from . import FP64
from . import sweep as SW
from .surrogate import derive, sweep_args
from .train_surrogate import load


def run(weights, count=32, N=4096, NZ=25, max_rt=20000, seed=7, precision=FP64):
    sur, meta = load(weights)
    rng = np.random.default_rng(seed)
    pts = []
    for k in range(count):
        nth = rng.uniform(1.2, 5.0)
        delta = rng.uniform(-np.pi, np.pi)
        lcr = rng.uniform(5, 40)
        gdd = rng.uniform(-500, 500)
        R = rng.uniform(0.90, 0.99)
        pts.append((N, nth, delta, lcr, gdd, R))
    ap0 = [derive(sweep_args(*p[1:], N=N, NZ=NZ)).Ap0 for p in pts]
    warm = sur.warm_fields([p[1:] for p in pts], N, ap0, seed=seed)
    mean = sur.warm_fields([p[1:] for p in pts], N, ap0, seed=seed, use_mean=True)
    out = {"meta": meta, "points": pts}
    for name, kw in (("cold_a", dict(seed=1000)), ("cold_b", dict(seed=5000)), ("warm", dict(warm=warm)),
                     ("mean", dict(warm=mean))):
        if "warm" in kw:  # warm starts fall back to cold_a's start when they diverge
            r = SW.run_warm(pts, kw["warm"], seed=1000, NZ=NZ, max_rt=max_rt, precision=precision)
        else:
            r = SW.run_batched(pts, NZ=NZ, max_rt=max_rt, precision=precision, **kw)
        out[name] = dict(rts=r["rts"].tolist(), states=r["states"].tolist(), power=r["power"].tolist(),
                         run_s=r["run_s"], cold_fallbacks=int(len(r.get("fallback", ()))))
        r["batch"].close()
        if r.get("fallback_batch") is not None:
            r["fallback_batch"].close()
        print(name, "run_s", round(r["run_s"], 2), "states", np.bincount(r["states"] + 1, minlength=4).tolist(),
              flush=True)
    ca, cb, w, mn = (np.array(out[k]["power"]) for k in ("cold_a", "cold_b", "warm", "mean"))
    sa, sb, sw, sm = (np.array(out[k]["states"]) for k in ("cold_a", "cold_b", "warm", "mean"))
    ra, rw, rm = np.array(out["cold_a"]["rts"]), np.array(out["warm"]["rts"]), np.array(out["mean"]["rts"])
    ok = (sa == 1) & (sw == 1)
    okm = (sa == 1) & (sm == 1)
    okc = (sa == 1) & (sb == 1)
    summ = dict(
        points=count, steady_cold=int((sa == 1).sum()), steady_warm=int((sw == 1).sum()),
        off_cold=int((sa == 2).sum()), off_warm=int((sw == 2).sum()),
        same_outcome=int((sa == sw).sum()),
        mean_rts_cold=float(np.mean(np.where(ra > 0, ra, max_rt))),
        mean_rts_warm=float(np.mean(np.where(rw > 0, rw, max_rt))),
        mean_rts_mean_seed=float(np.mean(np.where(rm > 0, rm, max_rt))),
        steady_mean_seed=int((sm == 1).sum()), off_mean_seed=int((sm == 2).sum()),
        same_outcome_mean_seed=int((sa == sm).sum()),
        mean_rts_cold_steady=float(np.mean(ra[sa == 1])) if (sa == 1).any() else None,
        mean_rts_warm_on_cold_steady=float(np.mean(np.where(rw > 0, rw, max_rt)[sa == 1])) if (sa == 1).any() else None,
        rel_power_warm_vs_cold=(np.abs(w - ca) / ca)[ok].tolist(),
        rel_power_mean_seed_vs_cold=(np.abs(mn - ca) / ca)[okm].tolist(),
        rel_power_cold_b_vs_cold_a=(np.abs(cb - ca) / ca)[okc].tolist(),
        run_s={k: out[k]["run_s"] for k in ("cold_a", "cold_b", "warm", "mean")},
    )
    out["summary"] = summ
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", required=True)
    ap.add_argument("--count", type=int, default=32)
    ap.add_argument("--max-rt", type=int, default=20000)
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    t = time.perf_counter()
    res = run(a.weights, count=a.count, max_rt=a.max_rt)
    res["wall_s"] = time.perf_counter() - t
    print(json.dumps(res["summary"]))
    if a.out:
        json.dump(res, open(a.out, "w"))


if __name__ == "__main__":
    main()
