"""What the trained surrogate gains, at the workload's grid: the same points
of the three-equation sPPLT family run from noise (two independent seeds),
from the surrogate's warm start, and from the mean training target (the same
seed for every point: the ablation that separates what the network adds from
what any seed at a typical mode and power adds), batched, to the steady-state
test.  Reports per arm the finish states and round trips, and on the points
that reach steady state in both runs: the round trips saved against the cold
start, the steady power against the cold start's, and the field error against
the cold start's fields (signal and idler aligned by the free gauge phase),
next to the cold-vs-cold spread of the same numbers.

    python -m cuopo_amd.ab_surrogate --out ab.json
"""
import argparse
import json
import os
import time

import numpy as np

from . import FP32C, FP64
from . import surrogate as S
from .train_surrogate import load

HERE = os.path.dirname(os.path.abspath(__file__))


def field_errors(a, b):
    """Relative L2 error of fields a = (Ap, As, Ai) against b, after rotating
    a's signal by e^{-i phi} and idler by e^{+i phi} (the gauge the pump
    leaves free), phi from the signals' overlap."""
    phi = np.angle(np.vdot(b[1], a[1]))
    al = (a[0], a[1] * np.exp(-1j * phi), a[2] * np.exp(1j * phi))
    return [float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(al, b)]


def dominant_mode(A):
    return int(np.argmax(np.abs(np.fft.fft(A))))


def compare(arm, ref):
    """Round trips, power and fields of `arm` against `ref` (both run_family
    results) on the points steady in both; the same state means the same
    dominant signal mode, and the power and field errors are also given on
    those points alone."""
    both = np.flatnonzero((arm["states"] == 1) & (ref["states"] == 1))
    out = {"points_steady_in_both": int(both.size)}
    if not both.size:
        return out
    r_a, r_r = arm["rts"][both].astype(float), ref["rts"][both].astype(float)
    pw = np.abs(arm["power"][both] - ref["power"][both]) / ref["power"][both]
    same = np.array([dominant_mode(arm["fields"][k][1]) == dominant_mode(ref["fields"][k][1]) for k in both])
    fe = np.array([field_errors(arm["fields"][k], ref["fields"][k]) for k in both])
    out.update(rts_mean=float(r_a.mean()), ref_rts_mean=float(r_r.mean()),
               rts_saved_fraction=float(1 - r_a.sum() / r_r.sum()),
               rts_median=float(np.median(r_a)), ref_rts_median=float(np.median(r_r)),
               rel_power_median=float(np.median(pw)), rel_power_max=float(pw.max()),
               rel_power_within_1e4=int((pw <= 1e-4).sum()), same_mode=int(same.sum()))
    if same.any():
        out["same_mode_points"] = {
            "rel_power_median": float(np.median(pw[same])), "rel_power_max": float(pw[same].max()),
            "rel_power_within_1e4": int((pw[same] <= 1e-4).sum()),
            "field_err_median": {f: float(np.median(fe[same, j])) for j, f in enumerate(("pump", "signal", "idler"))},
            "field_err_max": {f: float(np.max(fe[same, j])) for j, f in enumerate(("pump", "signal", "idler"))}}
    return out


def summary(r, cap):
    st = r["states"]
    rts = np.where(r["rts"] > 0, r["rts"], cap)
    return {"steady": int((st == 1).sum()), "off": int((st == 2).sum()), "capped": int((st == 0).sum()),
            "diverged": int((st == -1).sum()), "mean_round_trips": float(rts.mean()), "run_s": r["run_s"]}


def run(weights, count=32, N=1 << 14, NZ=100, max_rt=20000, seed=7, precision=FP64):
    sur, meta = load(weights)
    pts = S.sample_points(count, seed)  # training used other seeds
    ap0 = [S.derive(S.family_args(*p, N=N, NZ=NZ)).Ap0 for p in pts]
    arms = {"cold_a": dict(seed=1000), "cold_b": dict(seed=5000),
            "warm": dict(warm=sur.warm_fields(pts, N, ap0, seed=seed)),
            "mean": dict(warm=sur.warm_fields(pts, N, ap0, seed=seed, use_mean=True))}
    res = {}
    out = {"meta": meta, "count": count, "N": N, "NZ": NZ, "max_rt": max_rt,
           "precision": "FP64" if precision == FP64 else "FP32C", "arms": {}}
    for name, kw in arms.items():
        t = time.perf_counter()
        r = S.run_family(pts, N=N, NZ=NZ, max_rt=max_rt, precision=precision, **kw)
        r["run_s"] = time.perf_counter() - t
        res[name] = r
        out["arms"][name] = summary(r, max_rt)
        print(name, json.dumps(out["arms"][name]), flush=True)
    for name in ("cold_b", "warm", "mean"):
        out["arms"][name]["vs_cold_a"] = compare(res[name], res["cold_a"])
        print(name, "vs cold_a", json.dumps(out["arms"][name]["vs_cold_a"]), flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=os.path.join(HERE, "data", "surrogate.npz"))
    ap.add_argument("--count", type=int, default=32)
    ap.add_argument("--N", type=int, default=1 << 14)
    ap.add_argument("--nz", type=int, default=100)
    ap.add_argument("--max-rt", type=int, default=20000)
    ap.add_argument("--fp32c", action="store_true")
    ap.add_argument("--out", default="")
    a = ap.parse_args(argv)
    out = run(a.weights, a.count, a.N, a.nz, a.max_rt, precision=FP32C if a.fp32c else FP64)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
