"""Train the warm-start surrogate on steady states the solver itself produces.

    python -m cuopo_amd.train_surrogate --points 13312 \
        --dataset-out cuopo_amd/data/surrogate_dataset.npz --out cuopo_amd/data/surrogate.npz
    python -m cuopo_amd.train_surrogate --dataset cuopo_amd/data/surrogate_dataset.npz   # retrain only
    python -m cuopo_amd.train_surrogate --dataset cuopo_amd/data/surrogate_dataset.npz --append 3328 \
        --dataset-out cuopo_amd/data/surrogate_dataset.npz                             # more points, retrain

Steady states come from batched FP32-complex cold starts (seeded 1e-20 noise,
as the workload starts) of random points of the three-equation sPPLT family
(cuopo_amd.surrogate), run to the convergence test: intracavity power change
below 1e-6 over 50 round trips, cap 20,000 round trips.  Only points that
reach that steady state are kept.  The grid is reduced from the workload's
2^14 samples x 100 slices to 4096 x 25: a cold start needs ~10^4 round trips,
so 10^4 steady states at the full grid would take about 80 GPU-minutes, and
the quantities the network predicts (the lasing mode pair in units of the
free spectral range, and the power per sample) do not depend on the grid
beyond the discretisation error (the modes lie within +-512 bins either way).
Training runs on the GPU through opo_mlp_* (BF16 MFMA, FP32 accumulation,
Adam).  The output file holds the FP32 master weights, the mean training
target (the ablation seed) and the run's bookkeeping.
"""
import argparse
import json
import os
import time

import numpy as np

from . import FP32C
from . import surrogate as S

HERE = os.path.dirname(os.path.abspath(__file__))


def validate(sur, X, ds, idx, mean_y, N, NZ):
    """Mode and power accuracy of the network (and of the mean target) on the
    points idx of a dataset."""
    sub = {k: ds[k][idx] for k in ("bins_s", "amp_s", "P_s", "bins_i", "amp_i", "P_i", "scale", "points")}
    Y = S.targets(sub, N, NZ)
    pred = sur.predict(X[idx])
    nb = S.NBINS
    true_bin = np.argmax(Y[:, :nb], axis=1)  # the dominant signal bin in the resonance frame
    out = {}
    for name, P in (("network", pred), ("mean_target", np.tile(mean_y, (len(idx), 1)))):
        pb = np.argmax(P[:, :nb], axis=1)
        dm = np.abs(pb - true_bin)
        lp_true = 10 * Y[:, nb:2 * nb].mean(axis=1)
        lp_pred = 10 * P[:, nb:2 * nb].mean(axis=1)
        out[name] = {"signal_mode_exact": float(np.mean(dm == 0)), "signal_mode_within_1": float(np.mean(dm <= 1)),
                     "signal_mode_within_3": float(np.mean(dm <= 3)),
                     "log10_power_rmse": float(np.sqrt(np.mean((lp_pred - lp_true) ** 2))),
                     "mse": float(np.mean((P - Y) ** 2))}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=13312)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--nz", type=int, default=25)
    ap.add_argument("--max-rt", type=int, default=20000)
    ap.add_argument("--noise", type=float, default=1e-20)
    ap.add_argument("--chunk", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--dataset", default="", help="train on this dataset instead of generating one")
    ap.add_argument("--append", type=int, default=0,
                    help="with --dataset: run this many more points (seed + 1) and add their steady states")
    ap.add_argument("--dataset-out", default="")
    ap.add_argument("--out", default=os.path.join(HERE, "data", "surrogate.npz"))
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    if a.dataset:
        z = np.load(a.dataset, allow_pickle=False)
        ds = {k: z[k] for k in z.files}
        gen = json.loads(str(ds.pop("meta")))
        if a.append > 0:
            seed2 = gen.get("seeds", [gen["seed"]])[-1] + 1
            more = S.generate_dataset(a.append, seed=seed2, chunk=a.chunk, N=gen["N"], NZ=gen["NZ"],
                                      max_rt=gen["max_rt"], noise_sigma=gen["noise"], precision=FP32C,
                                      log=lambda n, c: print(f"dataset: {n} more points run, {c}", flush=True))
            for k, v in more.items():
                if v.ndim >= 1:
                    ds[k] = np.concatenate([ds[k], v]) if ds[k].size else v
            gen["points"] += a.append
            gen["seeds"] = gen.get("seeds", [gen["seed"]]) + [seed2]
            gen["dataset_s"] += time.perf_counter() - t0
            for k in ("steady", "off", "capped", "diverged"):
                gen[k] += int(more[k])
            if a.dataset_out:
                np.savez_compressed(a.dataset_out, meta=json.dumps(gen), **ds)
    else:
        ds = S.generate_dataset(a.points, seed=a.seed, chunk=a.chunk, N=a.N, NZ=a.nz, max_rt=a.max_rt,
                                noise_sigma=a.noise, precision=FP32C,
                                log=lambda n, c: print(f"dataset: {n} points run, {c}", flush=True))
        gen = dict(points=a.points, N=a.N, NZ=a.nz, max_rt=a.max_rt, noise=a.noise, seed=a.seed,
                   dataset_s=time.perf_counter() - t0,
                   **{k: int(ds[k]) for k in ("steady", "off", "capped", "diverged")})
        if a.dataset_out:
            np.savez_compressed(a.dataset_out, meta=json.dumps(gen),
                                **{k: v for k, v in ds.items() if v.ndim >= 1})
        ds = {k: v for k, v in ds.items() if v.ndim >= 1}
    m = len(ds["scale"])
    print(f"dataset: {m} steady states of {gen['points']} points ({gen})", flush=True)
    X = ds["X"].astype(np.float32)
    Y = S.targets(ds, gen["N"], gen["NZ"])
    n_val = max(1, m // 10)
    tr, va = np.arange(m - n_val), np.arange(m - n_val, m)
    sur = S.Surrogate(seed=a.seed, max_batch=a.batch)
    t1 = time.perf_counter()
    hist = sur.fit(X[tr], Y[tr], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                   log=lambda e, l: print(f"epoch {e} loss {l:.3e}", flush=True))
    t_train = time.perf_counter() - t1
    val = validate(sur, X, ds, va, sur.mean_y, gen["N"], gen["NZ"])
    print("validation:", json.dumps(val), flush=True)
    meta = dict(target=S.TARGET_KIND, family="three-equation sPPLT (532 -> 1000 + 1137 nm), CW", **gen,
                steady_states=m, train=len(tr), validation=len(va), epochs=a.epochs, batch=a.batch, lr=a.lr,
                train_s=t_train, train_loss_first=float(hist[0]), train_loss_last=float(hist[-1]),
                val=val, cold_rts_median=float(np.median(ds["rts"])), cold_rts_mean=float(np.mean(ds["rts"])))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    np.savez_compressed(a.out, weights=sur.net.get_weights(), shape=np.array(sur.net.shape),
                        mean_y=sur.mean_y, meta=json.dumps(meta))
    print(json.dumps(meta))


def load(path, device=0, max_batch=256):
    """A Surrogate with weights from `path` (written by main)."""
    z = np.load(path, allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    if meta.get("target") != S.TARGET_KIND:
        raise ValueError(f"{path}: weights for target encoding {meta.get('target')!r}, "
                         f"this code decodes {S.TARGET_KIND!r}; retrain with python -m cuopo_amd.train_surrogate")
    sur = S.Surrogate(device=device, max_batch=max_batch)
    sur.net.set_weights(z["weights"])
    sur.mean_y = np.asarray(z["mean_y"], np.float32)
    return sur, meta


if __name__ == "__main__":
    main()
