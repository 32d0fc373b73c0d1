"""Train the warm-start surrogate on steady states the solver itself produces.

    python -m cuopo_amd.train_surrogate --points 4096 --out cuopo_amd/data/surrogate.npz

Steady states come from batched FP32-complex sweeps (grid 4096, 25 slices per
pass) over the surrogate's input ranges, seeded with noise large enough that
an above-threshold point builds up within a few hundred round trips (the
steady state is an attractor and does not depend on the seed amplitude).
Only points that finished (steady or off) are kept.
Training runs on the GPU through opo_mlp_* (BF16 MFMA, FP32 accumulation,
Adam).  The output file holds the FP32 master weights and the run's
bookkeeping (timings, losses, how many points reached steady state).
"""
import argparse
import json
import os
import time

import numpy as np

from . import FP32C
from . import surrogate as S


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--N", type=int, default=4096)
    ap.add_argument("--nz", type=int, default=25)
    ap.add_argument("--max-rt", type=int, default=4000)
    ap.add_argument("--seed-noise", type=float, default=1e-4)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                  "surrogate.npz"))
    ap.add_argument("--dataset-out", default="")
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    X, Y, scales, R, ST = S.generate_dataset(a.points, seed=a.seed, N=a.N, NZ=a.nz, max_rt=a.max_rt,
                                             check_every=50, tol=1e-6, precision=FP32C,
                                             noise_sigma=a.seed_noise)
    t_gen = time.perf_counter() - t0
    print(f"dataset: {X.shape[0]} finished of {a.points} points in {t_gen:.1f} s "
          f"({int((ST == 1).sum())} steady, {int((ST == 2).sum())} off)", flush=True)
    if a.dataset_out:
        np.savez_compressed(a.dataset_out, X=X, Y=Y, S=scales, R=R, ST=ST)
    n_val = max(1, X.shape[0] // 10)
    sur = S.Surrogate(seed=a.seed, max_batch=a.batch)
    t1 = time.perf_counter()
    hist = sur.fit(X[:-n_val], Y[:-n_val], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                   log=lambda e, l: print(f"epoch {e} loss {l:.3e}", flush=True))
    t_train = time.perf_counter() - t1
    pred = sur.predict(X[-n_val:])
    Yv = Y[-n_val:]
    nb = Y.shape[1] // 4
    val = float(np.mean((pred - Yv) ** 2))
    var = float(np.mean((Yv - Y[:-n_val].mean(axis=0)) ** 2))
    # the power part alone: predicted vs true log10(P/Ap0^2) of the signal
    lp_true = 10 * Yv[:, nb:2 * nb].mean(axis=1)
    lp_pred = 10 * pred[:, nb:2 * nb].mean(axis=1)
    lp_mean = 10 * Y[:-n_val, nb:2 * nb].mean()
    lp_rmse = float(np.sqrt(np.mean((lp_pred - lp_true) ** 2)))
    lp_rmse_mean = float(np.sqrt(np.mean((lp_mean - lp_true) ** 2)))
    print(f"train {t_train:.1f} s; validation MSE {val:.3e} (predicting the training mean: {var:.3e}); "
          f"signal log10 power RMSE {lp_rmse:.3f} (training mean: {lp_rmse_mean:.3f})", flush=True)
    meta = dict(target=S.TARGET_KIND, points=int(a.points), finished=int(X.shape[0]),
                steady=int((ST == 1).sum()), off=int((ST == 2).sum()), N=a.N, NZ=a.nz, max_rt=a.max_rt,
                seed_noise=a.seed_noise, epochs=a.epochs, batch=a.batch, lr=a.lr, dataset_s=t_gen,
                train_s=t_train, train_loss_first=float(hist[0]), train_loss_last=float(hist[-1]),
                val_mse=val, val_mse_of_training_mean=var, val_log10_power_rmse=lp_rmse,
                val_log10_power_rmse_of_training_mean=lp_rmse_mean)
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
