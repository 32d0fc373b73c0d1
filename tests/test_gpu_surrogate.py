"""What the trained surrogate gains over the method without it (GPU).

The committed network (cuopo_amd/data/surrogate.npz, trained on the GPU by
python -m cuopo_amd.train_surrogate) is run on sweep points it never saw: the
training generator's settings with another seed.  Without the network, the
best input-independent guess of a point's steady state is the mean training
target (the ablation the A/B uses).  The network's gain is in the steady
intracavity power: its predicted log10(P / Ap0^2) must beat that mean's.  (The
spectral shape is multistable and not predictable from the inputs, and the
warm start saves no round trips; both are measured by the A/B script.)"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuopo_amd", "data",
                       "surrogate.npz")


def test_network_predicts_steady_power_better_than_the_mean_target(gpu):
    from cuopo_amd import FP32C
    from cuopo_amd import surrogate as S
    from cuopo_amd.train_surrogate import load
    sur, meta = load(WEIGHTS, max_batch=64)
    # the training generator's settings (meta), fresh points: training used seed 1
    X, Y, scales, R, ST = S.generate_dataset(64, seed=2024, N=meta["N"], NZ=meta["NZ"], max_rt=meta["max_rt"],
                                             check_every=50, tol=1e-6, precision=FP32C,
                                             noise_sigma=meta["seed_noise"])
    assert X.shape[0] >= 16, f"only {X.shape[0]} of 64 points finished"
    nb = Y.shape[1] // 4
    true = 10 * Y[:, nb:2 * nb].mean(axis=1)
    pred = 10 * sur.predict(X)[:, nb:2 * nb].mean(axis=1)
    mean = 10 * float(sur.mean_y[nb:2 * nb].mean())
    rmse_net = float(np.sqrt(np.mean((pred - true) ** 2)))
    rmse_mean = float(np.sqrt(np.mean((mean - true) ** 2)))
    print(f"finished {X.shape[0]} of 64 ({int((ST == 1).sum())} steady, {int((ST == 2).sum())} off); "
          f"signal log10 power RMSE: network {rmse_net:.3f}, mean target {rmse_mean:.3f} decades")
    sur.net.close()
    assert np.isfinite(pred).all()
    assert rmse_net < rmse_mean
