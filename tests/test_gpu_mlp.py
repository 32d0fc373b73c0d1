"""The surrogate MLP's BF16-MFMA kernels against the FP64 CPU oracle
(oracle/mlp_oracle.py) on identical BF16 operands."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def make(shape, batch, seed):
    from cuopo_amd.surrogate import MLP
    net = MLP(n_in=shape[0], width=shape[1], depth=shape[2], n_out=shape[3], max_batch=batch, seed=seed)
    rng = np.random.default_rng(seed)
    theta = (rng.standard_normal(net.nparams) * 0.2).astype(np.float32)
    net.set_weights(theta)
    return net, theta.astype(np.float64), rng


@pytest.mark.parametrize("shape,batch", [((5, 64, 1, 80), 33), ((5, 96, 2, 70), 64), ((5, 256, 4, 4096), 128)])
def test_forward_matches_fp64_oracle(gpu, shape, batch):
    """Ragged shapes (not multiples of the 64x64 tile) exercise the MFMA lane
    layout and the tile edges."""
    net, theta, rng = make(shape, batch, 1)
    x = rng.uniform(-1, 1, (batch, shape[0])).astype(np.float32)
    y = net.forward(x)
    ref, _, _ = M.forward(theta, shape, x, quantize=True)
    assert rel(y, ref) < 5e-3


@pytest.mark.parametrize("shape,batch", [((5, 64, 2, 80), 48), ((5, 256, 4, 4096), 128)])
def test_gradients_match_fp64_oracle(gpu, shape, batch):
    net, theta, rng = make(shape, batch, 2)
    x = rng.uniform(-1, 1, (batch, shape[0])).astype(np.float32)
    t = rng.standard_normal((batch, shape[3])).astype(np.float32)
    loss, g = net.gradients(x, t)
    rloss, rg = M.loss_and_grad(theta, shape, x, t, quantize=True)
    assert loss == pytest.approx(rloss, rel=1e-3)
    cos = float(np.dot(g, rg) / np.linalg.norm(g) / np.linalg.norm(rg))
    print(f"shape={shape} grad rel L2={rel(g, rg):.2e} cos={cos:.6f}")
    assert cos > 0.999
    assert rel(g, rg) < 3e-2
    # per layer: weight and bias gradients separately
    din, dout, offs, _ = M.layers(*shape)
    for (wo, bo), a, b in zip(offs, din, dout):
        assert rel(g[wo:wo + a * b], rg[wo:wo + a * b]) < 5e-2
        assert rel(g[bo:bo + b], rg[bo:bo + b]) < 5e-2


def test_training_fits_a_teacher(gpu):
    """Adam on the GPU drives the loss on a smooth synthetic regression (a
    random teacher network) down by more than 20x."""
    from cuopo_amd.surrogate import MLP
    shape = (5, 128, 3, 256)
    teacher, ttheta, rng = make(shape, 256, 3)
    X = rng.uniform(-1, 1, (1024, 5)).astype(np.float32)
    Y = teacher.forward(X)
    net = MLP(n_in=5, width=128, depth=3, n_out=256, max_batch=256, seed=7)
    first = net.train_step(X[:256], Y[:256], 1e-3)
    for ep in range(150):
        for s in range(0, 1024, 256):
            last = net.train_step(X[s:s + 256], Y[s:s + 256], 1e-3)
    print("loss", first, "->", last)
    assert last < first / 20


def test_mlp_closes_on_leaving_a_with_block(gpu):
    from cuopo_amd.surrogate import MLP
    with MLP(n_in=5, width=64, depth=1, n_out=80, max_batch=8, seed=3) as net:
        y = net.forward(np.zeros((8, 5), np.float32))
        assert np.isfinite(y).all()
    assert net._h is None
