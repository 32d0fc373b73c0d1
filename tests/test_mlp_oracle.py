"""CPU pins for the MLP oracle (oracle/mlp_oracle.py) and the surrogate's
spectrum encodings."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

SHAPE = (5, 16, 3, 12)


def test_bf16_rounding_known_answers():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 2 ** -8) == 1.0          # tie -> even (mantissa ends in 0)
    assert M.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6  # tie -> even upward
    assert M.bf16(1.0 + 2 ** -7) == 1.0 + 2 ** -7  # representable
    assert M.bf16(-3.14159) == -3.140625


def test_gradient_matches_finite_differences():
    """Backprop of the unquantised FP64 network against central differences
    of its loss: pins the gradient formula (transposes, GELU', bias sums)."""
    rng = np.random.default_rng(0)
    _, _, _, n = M.layers(*SHAPE)
    theta = rng.standard_normal(n) * 0.3
    x = rng.standard_normal((7, SHAPE[0]))
    t = rng.standard_normal((7, SHAPE[3]))
    loss, g = M.loss_and_grad(theta, SHAPE, x, t, quantize=False)
    eps = 1e-6
    for i in rng.choice(n, 40, replace=False):
        tp = theta.copy(); tp[i] += eps
        tm = theta.copy(); tm[i] -= eps
        fd = (M.loss_and_grad(tp, SHAPE, x, t, False)[0] - M.loss_and_grad(tm, SHAPE, x, t, False)[0]) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_gelu_derivative():
    z = np.linspace(-4, 4, 101)
    fd = (M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(M.gelu_grad(z), fd, atol=1e-8)


def test_spectrum_encoding_round_trip():
    import cuopo_amd.surrogate as S
    rng = np.random.default_rng(1)
    N = 4096
    spec = np.zeros(N, complex)
    spec[N // 2 - 100:N // 2 + 100] = rng.standard_normal(200) + 1j * rng.standard_normal(200)
    A = np.fft.ifft(np.fft.ifftshift(spec)) * N
    c = S.coarse_spectrum(A)
    assert c.size == 1024
    assert np.allclose(S.field_from_coarse(c, N), A, atol=1e-12)
    # gauge fix keeps the signal-idler phase sum and makes the peak real positive
    I = rng.standard_normal(1024) + 1j * rng.standard_normal(1024)
    s2, i2 = S.gauge_fix(c, I)
    k = np.argmax(np.abs(c))
    assert abs(s2[k].imag) < 1e-12 and s2[k].real > 0
    assert np.allclose(s2 * i2, c * I)
    y = S.encode_target(A, 0.5 * A, 2.0)
    s3, i3 = S.decode_target(y, 2.0)
    assert np.allclose(s3, np.abs(c), rtol=1e-5, atol=1e-6 * np.abs(c).max())
    assert np.allclose(i3, 0.5 * np.abs(c), rtol=1e-5, atol=1e-6 * np.abs(c).max())
    # shape halves have unit norm, power halves are constant
    nb = 1024
    assert abs(np.linalg.norm(y[:nb]) - 1) < 1e-5 and np.ptp(y[nb:2 * nb]) == 0
    # an off point: zero shape, power floor, decodes to zero fields
    z = S.encode_target(np.zeros(N, complex), np.zeros(N, complex), 2.0)
    assert np.all(z[:nb] == 0) and np.allclose(z[nb:2 * nb], S.LOG_FLOOR / 10)
    s4, i4 = S.decode_target(z, 2.0)
    assert not s4.any() and not i4.any()


def test_feature_encoding_range():
    import cuopo_amd.surrogate as S
    lo = S.encode_features(*S.RANGES[:, 0])
    hi = S.encode_features(*S.RANGES[:, 1])
    assert np.allclose(lo, -1) and np.allclose(hi, 1)
