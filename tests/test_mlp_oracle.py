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
    """Coarse spectrum <-> field, and the compact steady-state target in the
    resonance frame decodes back to the same carrier modes and powers."""
    import cuopo_amd.surrogate as S
    rng = np.random.default_rng(1)
    N = 4096
    spec = np.zeros(N, complex)
    spec[N // 2 - 100:N // 2 + 100] = rng.standard_normal(200) + 1j * rng.standard_normal(200)
    A = np.fft.ifft(np.fft.ifftshift(spec)) * N
    c = S.coarse_spectrum(A)
    assert c.size == 1024
    assert np.allclose(S.field_from_coarse(c, N), A, atol=1e-12)
    # a single mode pair (m, -m) = (37, -37) with a weak neighbour
    j = np.arange(N)
    As = 0.3 * np.exp(2j * np.pi * 37 * j / N) + 0.001 * np.exp(2j * np.pi * 38 * j / N)
    Ai = 0.25 * np.exp(-2j * np.pi * 37 * j / N + 0.4j)
    bs, a_s, P_s = S.top_modes(As)
    bi, a_i, P_i = S.top_modes(Ai)
    assert bs[0] == 512 + 37 and bi[0] == 512 - 37
    assert abs(P_s - (0.09 + 1e-6)) < 1e-12 and abs(P_i - 0.0625) < 1e-12
    nb = 1024
    for m0 in (0, 30, -5):
        y = S.encode_target(bs, a_s, P_s, bi, a_i, P_i, scale=2.0, m0=m0)
        # shapes have unit norm and sit at the offset from the resonance pair
        assert abs(np.linalg.norm(y[:nb]) - 1) < 1e-6 and np.argmax(y[:nb]) == 512 + 37 - m0
        assert np.argmax(y[2 * nb:3 * nb]) == 512 - 37 + m0
        assert np.ptp(y[nb:2 * nb]) == 0  # replicated log power
        (ds, das), (di, dai) = S.decode_target(y, 2.0, m0)
        assert (ds, di) == (512 + 37, 512 - 37)
        assert abs(das - np.sqrt(P_s)) < 1e-6 * np.sqrt(P_s) and abs(dai - 0.25) < 1e-6


def test_resonance_mode_cancels_the_detuning_phase():
    """The resonance pair's idler walk-off phase per pass cancels the two
    mirror phases 2 delta (mod 2 pi) to within half a mode, on the branch of
    smallest |m|, for points of the sPPLT family."""
    import math

    import cuopo_amd.surrogate as S
    for pt in S.sample_points(50, seed=9):
        d = S.derive(S.family_args(*pt, N=4096, NZ=25))
        m0 = S.resonance_mode(d, pt[1])
        c = (1 / d.vi - 1 / d.vs) * d.Lcr * 2 * math.pi / d.t_rt * d.N / (d.N - 1)
        r = (2 * pt[1] + c * m0 + math.pi) % (2 * math.pi) - math.pi
        assert abs(r) <= 0.5 * abs(c) + 1e-12
        assert abs(c * m0) <= math.pi + 0.5 * abs(c)


def test_feature_encoding_range():
    import cuopo_amd.surrogate as S
    lo = S.encode_features(*S.RANGES[:, 0])
    hi = S.encode_features(*S.RANGES[:, 1])
    assert np.allclose(lo, -1) and np.allclose(hi, 1)
