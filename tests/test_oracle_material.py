"""Pins for the oracle's setup restatement (src/cuOPO.cu:92-269,
src/headers/ppln.cuh, spplt.cuh, functions.cuh:12-122) and agreement of the
product's own setup code with it.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import c1, c2, c5, to_args

CRYSTALS = [O.PPLN, O.SPPLT]


@pytest.mark.parametrize("cr", CRYSTALS)
@pytest.mark.parametrize("L", [0.532, 1.0, 1.064, 1.55, 3.1])
@pytest.mark.parametrize("T", [20.0, 60.0, 140.0])
def test_dispersion_derivatives_match_finite_differences(oracle_lib, cr, L, T):
    """dn/dL and d2n/dL2 are analytic derivatives of n (ppln.cuh:83-134,
    spplt.cuh:71-106): a sign or factor slip shows up against central
    differences of n itself."""
    lib = oracle_lib.lib()
    h = 1e-4
    n = lambda x: lib.orc_n(cr, x, T)
    fd1 = (n(L + h) - n(L - h)) / (2 * h)
    fd2 = (n(L + h) - 2 * n(L) + n(L - h)) / h ** 2
    assert lib.orc_dndl(cr, L, T) == pytest.approx(fd1, rel=1e-6)
    assert lib.orc_d2ndl2(cr, L, T) == pytest.approx(fd2, rel=1e-4)
    # group velocity and GVD follow from them (ppln.cuh:171-182)
    C = 299792458e6 / 1e12
    assert lib.orc_GV(cr, L, T) == pytest.approx(C / (n(L) - L * fd1), rel=1e-7)
    assert lib.orc_GVD(cr, L, T) == pytest.approx(L ** 3 * fd2 / (2 * math.pi * C ** 2), rel=1e-4)


@pytest.mark.parametrize("cr", CRYSTALS)
def test_refractive_index_is_physical(oracle_lib, cr):
    lib = oracle_lib.lib()
    for L in (0.532, 0.8, 1.064, 1.55, 3.0):
        n = lib.orc_n(cr, L, 40.0)
        assert 2.05 < n < 2.35
        assert lib.orc_dndl(cr, L, 40.0) < 0  # normal dispersion below the IR resonance
    # n decreases with wavelength, increases with temperature near 1 um
    assert lib.orc_n(cr, 1.0, 40.0) > lib.orc_n(cr, 1.1, 40.0)
    assert lib.orc_n(cr, 1.0, 120.0) > lib.orc_n(cr, 1.0, 20.0)


def test_phase_matching_temperature_matches_sweep_script_pair(oracle_lib):
    """src/cuOPO.sh:69-71 pairs PPLN, 1550 nm pump, grating 35.01 um with a
    phase-matching temperature of 32.88 C; the reference scan
    (functions.cuh:23-54, 5000 points over 10-150 C) must land within one
    grid step of it."""
    T = oracle_lib.lib().orc_find_pm_temp(O.PPLN, 1.55, 3.1, 35.01, 10, 150, 0)
    assert abs(T - 32.88) < 140 / 4999


def test_phase_matching_scan_first_minimum_and_grid(oracle_lib):
    """The scan returns a point of linspace(10, 150, 5000) that minimises the
    mismatch, checked by a brute-force scan in numpy over the same grid."""
    lib = oracle_lib.lib()
    for cr, lp, Lam in ((O.PPLN, 0.532, 6.97), (O.SPPLT, 0.532, 7.97), (O.PPLN, 1.064, 29.5)):
        grid = 10 + np.arange(5000) * (150 - 10) / 4999
        dk = np.array([2 * np.pi / lp * abs(lib.orc_n(cr, 2 * lp, t) - lib.orc_n(cr, lp, t)) - 2 * np.pi / Lam
                       for t in grid])
        T = lib.orc_find_pm_temp(cr, lp, 2 * lp, Lam, 10, 150, 0)
        assert T == grid[int(np.argmin(np.abs(dk)))]


def test_three_wave_scan_phase_matches(oracle_lib):
    """The opt-in three-wave scan drives 2pi(np/lp - ns/ls - ni/li - 1/Lambda)
    to its grid minimum: the residual is below one grid step's change."""
    p = O.setup(c2(N=1024, NZ=4, NRT=2))
    assert 10 < p.Temp < 150
    assert abs(p.dk) * p.Lcr < 1.0


def test_threshold_degenerate_limit(oracle_lib):
    """The three-equation DRO threshold (cuOPO.cu:217) reduces to the
    two-equation degenerate one (cuOPO.cu:222) when idler = signal."""
    cfg2 = c1(N=1024, NZ=4, NRT=2)
    p2 = O.setup(cfg2)
    cfg3 = c1(N=1024, NZ=4, NRT=2)
    cfg3.three_eqs = True
    p3 = O.setup(cfg3)
    assert p3.li == pytest.approx(p3.ls)
    # Ri is fixed at 0.98, so set R_s = 98% to make the loss terms equal
    assert p2.alphas == pytest.approx(p3.alphai)
    assert p3.Ith == pytest.approx(p2.Ith, rel=1e-12)


def test_pump_amplitude_closed_form(oracle_lib):
    """I = n eps0 c |A|^2 / 2 (cuOPO.cu:229) and P = I * pi w0^2."""
    for cfg in (c1(N=1024, NZ=4, NRT=2), c2(N=1024, NZ=4, NRT=2)):
        p = O.setup(cfg)
        eps0, C = 8.8541878128e-6, 299.792458
        assert 0.5 * p.np * eps0 * C * p.Ap0 ** 2 == pytest.approx(cfg.Nth * p.Ith, rel=1e-13)
        assert p.Power == pytest.approx(cfg.Nth * p.Pth, rel=1e-13)
        assert p.spot == pytest.approx(math.pi * 55 ** 2)


def test_grids_closed_form(oracle_lib):
    """T = linspace(-t_rt/2, t_rt/2, N); w = 2 pi fftshift(linspace(-N dF/2,
    N dF/2, N)): denominator N-1 stretches w by N/(N-1), w[0] = pi N dF/(N-1)."""
    p = O.setup(c1(N=1024, NZ=4, NRT=2))
    T, w = O.grids(p)
    N, dF = p.N, p.dF
    assert T[0] == -0.5 * p.t_rt and T[-1] == pytest.approx(0.5 * p.t_rt)
    assert np.allclose(np.diff(T), p.t_rt / (N - 1), rtol=1e-9)
    assert w[0] == pytest.approx(math.pi * N * dF / (N - 1), rel=1e-12)
    step = 2 * math.pi * N * dF / (N - 1)
    assert np.allclose(np.diff(w[: N // 2]), step, rtol=1e-9)
    assert np.allclose(np.diff(w[N // 2:]), step, rtol=1e-9)
    assert w[N // 2] == pytest.approx(-math.pi * N * dF, rel=1e-12)
    assert w[N // 2 - 1] == pytest.approx(math.pi * N * dF, rel=1e-12)


def test_ns_pump_is_gaussian(oracle_lib):
    p = O.setup(c5(N=256, NZ=4, NRT=40))
    Ap = O.pump_input(p)
    Tp, _ = O.grids_long(p)
    assert Ap.size == p.N * p.NRT
    assert np.max(np.abs(Ap.imag)) == 0
    k = np.argmax(Ap.real)
    assert Ap.real[k] == pytest.approx(p.Ap0, rel=1e-6)
    # intensity FWHM = 10 ns: |A|^2 halves at |t| = FWHM/2
    half = np.exp(-(5000.0 ** 2) / (2 * p.sigmap ** 2)) ** 2
    assert half == pytest.approx(0.5, rel=1e-12)


def test_eom_parameters(oracle_lib):
    """df = fm sqrt(Nth-1) alpha_s/(pi m) FSR, fpm = FSR - df (cuOPO.cu:265-269)."""
    cfg = c5(N=256, NZ=4, NRT=3)
    p = O.setup(cfg)
    assert p.mod_depth == pytest.approx(0.8)
    assert p.fpm == pytest.approx(p.FSR - p.df)
    assert p.df == pytest.approx(cfg.fm * math.sqrt(cfg.Nth - 1) * p.alphas / (math.pi * p.mod_depth) * p.FSR)
    cfg.Nth = 0.5
    with pytest.raises(ValueError):
        O.setup(cfg)


@pytest.mark.parametrize("make", [c1, c2, c5])
def test_product_setup_equals_oracle(oracle_lib, make):
    import cuopo_amd as P
    cfg = make(N=1024, NZ=8, NRT=5)
    o = O.setup(cfg).as_dict()
    d = P.derive(to_args(cfg)).as_dict()
    for k, v in o.items():
        if isinstance(v, float):
            assert d[k] == pytest.approx(v, rel=1e-14, abs=1e-300), k
        else:
            assert d[k] == v, k
    po = O.setup(cfg)
    Tg, wg = P.grids(P.derive(to_args(cfg)))
    To, wo = O.grids(po)
    assert np.array_equal(Tg, To) and np.array_equal(wg, wo)


def test_noise_generator(oracle_lib):
    """splitmix64 + Box-Muller: product equals oracle to 1 ulp, an
    independent numpy splitmix64 gives the same integers, moments match."""
    import cuopo_amd as P
    a = P.noise(42, 1, 1e-20, 20000)
    b = O.noise(42, 1, 1e-20, 20000)
    assert np.max(np.abs(a - b) / np.abs(b)) < 3e-16
    assert abs(np.mean(b.real)) < 3e-22 and abs(np.std(b.real) - 1e-20) < 3e-22
    assert abs(np.mean(b.imag)) < 3e-22 and abs(np.std(b.imag) - 1e-20) < 3e-22
    # independent splitmix64
    M = (1 << 64) - 1
    s = 42 ^ ((1 * 0xD1B54A32D192ED03) & M)

    def nxt():
        nonlocal s
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    for j in range(5):
        u1 = ((nxt() >> 11) + 1) * 2.0 ** -53
        u2 = ((nxt() >> 11) + 1) * 2.0 ** -53
        r = math.sqrt(-2 * math.log(u1))
        assert b[j] == pytest.approx(1e-20 * r * complex(math.cos(2 * math.pi * u2), math.sin(2 * math.pi * u2)),
                                     rel=1e-15)
    # different streams differ
    assert not np.allclose(O.noise(42, 0, 1.0, 16), O.noise(42, 1, 1.0, 16))


@pytest.mark.parametrize("cr", CRYSTALS)
def test_product_material_abi_equals_oracle(oracle_lib, cr):
    """The exported material / phase-matching C ABI (opo_n, opo_group_velocity,
    opo_gvd, opo_find_pm_temp) equals the oracle's restatement over a grid of
    wavelengths and temperatures, and its group velocity and GVD follow from
    central differences of its own opo_n (host code: no GPU needed)."""
    import cuopo_amd as P
    L_, O_ = P.lib(), oracle_lib.lib()
    C = 299792458e6 / 1e12
    h = 1e-4
    for L in (0.532, 0.8, 1.0, 1.064, 1.55, 2.2, 3.1):
        for T in (15.0, 37.5, 70.0, 120.0):
            n = L_.opo_n(cr, L, T)
            assert n == pytest.approx(O_.orc_n(cr, L, T), rel=1e-14)
            assert L_.opo_group_velocity(cr, L, T) == pytest.approx(O_.orc_GV(cr, L, T), rel=1e-13)
            assert L_.opo_gvd(cr, L, T) == pytest.approx(O_.orc_GVD(cr, L, T), rel=1e-12)
            fd1 = (L_.opo_n(cr, L + h, T) - L_.opo_n(cr, L - h, T)) / (2 * h)
            fd2 = (L_.opo_n(cr, L + h, T) - 2 * n + L_.opo_n(cr, L - h, T)) / h ** 2
            assert L_.opo_group_velocity(cr, L, T) == pytest.approx(C / (n - L * fd1), rel=1e-7)
            assert L_.opo_gvd(cr, L, T) == pytest.approx(L ** 3 * fd2 / (2 * math.pi * C ** 2), rel=1e-4)
    for lp, ls, lam, mode in ((0.532, 1.064, 6.97, 0), (1.55, 3.1, 35.01, 0), (0.532, 1.0, 7.97, 1),
                              (0.532, 1.064, 7.97, 0)):
        assert P.find_pm_temp(cr, lp, ls, lam, pm_mode=mode) == O_.orc_find_pm_temp(cr, lp, ls, lam, 10, 150, mode)
