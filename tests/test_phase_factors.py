"""The row stage's factorised linear-step phase (opo_linear_factors) against
the phase of the linear operator written out directly from
src/headers/cwes3.cuh:112-125 / cwes2.cuh:108-118:
  X_q *= exp(i dz w_q (D + w_q b2/2 + w_q^2 b3/6)) exp(-alpha dz/2) / N
(1/N of CUFFTscale, functions.cuh:126-135, folded in as the solver does).
CPU only: the factor tables are built on the host."""
import ctypes as C

import numpy as np
import pytest

import cuopo_amd as P
from cuopo_amd import sweep as SW
from cuopo_amd.surrogate import sweep_args
from tests.helpers import c1, c2, c5, to_args


def direct_phase(p, w):
    """Per field (pump, signal[, idler]) the N-point phase factor in natural frequency order."""
    z = p.dz
    if p.three_eqs:  # cwes3.cuh:117-119 (pump walk-off against the idler)
        coef = [(1 / p.vp - 1 / p.vi, p.b2p, p.b3p, p.alpha_crp), (0.0, p.b2s, p.b3s, p.alpha_crs),
                (1 / p.vi - 1 / p.vs, p.b2i, p.b3i, p.alpha_cri)]
    else:  # cwes2.cuh:112-113
        coef = [(1 / p.vp - 1 / p.vs, p.b2p, p.b3p, p.alpha_crp), (0.0, p.b2s, p.b3s, p.alpha_crs)]
    return [np.exp(1j * z * w * (D + 0.5 * w * b2 + w * w * b3 / 6)) * np.exp(-0.5 * al * z) / len(w)
            for D, b2, b3, al in coef]


def factors(p, w, N1, xl, E=8):
    N = p.N
    nf = 3 if p.three_eqs else 2
    FB = E * N1 + N // N1 + ((N // E) >> xl) + (1 << xl)
    out = np.zeros(2 * nf * FB)
    rc = P.lib().opo_linear_factors(C.byref(p), w.ctypes.data, N1, E, xl, out.ctypes.data)
    return rc, out.view(np.complex128).reshape(nf, FB)


def rebuild(tab, N, N1, xl, E=8):
    """Phase of natural frequency q = k1 + N1 k2 from the tables: the row
    stage's thread t holds k2 = t + m N2/E."""
    N2 = N // N1
    T2 = N2 // E
    q = np.arange(N)
    k1, k2 = q % N1, q // N1
    t, m = k2 % T2, k2 // T2
    e = k1 * t
    Z, S = tab[: E * N1], tab[E * N1: E * N1 + N2]
    XH = tab[E * N1 + N2: E * N1 + N2 + ((N // E) >> xl)]
    XL = tab[E * N1 + N2 + ((N // E) >> xl):]
    return Z[E * k1 + m] * S[k2] * XH[e >> xl] * XL[e & ((1 << xl) - 1)]


def product_setup(args):
    d = P.derive(args)
    p = P.params_from(d, P.FP64)
    T, w = P.grids(d)
    return p, w


CASES = {
    "c1_ppln_2eq": lambda: to_args(c1(N=4096)),
    "c2_spplt_3eq": lambda: to_args(c2(N=16384)),
    "c5_ns_3eq_2^18": lambda: to_args(c5(N=1 << 18)),
    "sweep_2^15": lambda: sweep_args(4.2, -2.5, 20.0, 0.0, 0.95, N=1 << 15, NZ=100, NRT=100),
    "sweep_gdd_2^12": lambda: sweep_args(1.7, 1.0, 35.0, 300.0, 0.92, N=1 << 12, NZ=100, NRT=100),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_factorised_phase_matches_direct(case):
    p, w = product_setup(CASES[case]())
    N = p.N
    L = int(np.log2(N))
    ref = direct_phase(p, w)
    for N1 in sorted({16, 64, 128, min(1 << (L // 2), 64)}):
        if N // N1 < 16:
            continue
        for E in (4, 8):
            for xl in {(L - 2) // 2, (L - 4) // 2}:
                rc, tab = factors(p, w, N1, xl, E)
                assert rc == 1, (case, N1, xl, E)
                for f, r in enumerate(ref):
                    got = rebuild(tab[f], N, N1, xl, E)
                    err = np.max(np.abs(got / r - 1))
                    assert err < 2e-13, (case, N1, xl, E, f, err)


def test_factorised_phase_is_not_trivial():
    """The cross term and both halves matter: dropping either breaks the match."""
    p, w = product_setup(to_args(c2(N=16384)))
    N, N1, xl = p.N, 64, 5
    rc, tab = factors(p, w, N1, xl)
    ref = direct_phase(p, w)[0]
    assert np.max(np.abs(rebuild(tab[0], N, N1, xl) / ref - 1)) < 1e-12
    no_cross = tab[0].copy()
    no_cross[8 * N1 + N // N1:] = 1.0
    assert np.max(np.abs(rebuild(no_cross, N, N1, xl) / ref - 1)) > 1e-3
    one_half = tab[0].copy().reshape(-1)
    z = one_half[: 8 * N1].reshape(N1, 8)
    z[:, 4:] = z[:, :4]
    assert np.max(np.abs(rebuild(one_half, N, N1, xl) / ref - 1)) > 1e-3


def test_generic_grid_and_third_order_fall_back():
    p, w = product_setup(to_args(c1(N=4096)))
    w2 = w * (1 + 1e-9 * np.sin(np.arange(p.N)))
    assert factors(p, w2, 64, 6)[0] == 0
    p.b3s = 1e-6
    assert factors(p, w, 64, 6)[0] == 0
    assert P.lib().opo_linear_factors(None, w.ctypes.data, 64, 8, 6, w.ctypes.data) == P.ERR_INVALID
    assert P.lib().opo_linear_factors(C.byref(p), w.ctypes.data, 64, 5, 6, w.ctypes.data) == P.ERR_INVALID
