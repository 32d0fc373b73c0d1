"""Pins for the oracle's propagation restatement: closed forms of the
nonlinear half step (three equations), the linear step on single tones, the
boundary operators, pump slot arithmetic, physics invariants, and an
independent NumPy restatement of whole round trips.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import np_restatement as R
from tests.helpers import c1, c2, c5, inputs, random_fields, rel_l2


def test_nl_half_three_eq_closed_form(oracle_lib):
    """cwes3.cuh:153-172 evaluates dAdz on the unchanged fields in every stage,
    so one half step is A + (h/6) i kappa P(A) (E(z) + 4E(z+dz/4) + E(z+dz/2))."""
    cfg = c2(N=256, NZ=10, NRT=2)
    p = O.setup(cfg)
    Ap, As, Ai = random_fields(256, True, seed=1, amp_p=50.0, amp_s=20.0)
    a0 = [x.copy() for x in (Ap, As, Ai)]
    z = 3 * p.dz
    O.nl_half(p, Ap, As, Ai, z)
    h = p.dz / 2
    E = lambda s: np.exp(1j * s * p.dk * z) + 4 * np.exp(1j * s * p.dk * (z + p.dz / 4)) + \
        np.exp(1j * s * p.dk * (z + p.dz / 2))
    ep = a0[0] + h / 6 * 1j * p.kp * a0[1] * a0[2] * E(-1)
    es = a0[1] + h / 6 * 1j * p.ks * a0[0] * np.conj(a0[2]) * E(+1)
    ei = a0[2] + h / 6 * 1j * p.ki * a0[0] * np.conj(a0[1]) * E(+1)
    assert rel_l2(Ap - a0[0], ep - a0[0]) < 1e-13
    assert rel_l2(As - a0[1], es - a0[1]) < 1e-13
    assert rel_l2(Ai - a0[2], ei - a0[2]) < 1e-13


def test_nl_half_two_eq_stage_chain(oracle_lib):
    """cwes2.cuh:147-166 on a tiny grid against the NumPy stage chain
    (stages on A + 0.5 k1, A + 0.5 k2, A + k3; combine with h/6)."""
    p = O.setup(c1(N=256, NZ=10, NRT=2))
    Ap, As, _ = random_fields(256, False, seed=2, amp_p=3e3, amp_s=1e3)  # stage increments matter
    ref = R.nl_half(p, [Ap.copy(), As.copy(), None], 0.7 * p.dz)
    O.nl_half(p, Ap, As, None, 0.7 * p.dz)
    assert rel_l2(Ap, ref[0]) < 1e-13 and rel_l2(As, ref[1]) < 1e-13


def test_nl_half_vanishes_without_signal(oracle_lib):
    for cfg in (c1(N=256, NZ=4, NRT=2), c2(N=256, NZ=4, NRT=2)):
        p = O.setup(cfg)
        Ap = np.full(256, 0.3 + 0.1j)
        As = np.zeros(256, complex)
        Ai = np.zeros(256, complex) if cfg.three_eqs else None
        O.nl_half(p, Ap, As, Ai, 0.0)
        assert np.all(Ap == 0.3 + 0.1j) and np.all(As == 0)


@pytest.mark.parametrize("three", [False, True])
def test_linear_step_single_tone(oracle_lib, three):
    """A tone e^{-2 pi i m j/N} is an eigenvector of IFFT+/N -> phase -> FFT-:
    it comes back multiplied by exp(i dz w_m (D + w_m b2/2)) exp(-alpha dz/2)
    with the field's own walk-off D (pump 1/vp-1/vi in three equations,
    1/vp-1/vs in two; signal 0; idler 1/vi-1/vs)."""
    cfg = c2(N=512, NZ=20, NRT=2) if three else c1(N=512, NZ=20, NRT=2)
    p = O.setup(cfg)
    T, w = O.grids(p)
    j = np.arange(512)
    tones = {0: 5, 1: 130, 2: 300}
    fields = [np.exp(-2j * np.pi * tones[f] * j / 512) for f in range(3 if three else 2)]
    ins = [x.copy() for x in fields]
    O.linear(p, w, fields[0], fields[1], fields[2] if three else None)
    if three:
        D = [1 / p.vp - 1 / p.vi, 0.0, 1 / p.vi - 1 / p.vs]
        b2 = [p.b2p, p.b2s, p.b2i]
        al = [p.alpha_crp, p.alpha_crs, p.alpha_cri]
    else:
        D = [1 / p.vp - 1 / p.vs, 0.0]
        b2 = [p.b2p, p.b2s]
        al = [p.alpha_crp, p.alpha_crs]
    for f in range(len(fields)):
        wm = w[tones[f]]
        L = np.exp(1j * p.dz * wm * (D[f] + 0.5 * wm * b2[f])) * np.exp(-0.5 * al[f] * p.dz)
        assert rel_l2(fields[f], ins[f] * L) < 1e-13


def test_boundary_closed_forms(oracle_lib):
    """EOM multiplies by exp(i m sin(2 pi fpm T_j)); the mirror by sqrt(R)
    e^{i pi delta}; GDD is the linear step's eigen-map with exp(i w^2 GDD/2)."""
    cfg = c5(N=512, NZ=4, NRT=3)
    cfg.regime = O.CW
    cfg.delta = 0.37
    p = O.setup(cfg)
    T, w = O.grids(p)
    rng = np.random.default_rng(5)
    A = rng.standard_normal(512) + 1j * rng.standard_normal(512)
    B = A.copy()
    O.eom(p, T, B)
    assert rel_l2(B, A * np.exp(1j * p.mod_depth * np.sin(2 * np.pi * p.fpm * T))) < 1e-15
    B = A.copy()
    O.mirror(p, p.Rs, B)
    assert rel_l2(B, A * np.sqrt(p.Rs) * np.exp(1j * np.pi * 0.37)) < 1e-15
    m = 77
    tone = np.exp(-2j * np.pi * m * np.arange(512) / 512)
    B = tone.copy()
    O.gdd(p, w, B)
    assert rel_l2(B, tone * np.exp(0.5j * w[m] ** 2 * p.GDD)) < 1e-13
    assert p.GDD == pytest.approx(-1.0 * p.b2s * p.Lcr)


def test_pump_slots(oracle_lib):
    """ReadPump (functions.cuh:191-213): slot nn for 0 <= nn < NRT-1; the last
    round trip reads N samples starting one sample before slot NRT-1."""
    cfg = c5(N=256, NZ=4, NRT=5)
    p = O.setup(cfg)
    Ap_in = np.arange(256 * 5).astype(complex)
    for nn in range(4):
        assert np.array_equal(O.read_pump(p, Ap_in, nn), Ap_in[nn * 256:(nn + 1) * 256])
    last = O.read_pump(p, Ap_in, 4)
    assert np.array_equal(last, Ap_in[4 * 256 - 1: 5 * 256 - 1])
    cfg.regime = O.CW
    p = O.setup(cfg)
    assert np.array_equal(O.read_pump(p, Ap_in, 3), Ap_in[:256])


def test_pump_alone_propagates_linearly(oracle_lib):
    """With no signal or idler the crystal pass is linear: a CW pump sits in the
    k = 0 bin, whose w[0] = pi N dF/(N-1) != 0 (Q7), and picks up NZ times
    that bin's phase and attenuation."""
    cfg = c2(N=256, NZ=12, NRT=2)
    p = O.setup(cfg)
    T, w = O.grids(p)
    Ap = np.full(256, p.Ap0 + 0j)
    As = np.zeros(256, complex)
    Ai = np.zeros(256, complex)
    O.crystal_pass(p, w, Ap, As, Ai)
    w0 = w[0]
    L = np.exp(1j * p.dz * w0 * ((1 / p.vp - 1 / p.vi) + 0.5 * w0 * p.b2p)) * np.exp(-0.5 * p.alpha_crp * p.dz)
    assert rel_l2(Ap, np.full(256, p.Ap0) * L ** cfg.NZ) < 1e-12
    assert np.all(As == 0) and np.all(Ai == 0)


def test_below_and_above_threshold(oracle_lib):
    """Noise decays below threshold and grows above it."""
    for nth, grows in ((0.5, False), (30.0, True)):
        cfg = c1(N=256, NZ=4, NRT=300)
        cfg.Nth = nth
        p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=3, sigma=1e-10)
        _, pw = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 300, power=True)
        assert (pw[-1] > 10 * pw[0]) == grows


@pytest.mark.parametrize("which", ["c1", "c2", "c5"])
def test_round_trips_match_numpy_restatement(oracle_lib, which):
    """Three full round trips of the C oracle against the independent NumPy
    restatement (both switch sets, CW and NS with EOM and GDD)."""
    make = {"c1": c1, "c2": c2, "c5": c5}[which]
    cfg = make(N=256, NZ=4, NRT=4)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=8, sigma=1e-2)
    ref, rpw = R.run(p, w, T, Ap_in, As, Ai, 3)
    _, opw = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 3, power=True)
    assert rel_l2(As, ref[1]) < 1e-12
    assert rel_l2(Ap, ref[0]) < 1e-12
    if Ai is not None:
        assert rel_l2(Ai, ref[2]) < 1e-12
    assert np.allclose(opw, rpw, rtol=1e-12)
