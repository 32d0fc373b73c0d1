"""GPU parity: each HIP stage and the full round trip against the FP64 oracle
on identical seeded inputs, through the C ABI (cuopo_amd -> libcuopo_amd.so).

Tolerances (relative L2 per field, from the solver's stated parity bar):
  single stages (linear step, NL half step, boundary) ... 1e-12
  one crystal pass / one round trip ...................... 1e-9
  1000 round trips ........................................ 1e-7
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import c1, c2, c5, inputs, random_fields, rel_l2, to_args

pytestmark = pytest.mark.gpu


def make_solver(P, cfg, precision=0):
    d = P.derive(to_args(cfg))
    p = P.params_from(d, precision)
    return P.Solver(p)


def oracle_params(cfg):
    return O.setup(cfg)


@pytest.mark.parametrize("three", [False, True])
@pytest.mark.parametrize("N", [256, 1024, 4096, 16384, 32768, 262144])
def test_linear_step(gpu, three, N):
    cfg = c2(N=N, NZ=10, NRT=3) if three else c1(N=N, NZ=10, NRT=3)
    p, T, w, Ap_in, _, _, _ = inputs(cfg)
    Ap, As, Ai = random_fields(N, three, seed=N)
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.linear_step()
    gp, gs, gi = s.download()
    O.linear(p, w, Ap, As, Ai)
    print(f"N={N} three={three} err p={rel_l2(gp, Ap):.2e} s={rel_l2(gs, As):.2e}")
    assert rel_l2(gp, Ap) < 1e-12
    assert rel_l2(gs, As) < 1e-12
    if three:
        assert rel_l2(gi, Ai) < 1e-12


@pytest.mark.parametrize("three", [False, True])
def test_linear_step_generic_grid(gpu, three):
    """A frequency grid that is not the reference's affine two-half grid takes
    the full phase-table path of the row stage (the default grid takes the
    factorised one); both must match the oracle."""
    N = 4096
    cfg = c2(N=N, NZ=10, NRT=3) if three else c1(N=N, NZ=10, NRT=3)
    p, T, w, Ap_in, _, _, _ = inputs(cfg)
    w = w * (1 + 1e-3 * np.sin(np.arange(N)))
    Ap, As, Ai = random_fields(N, three, seed=5)
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.linear_step()
    gp, gs, gi = s.download()
    O.linear(p, w, Ap, As, Ai)
    assert rel_l2(gp, Ap) < 1e-12
    assert rel_l2(gs, As) < 1e-12
    if three:
        assert rel_l2(gi, Ai) < 1e-12


@pytest.mark.parametrize("three", [False, True])
def test_nl_half(gpu, three):
    N = 4096
    cfg = c2(N=N, NZ=10, NRT=3) if three else c1(N=N, NZ=10, NRT=3)
    p, T, w, Ap_in, _, _, _ = inputs(cfg)
    Ap, As, Ai = random_fields(N, three, seed=7, amp_p=300.0, amp_s=100.0)  # strong coupling
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    sl = 7
    s.nl_half(sl, 2)
    gp, gs, gi = s.download()
    z = 0.0
    for _ in range(sl):
        z += p.dz
    A0 = Ap.copy()
    O.nl_half(p, Ap, As, Ai, z)
    O.nl_half(p, Ap, As, Ai, z)
    assert rel_l2(Ap, A0) > 1e-6  # the step did something
    assert rel_l2(gp, Ap) < 1e-12
    assert rel_l2(gs, As) < 1e-12
    if three:
        assert rel_l2(gi, Ai) < 1e-12


@pytest.mark.parametrize("three", [False, True])
def test_boundary_eom_gdd_mirror(gpu, three):
    N = 4096
    base = c5(N=N, NZ=4, NRT=3)
    base.three_eqs = three
    base.regime = O.CW
    p, T, w, Ap_in, _, _, _ = inputs(base)
    assert p.use_eom and p.GDD != 0
    Ap, As, Ai = random_fields(N, three, seed=3)
    s = make_solver(gpu, base)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.boundary()
    gp, gs, gi = s.download()
    O.eom(p, T, As)
    O.gdd(p, w, As)
    O.mirror(p, p.Rs, As)
    assert rel_l2(gp, Ap) == 0.0  # the pump is untouched by the boundary
    assert rel_l2(gs, As) < 1e-12
    if three:
        # idler: same EOM, GDD with beta2_s (src/cuOPO.cu:546), mirror with Ri
        O.eom(p, T, Ai)
        O.gdd(p, w, Ai)
        O.mirror(p, p.Ri, Ai)
        assert rel_l2(gi, Ai) < 1e-12


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_crystal_pass(gpu, which):
    cfg = c1(N=4096, NZ=32, NRT=3) if which == "c1" else c2(N=16384, NZ=20, NRT=3)
    p, T, w, Ap_in, Ap, _, _ = inputs(cfg)
    three = cfg.three_eqs
    _, As, Ai = random_fields(cfg.N, three, seed=11, amp_s=1e-2)
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.crystal_pass()
    gp, gs, gi = s.download()
    O.crystal_pass(p, w, Ap, As, Ai)
    print(f"{which}: p={rel_l2(gp, Ap):.2e} s={rel_l2(gs, As):.2e}")
    assert rel_l2(gp, Ap) < 1e-11
    assert rel_l2(gs, As) < 1e-11
    if three:
        assert rel_l2(gi, Ai) < 1e-11


@pytest.mark.parametrize("which", ["c1", "c2", "c5", "c5_2eq"])
def test_one_round_trip(gpu, which):
    if which == "c1":
        cfg = c1(N=4096, NZ=32, NRT=2000)
    elif which == "c2":
        cfg = c2(N=16384, NZ=100, NRT=5000)
    else:
        cfg = c5(N=4096, NZ=16, NRT=10)
        cfg.three_eqs = which == "c5"  # NS with two equations: the switches are independent
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=2)
    # start from a grown signal so the nonlinear terms matter
    As = As * 1e19
    if Ai is not None:
        Ai = Ai * 1e19
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(0, 1)
    gp, gs, gi = s.download()
    O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 1)
    print(f"{which}: p={rel_l2(gp, Ap):.2e} s={rel_l2(gs, As):.2e}")
    assert rel_l2(gp, Ap) < 1e-9
    assert rel_l2(gs, As) < 1e-9
    if Ai is not None:
        assert rel_l2(gi, Ai) < 1e-9


@pytest.mark.parametrize("three", [False, True])
def test_1000_round_trips(gpu, three):
    """1000 round trips on a grid the oracle finishes in seconds, covering
    build-up and saturation: two equations from 1e-20 noise; three equations
    (slower build-up) at 10x threshold from 1e-6 noise."""
    if three:
        cfg = c2(N=1024, NZ=8, NRT=1000)
        cfg.Nth = 10.0
        sigma = 1e-6
    else:
        cfg = c1(N=1024, NZ=8, NRT=1000)
        sigma = 1e-20
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=5, sigma=sigma)
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(0, 1000)
    gp, gs, gi = s.download()
    gpow = s.power_history(1000)
    _, opow = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 1000, power=True)
    print(f"three={three}: P0={opow[0]:.3e} P_end={opow[-1]:.3e} err s={rel_l2(gs, As):.2e} p={rel_l2(gp, Ap):.2e}")
    assert opow[-1] > 1e6 * opow[0]  # the oscillator built up
    assert rel_l2(gp, Ap) < 1e-7
    assert rel_l2(gs, As) < 1e-7
    if three:
        assert rel_l2(gi, Ai) < 1e-7
    assert np.max(np.abs(gpow - opow) / opow) < 1e-7


def test_ns_saving_and_last_slot(gpu):
    """NS regime: every round trip saved, pump read per slot, the final
    round trip reads one sample early (functions.cuh:206-209)."""
    cfg = c5(N=1024, NZ=4, NRT=6)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=9)
    As = As * 1e18
    Ai = Ai * 1e18
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.enable_save(True)
    s.run(0, cfg.NRT)
    sv = s.download_saved()
    osv, _ = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, cfg.NRT, save=True)
    for k in ("p", "s", "i"):
        assert rel_l2(sv[k], osv[k]) < 1e-9, k


def test_cw_saves_last_64(gpu):
    cfg = c1(N=256, NZ=4, NRT=70)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=4)
    s = make_solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.enable_save(True)
    s.run(0, cfg.NRT)
    sv = s.download_saved()
    osv, _ = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, cfg.NRT, save=True)
    assert sv["s"].shape == (64, 256)
    for k in ("p", "s"):
        assert rel_l2(sv[k], osv[k]) < 1e-9, k


def test_split_runs_equal_one_run(gpu):
    """opo_run(0, 5) then opo_run(5, 5) equals opo_run(0, 10)."""
    cfg = c5(N=1024, NZ=4, NRT=12)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=13)
    As, Ai = As * 1e18, Ai * 1e18
    a = make_solver(gpu, cfg); a.upload(w, T, Ap_in, Ap, As, Ai); a.run(0, 10)
    b = make_solver(gpu, cfg); b.upload(w, T, Ap_in, Ap, As, Ai); b.run(0, 5); b.run(5, 5)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)


def test_graph_equals_eager(gpu):
    cfg = c2(N=4096, NZ=6, NRT=20)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=21)
    As, Ai = As * 1e18, Ai * 1e18
    a = make_solver(gpu, cfg); a.upload(w, T, Ap_in, Ap, As, Ai); a.use_graph(True); a.run(0, 12)
    b = make_solver(gpu, cfg); b.upload(w, T, Ap_in, Ap, As, Ai); b.use_graph(False); b.run(0, 12)
    for x, y in zip(a.download(), b.download()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("col,row", [(64, 64), (128, 128), (256, 256), (64, 256)])
def test_tiles_do_not_change_results(gpu, col, row):
    cfg = c2(N=16384, NZ=4, NRT=3)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=1)
    As, Ai = As * 1e18, Ai * 1e18
    ref = make_solver(gpu, cfg); ref.upload(w, T, Ap_in, Ap, As, Ai); ref.run(0, 2)
    s = make_solver(gpu, cfg); s.set_tiles(col, row); s.upload(w, T, Ap_in, Ap, As, Ai); s.run(0, 2)
    for x, y in zip(s.download(), ref.download()):
        assert rel_l2(x, y) < 1e-14
