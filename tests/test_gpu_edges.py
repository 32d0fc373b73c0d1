"""Edge shapes of the round trip against the FP64 oracle (GPU, C ABI): one or
two slices per pass (no slice-interior kernel / one), the smallest grid
(N = 256, 16 x 16 four-step split), the largest (N = 2^18, the NS workload
with EOM and GDD at full size), FP32C at the largest grid, and a linear step
whose only effect is absorption (a size-independent closed form)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import c1, c2, c5, inputs, rel_l2, to_args

pytestmark = pytest.mark.gpu


def solver(P, cfg, precision=0):
    return P.Solver(P.params_from(P.derive(to_args(cfg)), precision))


def grown(As, Ai, f=1e19):
    return As * f, (Ai * f if Ai is not None else None)


@pytest.mark.parametrize("NZ", [1, 2])
@pytest.mark.parametrize("three", [False, True])
def test_one_and_two_slices(gpu, NZ, three):
    cfg = c2(N=2048, NZ=NZ, NRT=10) if three else c1(N=2048, NZ=NZ, NRT=10)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=9)
    As, Ai = grown(As, Ai)
    s = solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(0, 5)  # graph replay path (more than two round trips)
    gp, gs, gi = s.download()
    O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 5)
    assert rel_l2(gp, Ap) < 1e-9 and rel_l2(gs, As) < 1e-9
    if three:
        assert rel_l2(gi, Ai) < 1e-9


@pytest.mark.parametrize("three", [False, True])
def test_smallest_grid(gpu, three):
    cfg = c2(N=256, NZ=8, NRT=10) if three else c1(N=256, NZ=8, NRT=10)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=4)
    As, Ai = grown(As, Ai)
    s = solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(0, 3)
    gp, gs, gi = s.download()
    O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 3)
    assert rel_l2(gp, Ap) < 1e-9 and rel_l2(gs, As) < 1e-9
    if three:
        assert rel_l2(gi, Ai) < 1e-9


def test_largest_grid_ns_eom_gdd(gpu):
    """The N = 2^18 nanosecond workload (EOM and GDD on), 4 slices, 2 round
    trips: second round trip reads pump slot 1."""
    cfg = c5(N=1 << 18, NZ=4, NRT=3)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=6)
    assert p.use_eom and p.GDD != 0
    As, Ai = grown(As, Ai)
    s = solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(0, 2)
    gp, gs, gi = s.download()
    O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 2)
    for g, o in ((gp, Ap), (gs, As), (gi, Ai)):
        assert rel_l2(g, o) < 1e-9


def test_largest_grid_fp32c(gpu):
    cfg = c5(N=1 << 18, NZ=4, NRT=3)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=6)
    As, Ai = grown(As, Ai)
    s = solver(gpu, cfg, precision=gpu.FP32C)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(0, 1)
    gp, gs, gi = s.download()
    O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 1)
    # FP32-complex fields with FP64 phases and reductions: single-precision level
    for g, o in ((gp, Ap), (gs, As), (gi, Ai)):
        assert rel_l2(g, o) < 1e-5


@pytest.mark.parametrize("N", [256, 1 << 18])
def test_linear_step_absorption_only(gpu, N):
    """With every dispersion term zero the linear step multiplies each field by
    e^{-alpha dz/2} (cwes3.cuh:112-125): exact for any grid size."""
    cfg = c2(N=N, NZ=10, NRT=3)
    P = gpu
    d = P.derive(to_args(cfg))
    prm = P.params_from(d, 0)
    prm.vp = prm.vs = prm.vi = d.vs
    prm.b2p = prm.b2s = prm.b2i = 0.0
    T, w = P.grids(d)
    rng = np.random.default_rng(N)
    A = [rng.standard_normal(N) + 1j * rng.standard_normal(N) for _ in range(3)]
    s = P.Solver(prm)
    s.upload(w, T, P.pump_input(d), *A)
    s.linear_step()
    out = s.download()
    for f, (g, a) in enumerate(zip(out, A)):
        alpha = (prm.alpha_crp, prm.alpha_crs, prm.alpha_cri)[f]
        assert rel_l2(g, a * np.exp(-0.5 * alpha * prm.dz)) < 1e-13


def test_run_rejects_bad_round_trip_ranges(gpu):
    """opo_run rejects a negative start or count and a range past INT_MAX before
    any launch; the fields stay as uploaded."""
    P = gpu
    cfg = c1(N=1024, NZ=4, NRT=3)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=2)
    s = solver(P, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    before = s.download()
    for rt0, n in [(-1, 1), (0, -1), (2**31 - 10, 11)]:
        with pytest.raises(P.OpoError):
            s.run(rt0, n)
    for a, b in zip(before, s.download()):
        if a is not None:
            assert np.array_equal(a, b)


def test_handles_close_on_leaving_a_with_block(gpu):
    """Solver and Batch release their device memory when a with-block ends;
    a closed handle reaches the C ABI as a null handle and is refused."""
    P = gpu
    cfg = c1(N=1024, NZ=4, NRT=3)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=4)
    with solver(P, cfg) as s:
        s.upload(w, T, Ap_in, Ap, As, Ai)
        s.run(0, 2)
    assert s._h is None
    with pytest.raises(P.OpoError):
        s.run(0, 1)
    prm = P.params_from(P.derive(to_args(cfg)), 0)
    with P.Batch([prm, prm]) as b:
        for k in range(2):
            b.upload(k, w, T, Ap_in, Ap, As, Ai)
        b.run(2, check_every=0, tol=0.0)
    assert b._h is None


def test_full_pulsed_workload_grid_against_oracle(gpu):
    """The N = 2^18 nanosecond workload at its full 200 slices per pass, EOM
    and GDD on, all NRT = 3 round trips: every pump slot is read, the last
    one shifted by a sample (src/headers/functions.cuh:191-213), and every
    round trip is saved (src/cuOPO.cu:570-608)."""
    cfg = c5(N=1 << 18, NZ=200, NRT=3)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=8)
    As, Ai = grown(As, Ai)
    s = solver(gpu, cfg)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.enable_save(True)
    s.run(0, 3)
    gp, gs, gi = s.download()
    sv = s.download_saved()
    s.close()
    osv, _ = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 3, save=True)
    for g, o in ((gp, Ap), (gs, As), (gi, Ai)):
        assert rel_l2(g, o) < 1e-9
    for k in ("p", "s", "i"):
        assert rel_l2(sv[k], osv[k]) < 1e-9, k
