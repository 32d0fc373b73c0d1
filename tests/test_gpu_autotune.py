"""Per-group launch settings and the autotuner (GPU, through the C ABI).

The autotuner times candidate tiles / XCD packing / phase sources on the live
instances; afterwards the fields must be exactly as uploaded and the chosen
settings must give the oracle's round trip.  Every setting must give the same
fields: the phase sources (full table, factorised tables) agree to rounding."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import c1, c2, inputs, rel_l2, to_args

pytestmark = pytest.mark.gpu


def solver_for(P, cfg, seed=3, sigma=1e-3):
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=seed, sigma=sigma)
    d = P.derive(to_args(cfg))
    s = P.Solver(P.params_from(d, P.FP64))
    s.upload(w, T, Ap_in, Ap, As, Ai)
    return s, (p, T, w, Ap_in, Ap, As, Ai)


@pytest.mark.parametrize("three", [False, True])
def test_autotune_keeps_fields_and_parity(gpu, three):
    P = gpu
    cfg = c2(N=4096, NZ=8, NRT=3) if three else c1(N=4096, NZ=8, NRT=3)
    s, (p, T, w, Ap_in, Ap, As, Ai) = solver_for(P, cfg)
    before = s.download()
    t = s.autotune(iters=3)
    print("tuning", t)
    assert t["slice_us"] > 0 and t["col_tile"] >= 1 and t["row_tile"] >= 1
    after = s.download()
    for a, b in zip(before, after):
        if a is not None:
            assert np.array_equal(a, b)
    s.run(0, 1)
    gp, gs, gi = s.download()
    O.run(p, w, T, Ap_in, Ap, As, Ai, 0, 1)
    assert rel_l2(gp, Ap) < 1e-9 and rel_l2(gs, As) < 1e-9
    if three:
        assert rel_l2(gi, Ai) < 1e-9


def test_every_setting_gives_the_same_round_trip(gpu):
    P = gpu
    cfg = c2(N=16384, NZ=6, NRT=3)
    ref = None
    for col, row, xcd, ph, ce, re in [(0, 0, 0, 0, 8, 8), (0, 0, 0, 1, 8, 8), (64, 256, 0, 1, 8, 8),
                                      (128, 512, 1, 0, 8, 8), (32, 64, 1, 1, 8, 8), (32, 128, 0, 0, 8, 8),
                                      (0, 0, 0, 0, 4, 4), (128, 128, 0, 1, 4, 4), (256, 512, 1, 1, 4, 8),
                                      (64, 64, 0, 1, 8, 4), (32, 32, 0, 0, 4, 8), (32, 32, 0, 1, 4, 4)]:
        s, _ = solver_for(P, cfg)
        s.set_tuning(col_threads=col, row_threads=row, xcd_pack=xcd, phase_mode=ph, col_spt=ce, row_spt=re)
        t = s.get_tuning()
        assert t["factored"] == ph and t["xcd_pack"] == xcd and t["col_spt"] == ce and t["row_spt"] == re
        s.run(0, 2)
        out = s.download()
        s.close()
        if ref is None:
            ref = out
            continue
        for a, b in zip(out, ref):
            assert rel_l2(a, b) < 1e-13, (col, row, xcd, ph, ce, re)


def test_batch_autotune_mixed_sizes(gpu):
    P = gpu
    made = []
    for k, N in enumerate([2048, 4096, 2048, 8192, 4096]):
        cfg = c2(N=N, NZ=6, NRT=50)
        cfg.Nth = 3.0 + k
        d = P.derive(to_args(cfg))
        p = P.params_from(d, P.FP64)
        T, w = P.grids(d)
        Ap_in = P.pump_input(d)
        As = O.noise(50 + k, 0, 1e-3, N)
        Ai = O.noise(50 + k, 1, 1e-3, N)
        made.append((p, (w, T, Ap_in, Ap_in[:N].copy(), As, Ai)))
    outs = []
    for tune in (False, True):
        b = P.Batch([m[0] for m in made])
        for k, m in enumerate(made):
            b.upload(k, *m[1])
        if tune:
            tab = b.autotune(iters=3)
            assert sorted(t["N"] for t in tab) == [2048, 4096, 8192]
            assert all(t["slice_us"] > 0 for t in tab)
        b.run(5, check_every=0, tol=0.0)
        outs.append([b.download(k) for k in range(len(made))])
        b.close()
    for a, b in zip(*outs):
        for x, y in zip(a, b):
            assert rel_l2(x, y) < 1e-13


def test_set_tuning_rejects_bad_values(gpu):
    P = gpu
    s, _ = solver_for(P, c1(N=1024, NZ=4, NRT=3))
    with pytest.raises(P.OpoError):
        s.set_tuning(col_threads=48)
    with pytest.raises(P.OpoError):
        s.set_tuning(phase_mode=2)
    with pytest.raises(P.OpoError):
        s.set_tuning(col_spt=2)
    # a rejected setting leaves the previous one in place
    s.set_tuning(row_threads=128, row_spt=4)
    with pytest.raises(P.OpoError):
        s.set_tuning(row_threads=128, row_spt=16)
    assert s.get_tuning()["row_spt"] == 4


def test_xcd_pack_entry_point_and_timers(gpu):
    """opo_set_xcd_pack changes where blocks run, not the result; the stage and
    run timers report positive times, the run's no shorter than its slices."""
    P = gpu
    cfg = c2(N=16384, NZ=6, NRT=5)
    ref, _ = solver_for(P, cfg)
    ref.run(0, 2)
    s, _ = solver_for(P, cfg)
    s.set_xcd_pack(True)
    assert s.get_tuning()["xcd_pack"] == 1
    s.run(0, 2)
    for a, b in zip(s.download(), ref.download()):
        assert rel_l2(a, b) < 1e-13
    col_us, row_us = s.time_stages(5)
    assert col_us > 0 and row_us > 0
    s.run(0, 4)
    ms = s.last_run_ms()
    assert ms > 0 and ms * 1e3 >= 4 * cfg.NZ * min(col_us, row_us) * 0.5
    s.close()
    ref.close()
