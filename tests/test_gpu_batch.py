"""Batched sweeps (instances of several grid sizes in one launch per stage),
the convergence test with snapshots, the FP32-complex variant's error against
the FP64 oracle, and failure handling.  GPU tests through the C ABI."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import c1, c2, c5, inputs, rel_l2, to_args

pytestmark = pytest.mark.gpu


def instance(P, cfg, seed, sigma, precision=0):
    d = P.derive(to_args(cfg))
    p = P.params_from(d, precision)
    T, w = P.grids(d)
    Ap_in = P.pump_input(d)
    As = O.noise(seed, 0, sigma, d.N)
    Ai = O.noise(seed, 1, sigma, d.N) if cfg.three_eqs else None
    return p, dict(w=w, T=T, Ap_in=Ap_in, Ap=Ap_in[: d.N].copy(), As=As, Ai=Ai)


def sweep_configs(n=6):
    rng = np.random.default_rng(0)
    out = []
    for k in range(n):
        cfg = c2(N=[1024, 4096, 2048][k % 3], NZ=6, NRT=200)
        cfg.Nth = float(rng.uniform(1.2, 5.0)) * 3
        cfg.delta = float(rng.uniform(-np.pi, np.pi)) / np.pi
        cfg.R_pct = float(rng.uniform(90, 99))
        out.append(cfg)
    return out


def test_batch_equals_single_instances(gpu):
    P = gpu
    cfgs = sweep_configs(6)
    made = [instance(P, c, 100 + k, 1e-4) for k, c in enumerate(cfgs)]
    b = P.Batch([m[0] for m in made])
    for k, (p, x) in enumerate(made):
        b.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
    rts = b.run(40, check_every=0, tol=0.0)
    assert np.all(rts == -1)
    for k, (p, x) in enumerate(made):
        s = P.Solver(p)
        s.upload(x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
        s.run(0, 40)
        for a, c in zip(b.download(k), s.download()):
            assert rel_l2(a, c) < 1e-14, k


def test_batch_matches_oracle(gpu):
    P = gpu
    cfgs = sweep_configs(3)
    made = [instance(P, c, 7 + k, 1e-3) for k, c in enumerate(cfgs)]
    b = P.Batch([m[0] for m in made])
    for k, (p, x) in enumerate(made):
        b.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
    b.run(5, check_every=0, tol=0.0)
    for k, cfg in enumerate(cfgs):
        po = O.setup(cfg)
        T, w = O.grids(po)
        x = made[k][1]
        Ap, As, Ai = x["Ap"].copy(), x["As"].copy(), x["Ai"].copy()
        O.run(po, w, T, x["Ap_in"], Ap, As, Ai, 0, 5)
        gp, gs, gi = b.download(k)
        assert rel_l2(gs, As) < 1e-9 and rel_l2(gi, Ai) < 1e-9 and rel_l2(gp, Ap) < 1e-9


def test_convergence_snapshot_is_the_converged_state(gpu):
    """An instance declared converged at round trip r (relative power change
    over check_every round trips below tol) is frozen; its downloaded fields
    equal a single-instance run of exactly r round trips.  A below-threshold
    instance in the same batch stops as "off" once its decay rate is settled."""
    P = gpu
    cfg = c2(N=1024, NZ=8, NRT=3000)
    cfg.Nth = 10.0
    p, x = instance(P, cfg, 5, 1e-6)
    cfg2 = c2(N=1024, NZ=8, NRT=3000)
    cfg2.Nth = 0.5  # below threshold: decays, never converges
    p2, x2 = instance(P, cfg2, 6, 1e-6)
    b = P.Batch([p, p2])
    b.upload(0, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
    b.upload(1, x2["w"], x2["T"], x2["Ap_in"], x2["Ap"], x2["As"], x2["Ai"])
    rts, st = b.run_ex(3000, check_every=50, tol=1e-3, off_ratio=1e-6)
    assert rts[0] > 0 and rts[0] % 50 == 0 and st[0] == 1
    assert st[1] == 2 and 0 < rts[1] < 3000  # the below-threshold instance stops as "off"
    pw1 = b.power_history(1, int(rts[1]))
    assert pw1[-1] < 1e-6 * pw1[49] and pw1[-1] < pw1[-51]
    s = P.Solver(p)
    s.upload(x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
    s.run(0, int(rts[0]))
    pw = s.power_history(int(rts[0]))
    assert abs(pw[-1] - pw[-51]) <= 1e-3 * pw[-1]
    if len(pw) > 100:
        assert abs(pw[-51] - pw[-101]) > 1e-3 * pw[-51]  # not converged at the previous check
    for a, c in zip(b.download(0), s.download()):
        if a is not None:
            assert rel_l2(a, c) < 1e-14


def test_reference_instability_is_reproduced_and_detected(gpu):
    """Two equations on a coarse crystal grid (8 slices, 31x threshold) make the
    reference scheme itself blow up; the oracle and the HIP solver go
    non-finite at the same round trip, and the batched run reports the
    instance as diverged (opo_batch_run -> OPO_ERR_DIVERGED)."""
    P = gpu
    cfg = c1(N=1024, NZ=8, NRT=1100)
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=5, sigma=1e-20)
    _, opw = O.run(p, w, T, Ap_in, Ap.copy(), As.copy(), None, 0, 1100, power=True)
    o_bad = int(np.argmax(~np.isfinite(opw)))
    assert not np.isfinite(opw[o_bad])
    s = P.Solver(P.params_from(P.derive(to_args(cfg))))
    s.upload(w, T, Ap_in, Ap, As, None)
    s.run(0, 1100)
    gpw = s.power_history(1100)
    g_bad = int(np.argmax(~np.isfinite(gpw)))
    assert g_bad == o_bad
    assert np.allclose(gpw[:o_bad - 30], opw[:o_bad - 30], rtol=1e-6)
    pb, xb = instance(P, cfg, 5, 1e-20)
    b = P.Batch([pb])
    b.upload(0, xb["w"], xb["T"], xb["Ap_in"], xb["Ap"], xb["As"], xb["Ai"])
    with pytest.raises(P.OpoError) as e:
        b.run(1100, check_every=50, tol=1e-9)
    assert e.value.status == P.ERR_DIVERGED


@pytest.mark.parametrize("which", ["c1", "c2", "c5"])
def test_fp32c_error_against_fp64_oracle(gpu, which):
    """FP32-complex fields/FFT/RK with FP64 QPM phase, detuning phase and power
    reduction: the relative L2 error against the FP64 oracle is recorded per
    configuration (to $OPO_TEST_RECORD_DIR/fp32c_errors.json when that is set)
    and bounded."""
    P = gpu
    make = {"c1": lambda: c1(N=4096, NZ=32, NRT=200), "c2": lambda: c2(N=16384, NZ=20, NRT=200),
            "c5": lambda: c5(N=4096, NZ=16, NRT=40)}[which]
    cfg = make()
    nrt = 20
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=4, sigma=1e-2)
    s32 = P.Solver(P.params_from(P.derive(to_args(cfg)), P.FP32C))
    s32.upload(w, T, Ap_in, Ap, As, Ai)
    s32.run(0, 1)
    g1 = s32.download()
    A1 = [Ap.copy(), As.copy(), None if Ai is None else Ai.copy()]
    O.run(p, w, T, Ap_in, A1[0], A1[1], A1[2], 0, 1)
    s32.set_fields(Ap, As, Ai)
    s32.run(0, nrt)
    gN = s32.download()
    AN = [Ap.copy(), As.copy(), None if Ai is None else Ai.copy()]
    O.run(p, w, T, Ap_in, AN[0], AN[1], AN[2], 0, nrt)
    err1 = {f: rel_l2(g1[i], A1[i]) for i, f in enumerate("psi") if A1[i] is not None}
    errN = {f: rel_l2(gN[i], AN[i]) for i, f in enumerate("psi") if AN[i] is not None}
    print(which, "FP32C rel L2 after 1 RT", err1, f"after {nrt} RT", errN)
    rec_dir = os.environ.get("OPO_TEST_RECORD_DIR")
    if rec_dir:
        os.makedirs(rec_dir, exist_ok=True)
        path = os.path.join(rec_dir, "fp32c_errors.json")
        rec = json.load(open(path)) if os.path.exists(path) else {}
        rec[which] = {"after_1_rt": err1, f"after_{nrt}_rt": errN}
        json.dump(rec, open(path, "w"), indent=1)
    assert max(err1.values()) < 1e-5
    assert max(errN.values()) < 1e-3


def test_nonfinite_power_marks_instance_diverged(gpu):
    P = gpu
    cfg = c2(N=1024, NZ=4, NRT=100)
    p, x = instance(P, cfg, 1, 1e-3)
    bad = x["As"].copy()
    bad[17] = np.nan
    p2, x2 = instance(P, cfg, 2, 1e-3)
    b = P.Batch([p, p2])
    b.upload(0, x["w"], x["T"], x["Ap_in"], x["Ap"], bad, x["Ai"])
    b.upload(1, x2["w"], x2["T"], x2["Ap_in"], x2["Ap"], x2["As"], x2["Ai"])
    with pytest.raises(P.OpoError) as e:
        b.run(60, check_every=50, tol=1e-12)
    assert e.value.status == P.ERR_DIVERGED


def test_diverged_warm_start_falls_back_to_the_cold_start(gpu):
    """A warm start with a NaN sample makes its point diverge; run_warm runs
    that point again from its cold-start noise and reports it, and the other
    points keep their warm results.  The fallback's outcome equals the plain
    cold run of the point."""
    P = gpu
    from cuopo_amd import sweep as SW
    pts = [(1024, 3.0, 0.0, 20.0, 0.0, 0.95), (1024, 2.5, 0.3, 20.0, 0.0, 0.95), (2048, 3.0, -0.2, 20.0, 0.0, 0.95)]
    warm = [(P.noise(7 + k, 0, 1e-6, q[0]), P.noise(7 + k, 1, 1e-6, q[0])) for k, q in enumerate(pts)]
    bad = warm[1][0].copy()
    bad[5] = np.nan
    warm[1] = (bad, warm[1][1])
    kw = dict(NZ=4, max_rt=300, precision=P.FP64, autotune=False)
    plain = SW.run_batched(pts, warm=warm, **kw)
    assert plain["states"][1] == -1 and -1 not in (plain["states"][0], plain["states"][2])
    r = SW.run_warm(pts, warm, **kw)
    cold = SW.run_batched(pts, **kw)
    try:
        assert list(r["fallback"]) == [1]
        assert r["states"][1] == cold["states"][1] != -1
        assert r["power"][1] == pytest.approx(cold["power"][1], rel=1e-12)
        assert r["rts"][1] > cold["rts"][1] > 0 or cold["rts"][1] <= 0
        for k in (0, 2):
            assert r["states"][k] == plain["states"][k] and r["power"][k] == plain["power"][k]
    finally:
        for x in (plain, r, cold):
            x["batch"].close()
        if r["fallback_batch"] is not None:
            r["fallback_batch"].close()


def test_sweep_driver_passes_the_extrapolation_schedule_on(gpu):
    """run_batched(accel=(every, start)) runs exactly as a batch given the
    same schedule through Batch.set_accel (bit for bit)."""
    P = gpu
    from cuopo_amd import sweep as SW
    pts = [(1024, 3.0, 0.0, 20.0, 0.0, 0.95), (1024, 1.4, 0.3, 20.0, 0.0, 0.95), (2048, 3.0, -0.2, 20.0, 0.0, 0.95)]
    kw = dict(NZ=4, max_rt=1500, precision=P.FP64, sigma=1e-3)
    r = SW.run_batched(pts, autotune=False, accel=(100, 200), **kw)
    b, _ = SW.make_batch(pts, **kw)
    try:
        b.set_accel(100, 200)
        rts, st = b.run_ex(1500, check_every=50, tol=1e-6, off_ratio=1e-6)
        assert np.array_equal(r["rts"], rts) and np.array_equal(r["states"], st)
        assert np.array_equal(r["power"], b.power())
    finally:
        b.close()
        r["batch"].close()


def test_graph_cache_follows_the_convergence_settings(gpu):
    """The captured round-trip graph is reused only for the same convergence
    settings: a batch run with tol = 0.02, then re-run with tol = 1e-6, matches
    fresh batches run with each setting."""
    P = gpu
    cfgs = [c2(N=1024, NZ=4, NRT=400), c2(N=1024, NZ=4, NRT=400)]
    cfgs[1].Nth = 1.5
    made = [instance(P, c, 11 + k, 1e-3) for k, c in enumerate(cfgs)]

    def fresh():
        b = P.Batch([m[0] for m in made])
        for k, (_, x) in enumerate(made):
            b.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
        return b

    def result(b, tol):
        rts, st = b.run_ex(400, check_every=50, tol=tol)
        return rts.tolist(), st.tolist(), b.power().tolist()

    a = fresh()
    r1 = result(a, 0.02)
    for k, (_, x) in enumerate(made):
        a.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
    r2 = result(a, 1e-6)
    a.close()
    for tol, r in ((0.02, r1), (1e-6, r2)):
        b = fresh()
        assert result(b, tol) == r, tol
        b.close()
    assert r1 != r2


def test_graph_reuse_across_runs_and_off_ratio_changes(gpu, monkeypatch):
    """A batch re-run without new uploads replays the round-trip graph it
    captured; a different off_ratio (an argument of the captured tick kernel)
    makes it capture again.  Three runs in a row (off_ratio 1e-3, 1e-3, then
    1e-12) equal the same three runs of a batch that launches every kernel
    eagerly; the second run declares both instances off, the third (stricter
    collapse) runs them to the cap."""
    P = gpu
    cfgs = [c2(N=1024, NZ=4, NRT=1500), c2(N=1024, NZ=4, NRT=1500)]
    cfgs[0].Nth, cfgs[1].Nth = 0.5, 0.7  # below threshold: both decay and are declared off
    made = [instance(P, c, 21 + k, 1e-3) for k, c in enumerate(cfgs)]

    def fresh():
        b = P.Batch([m[0] for m in made])
        for k, (_, x) in enumerate(made):
            b.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
        return b

    def runs(b):
        out = []
        for off in (1e-3, 1e-3, 1e-12):
            rts, st = b.run_ex(1500, check_every=50, tol=1e-6, off_ratio=off)
            out.append((rts.tolist(), st.tolist(), b.power().tolist()))
        return out

    g = fresh()
    with_graph = runs(g)
    g.close()
    monkeypatch.setenv("OPO_USE_GRAPH", "0")  # read when the batch is created
    e = fresh()
    eager = runs(e)
    e.close()
    print(with_graph)
    assert with_graph == eager
    assert with_graph[1][1] == [2, 2]  # the replayed graph declares both off under 1e-3
    assert with_graph[2][1] == [0, 0]  # the new capture (1e-12) runs them to the cap


def test_invalid_calls_fail_cleanly(gpu):
    P = gpu
    cfg = c1(N=1024, NZ=4, NRT=5)
    s = P.Solver(P.params_from(P.derive(to_args(cfg))))
    with pytest.raises(P.OpoError) as e:
        s.run(0, 2)  # before upload
    assert e.value.status == P.ERR_STATE
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg)
    with pytest.raises(P.OpoError):
        s.upload(w, T, Ap_in[:10], Ap, As, Ai)  # pump too short
    with pytest.raises(P.OpoError):
        s.nl_half(99, 1)
    with pytest.raises(P.OpoError):
        s.set_tiles(48, 0)


def test_compaction_leaves_results_unchanged(gpu, monkeypatch):
    """A converging batch shrinks its launches to the still-active instances
    (OPO_COMPACT, on by default); instances that finished were skipped by
    every block before, so round trips, states, powers, power histories and
    fields equal those of a batch that keeps every instance in every launch,
    bit for bit.  A second run on the same batch starts from all instances.
    Rows of 256 threads make the compacted launches (at most 64 instances)
    regroup the rows into one-wavefront blocks; the tuned row tile is back
    after the run."""
    P = gpu
    cfgs = []
    for N in (1024, 2048):
        for nth in (0.5, 0.8, 2.0, 5.0, 10.0, 20.0):
            c = c2(N=N, NZ=4, NRT=2000)
            c.Nth = nth
            cfgs.append(c)
    made = [instance(P, c, 40 + k, 1e-6) for k, c in enumerate(cfgs)]

    def fresh():
        b = P.Batch([m[0] for m in made])
        for k, (_, x) in enumerate(made):
            b.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
        for N in (1024, 2048):
            b.set_tuning(N, row_threads=256)
        return b

    def result(b):
        rts, st = b.run_ex(2000, check_every=50, tol=1e-3, off_ratio=1e-6)
        hist = [b.power_history(k, 2000) for k in range(len(cfgs))]
        return rts.tolist(), st.tolist(), b.power().tolist(), hist, [b.download(k) for k in range(len(cfgs))]

    def same(r1, r2):
        assert r1[:3] == r2[:3]
        for h1, h2 in zip(r1[3], r2[3]):
            assert np.array_equal(h1, h2)
        for f1, f2 in zip(r1[4], r2[4]):
            for a, c in zip(f1, f2):
                assert (a is None and c is None) or np.array_equal(a, c)

    b = fresh()
    tuned = b.get_tuning()
    # rows of N2 = N / 32 points at these sizes: threads per row block = row_tile * N2 / row_spt
    assert all(t["row_tile"] * (t["N"] // 32) // t["row_spt"] > 64 for t in tuned)
    on = result(b)
    assert b.get_tuning() == tuned
    for k, (_, x) in enumerate(made):  # the same batch again, from the same inputs
        b.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
    again = result(b)
    b.close()
    monkeypatch.setenv("OPO_COMPACT", "0")  # read when the batch is created
    e = fresh()
    off = result(e)
    e.close()
    rts, st = on[0], on[1]
    assert 1 in st and 2 in st  # some steady, some off: instances finish at different round trips
    assert min(r for r in rts if r > 0) < max(rts)
    same(on, off)
    same(on, again)


def test_large_batch_runs_as_two_groups_with_unchanged_results(gpu, monkeypatch):
    """A batch of one grid size with 64 or more instances runs as two groups
    of half the instances on two streams (OPO_SPLIT=0 keeps one group); the
    round trips, states, powers and fields are those of the one-group batch,
    bit for bit, and a setting applied to the grid size reaches both groups."""
    P = gpu
    cfgs = []
    for k in range(64):
        c = c2(N=1024, NZ=4, NRT=1500)
        c.Nth = [0.6, 1.5, 3.0, 8.0][k % 4] * (1.0 + 0.01 * k)
        cfgs.append(c)
    made = [instance(P, c, 300 + k, 1e-6) for k, c in enumerate(cfgs)]

    def run():
        b = P.Batch([m[0] for m in made])
        for k, (_, x) in enumerate(made):
            b.upload(k, x["w"], x["T"], x["Ap_in"], x["Ap"], x["As"], x["Ai"])
        b.set_tuning(1024, col_threads=64, row_threads=128)
        tun = b.get_tuning()
        rts, st = b.run_ex(1500, check_every=50, tol=1e-3, off_ratio=1e-6)
        out = (rts.tolist(), st.tolist(), b.power().tolist(), [b.download(k) for k in range(len(cfgs))], tun)
        b.close()
        return out

    two = run()
    monkeypatch.setenv("OPO_SPLIT", "0")  # read when the batch is created
    one = run()
    assert [t["count"] for t in two[4]] == [32, 32] and [t["count"] for t in one[4]] == [64]
    assert all(t["row_tile"] == one[4][0]["row_tile"] and t["col_tile"] == one[4][0]["col_tile"] for t in two[4])
    assert 1 in two[1] and 2 in two[1]  # instances finish at different round trips
    assert two[:3] == one[:3]
    for f1, f2 in zip(two[3], one[3]):
        for a, c in zip(f1, f2):
            assert (a is None and c is None) or np.array_equal(a, c)

