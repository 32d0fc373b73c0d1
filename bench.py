"""Benchmark of the SSFM round-trip solver on one MI355X.

Headline: sweep points per second to steady state.  A 256-point sweep of the
continuous-wave three-equation sPPLT family (pump 532 nm, signal 1000 nm,
idler 1137 nm; pump 1.2-5x threshold, detuning -pi..pi rad, crystal 5-40 mm,
GDD -500..500 fs^2, reflectivity 0.90-0.99) at the workload's grid (2^14
complex128 samples per round trip, 100 z-slices per pass) runs as one FP64
batch from the learned warm start (the MLP's predicted mode pair and power)
to the steady-state test: intracavity power change below 1e-6 over 50 round
trips, cap 20,000 round trips (a cold start from 1e-20 noise needs ~10^4),
with the batch's round-trip extrapolation on (every 150 round trips from
300 on, a point relaxing geometrically towards its steady state jumps to the
limit).  Counted: points that pass that test.  Beside it,
the plain per-instance FP64 solver cold-starts from 1e-20 noise on a subset
of the same points, one point after another, to the same test or cap, and
its steady points per second are measured, as is the warm start's steady
power against the cold start's where both reach steady state.

Prints one JSON line: the headline, the roofline of the batch's dominant
kernel (per-launch kernel time measured with start/stop events on the
batch's stream) and a CPU baseline (the FP64 oracle on this host's cores, a
bounded sample scaled to the same unit).  Under "extra": the single-instance
round trips/s of the 2^14-sample workload (with its kernel roofline and the
PCIe-inclusive rate), the N = 2^18 pulsed single instance, the 256-instance
degenerate-PPLN batch (FP64 round trips, and FP32C to steady state from
noise), and the other CPU baselines.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s


def workload_c2(N=1 << 14, NZ=100, NRT=5000):
    import numpy as np
    import cuopo_amd as P
    # grating period of the sweep script's sPPLT line (src/cuOPO.sh:75) with the
    # opt-in three-wave phase-matching scan
    return P.Args(three_eqs=True, regime=P.CW, crystal=P.SPPLT, N=N, NZ=NZ, NRT=NRT, lp_nm=532.0,
                  Lambda=7.97, Nth=3.0, R_pct=98.0, delta=0.3 / np.pi, lcr_mm=20.0, ls_um=1.0,
                  pm_mode=P.PM_THREE_WAVE)


def workload_c1(N=1 << 12, NZ=32, NRT=2000):
    """Continuous-wave two-equation MgO:PPLN OPO (532 -> 1064 nm), 5 mm
    crystal, 2 W of pump (the threshold ratio follows from the threshold)."""
    import cuopo_amd as P
    a = P.Args(three_eqs=False, regime=P.CW, crystal=P.PPLN, N=N, NZ=NZ, NRT=NRT, lp_nm=532.0, Lambda=6.97,
               Nth=1.0, R_pct=98.0, lcr_mm=5.0)
    a.Nth = 2.0 / P.derive(a).Pth
    return a


def workload_c5(N=1 << 18, NZ=200, NRT=40):
    """Nanosecond-pulsed three-equation MgO:PPLN OPO, EOM and GDD compensation
    on (4x threshold, EOM depth 0.8 rad, modulation at 4 x FSR), 5 mm crystal."""
    import numpy as np
    import cuopo_amd as P
    return P.Args(three_eqs=True, regime=P.NS, crystal=P.PPLN, N=N, NZ=NZ, NRT=NRT, lp_nm=532.0, Lambda=6.97,
                  Nth=4.0, R_pct=98.0, gdd_pct=100.0, upm=1, md=0.8 / np.pi, fm=4.0, lcr_mm=5.0)


TUNING = None  # (col_threads, row_threads, xcd_pack, phase_mode[, col_spt, row_spt]) instead of autotuning


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tune(obj):
    """Autotune a Solver or Batch, or apply the fixed --tuning settings."""
    if TUNING is None:
        return obj.autotune(iters=10)
    c, r, x, ph = TUNING[:4]
    ce, re_ = TUNING[4:6] if len(TUNING) == 6 else (0, 0)
    kw = dict(col_threads=c, row_threads=r, xcd_pack=x, phase_mode=ph, col_spt=ce, row_spt=re_)
    if hasattr(obj, "count"):
        for t in obj.get_tuning():
            obj.set_tuning(t["N"], **kw)
        return obj.get_tuning()
    obj.set_tuning(**kw)
    return obj.get_tuning()


def b_rt(N, NZ, nf, n_res, sc):
    """Algorithmic bytes of one round trip: NZ slices reading and writing every
    field once, plus the boundary (resonant fields read+write, pump read)."""
    return NZ * 2 * nf * N * sc + (2 * n_res + 1) * N * sc


def c5_section(steps=20):
    """Plain FP64 single instance at N = 2^18 (bandwidth-meaningful shape)."""
    import numpy as np
    import cuopo_amd as P
    args = workload_c5(NRT=steps + 6)
    inst = P.setup_instance(args, P.FP64, seed=5, noise_sigma=1e-20)
    d = inst["derived"]
    s = P.Solver(inst["params"])
    s.upload(inst["w"], inst["T"], inst["Ap_in"], inst["Ap"], inst["As"], inst["Ai"])
    log("c5: tuning")
    tuning = tune(s)
    log("c5: running")
    s.run(0, 3)
    t0 = time.perf_counter()
    s.run(3, steps)
    wall = time.perf_counter() - t0
    log("c5: timing the stages")
    col_us, row_us = s.time_stages(50)
    s.close()
    bz = 2 * 3 * d.N * 16
    _, copy_us = P.copy_probe(3 * d.N * 16, 200)  # one copy launch of the same bytes
    rt_s = steps / wall
    return {"workload": "NS three-equation MgO:PPLN, N=2^18, 200 slices, EOM + GDD, FP64",
            "round_trips_per_s": rt_s, "col_stage_us": col_us, "row_stage_us": row_us,
            "col_frac": bz / (col_us * 1e-6) / HBM_PEAK, "row_frac": bz / (row_us * 1e-6) / HBM_PEAK,
            "copy_same_bytes_us": copy_us, "copy_same_bytes_frac": bz / (copy_us * 1e-6) / HBM_PEAK,
            # HBM bytes per dispatch from the committed counter summary (FETCH_SIZE + WRITE_SIZE)
            "col_traffic": load_traffic("c5_col_stage"), "row_traffic": load_traffic("c5_row_stage"),
            "round_trip_frac": b_rt(d.N, d.NZ, 3, 2, 16) * rt_s / HBM_PEAK, "tuning": tuning}


def c3_section(rts=10, fp32c=False):
    """The 256-point sweep family as one batch (FP64, or FP32C as the sweep
    runs it), fixed round-trip count."""
    import numpy as np
    import cuopo_amd as P
    from cuopo_amd import sweep as SW
    pts = SW.sweep_points(256, seed=3)
    sc = 8 if fp32c else 16
    b, made = SW.make_batch(pts, max_rt=rts + 10, precision=P.FP32C if fp32c else P.FP64)
    log("c3: tuning")
    tuning = tune(b)
    log("c3: running")
    b.run(2, check_every=0, tol=0.0)
    t0 = time.perf_counter()
    b.run(rts, check_every=0, tol=0.0)
    wall = time.perf_counter() - t0
    groups = []
    for N in SW.SWEEP_N:
        cnt = sum(1 for p in pts if p[0] == N)
        c, r = b.time_stages(N, 20)
        bz = 2 * 3 * N * sc * cnt
        _, cp = P.copy_probe(3 * N * sc * cnt, 50)  # one copy launch of the same bytes
        groups.append({"N": N, "count": cnt, "col_us": c, "row_us": r, "col_frac": bz / (c * 1e-6) / HBM_PEAK,
                       "row_frac": bz / (r * 1e-6) / HBM_PEAK, "copy_same_bytes_us": cp,
                       "copy_same_bytes_frac": bz / (cp * 1e-6) / HBM_PEAK})
    b.close()
    total = sum(b_rt(p[0], 100, 3, 2, sc) for p in pts)
    prec = "FP32C" if fp32c else "FP64"
    return {"workload": f"256 CW three-equation MgO:PPLN instances, N in 2^12..2^15, 100 slices, {prec}, one batch",
            "round_trips_per_s_per_instance": rts / wall, "round_trip_frac": total * rts / wall / HBM_PEAK,
            "groups": groups, "tuning": tuning}


def load_traffic(name):
    """HBM bytes per launch of `name` from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def cpu_host():
    """(CPU model, cores this process may use, threads the all-cores run uses:
    the affinity set capped at OMP_NUM_THREADS, 16 when unset)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    ncpu = len(os.sched_getaffinity(0))
    return model, ncpu, max(1, min(ncpu, int(os.environ.get("OMP_NUM_THREADS") or 16)))


def oracle_instance(args, seed=1):
    """The FP64 oracle's inputs for one workload (seeded 1e-20 noise)."""
    from oracle import oracle as O
    cfg = O.Config(**{k: getattr(args, k) for k in args.__dataclass_fields__})
    p = O.setup(cfg)
    T, w = O.grids(p)
    Ap_in = O.pump_input(p)
    As = O.noise(seed, 0, 1e-20, p.N)
    Ai = O.noise(seed, 1, 1e-20, p.N) if cfg.three_eqs else None
    return p, w, T, Ap_in, Ap_in[: p.N].copy(), As, Ai


def oracle_rate(inst, seconds, start=None):
    """Round trips/s of the oracle on one thread over at least `seconds`
    (timer around the round-trip calls only)."""
    from oracle import oracle as O
    p, w, T, Ap_in, Ap, As, Ai = inst
    if start is not None:
        start.wait()
    n = 0
    t0 = time.perf_counter()
    while True:
        O.run(p, w, T, Ap_in, Ap, As, Ai, n % p.NRT, 1)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            return n / dt


def median_rate(inst, seconds, reps=5):
    import numpy as np
    return float(np.median([oracle_rate(inst, seconds / reps) for _ in range(reps)]))


def cpu_c2_1core(args_c2, seconds=12.0):
    """The FP64 oracle (a scalar C transcription of the reference kernels with
    the project's radix-2 FFT) on one core, round trips of the 2^14-sample
    workload, median of 5 timings."""
    model, ncpu, _ = cpu_host()
    v = median_rate(oracle_instance(args_c2), seconds)
    return {"value": v, "unit": "round trips/s", "cores": 1, "kind": "port",
            "sample": f"median of 5 timings of >= {seconds / 5:.1f} s each, round trips of the same workload, "
                      f"FP64 oracle (radix-2 FFT, no FFTW on the host), single thread on {model} "
                      f"({ncpu} cores available)"}


def threaded_rates(insts, seconds):
    """Round trips/s of the oracle on each of len(insts) threads at once (the
    oracle releases the interpreter lock in its C calls)."""
    import threading
    rates = [0.0] * len(insts)
    start = threading.Barrier(len(insts))

    def work(k):
        rates[k] = oracle_rate(insts[k], seconds, start)

    th = [threading.Thread(target=work, args=(k,)) for k in range(len(insts))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return rates


def cpu_headline(cold, seconds=6.0):
    """CPU baseline in the headline's unit: the FP64 oracle runs the plain
    cold-start method one point per core on all cores.  Sample: every thread
    times round trips of one sweep point at the workload's grid for `seconds`;
    the per-core rate then prices the round trips each subset point needed in
    the per-instance FP64 cold start (the same algorithm), spread evenly over
    the threads, and the subset's steady points are counted."""
    from cuopo_amd import surrogate as S
    model, ncpu, nthr = cpu_host()
    pts = cold["points"]
    insts = [oracle_instance(S.family_args(*pts[k % len(pts)], N=1 << 14, NZ=100, NRT=SWEEP_CAP))
             for k in range(nthr)]
    rates = threaded_rates(insts, seconds)
    per_core = float(np.mean(rates))
    core_s = float(sum(cold["rts"]) / per_core)
    wall = core_s / nthr
    steady = int(sum(1 for st in cold["states"] if st == 1))
    return {"value": steady / wall, "unit": "sweep points/s to steady state", "cores": nthr, "kind": "port",
            "sample": f"{nthr} threads x {seconds:.0f} s of FP64 oracle round trips at 2^14 samples x 100 slices "
                      f"({per_core:.3f} round trips/s per core, radix-2 FFT, no FFTW) pricing the "
                      f"{int(sum(cold['rts']))} round trips the {len(pts)}-point cold-start subset ran to "
                      f"{steady} steady states; {model}, {ncpu} cores available",
            "round_trips_per_s_per_core": per_core}


def cpu_extra(seconds=8.0, c3_cold=None):
    """The other CPU baselines: C1 and C5 single instance on one core (C5 timed
    on 20 slices per pass and scaled to 200: the boundary work is counted ten
    times, so the C5 figure is slightly low), and the degenerate-PPLN sweep
    family with one instance per core on all cores: round trips/s per core per
    grid size and, when its cold sweep ran, the points/s (any finish state)
    those rates give for the round trips each point ran there."""
    from cuopo_amd import sweep as SW
    from cuopo_amd.surrogate import sweep_args
    model, ncpu, nthr = cpu_host()
    out = {"host": model, "cores_available": ncpu, "threads_all_cores": nthr}
    out["c1_1core_round_trips_per_s"] = median_rate(oracle_instance(workload_c1()), seconds / 2)
    c5 = workload_c5(NRT=40)
    c5.NZ = 20
    out["c5_1core_round_trips_per_s"] = median_rate(oracle_instance(c5), seconds / 2, reps=3) * 20 / 200
    pts = SW.sweep_points(256, seed=3)
    firsts = {N: next(p for p in pts if p[0] == N) for N in SW.SWEEP_N}
    Ns = [SW.SWEEP_N[k % len(SW.SWEEP_N)] for k in range(nthr)]
    insts = []
    for N in Ns:
        _, nth, delta, lcr, gdd, R = firsts[N]
        insts.append(oracle_instance(sweep_args(nth, delta, lcr, gdd, R, N=N, NZ=100, NRT=5000)))
    rates = threaded_rates(insts, seconds)
    per_n = {N: float(np.mean([r for r, n in zip(rates, Ns) if n == N])) for N in SW.SWEEP_N if N in Ns}
    c3 = {"round_trips_per_s_per_core": {str(N): v for N, v in per_n.items()}, "threads": nthr}
    if c3_cold is not None and len(per_n) == len(SW.SWEEP_N):
        sweep_pts, rts = c3_cold
        core_s = float(sum(rts[k] / per_n[p[0]] for k, p in enumerate(sweep_pts)))
        c3["finished_points_per_s"] = len(sweep_pts) / (core_s / nthr)
    out["c3_sweep_all_cores"] = c3
    return out


def c3_sweep_cold():
    """The degenerate-PPLN 256-point family (N in 2^12..2^15, 100 slices) as
    one FP32C batch from 1e-20 noise to the steady / off tests or 5000 round
    trips.  Its steady states are multimode and drift slowly, so few pass the
    test; reported as finish-state counts and round trips."""
    import cuopo_amd as P
    from cuopo_amd import sweep as SW
    pts = SW.sweep_points(256, seed=3)
    r = SW.run_batched(pts, precision=P.FP32C)
    r["batch"].close()
    st = r["states"]
    rts = np.where(r["rts"] > 0, r["rts"], 5000)
    return {"workload": "256 CW three-equation degenerate MgO:PPLN points, N in 2^12..2^15, 100 slices, FP32C, "
                        "from 1e-20 noise, cap 5000",
            "run_s": r["run_s"], "steady": int((st == 1).sum()), "off": int((st == 2).sum()),
            "capped": int((st == 0).sum()), "diverged": int((st == -1).sum()),
            "steady_points_per_s": float((st == 1).sum() / r["run_s"]),
            "finished_points_per_s": float(((st == 1) | (st == 2)).sum() / r["run_s"]),
            "mean_round_trips": float(rts.mean())}, (pts, rts)


FAMILY_N, FAMILY_NZ, SWEEP_CAP = 1 << 14, 100, 20000
SWEEP_ACCEL = (150, 300)  # extrapolation every 150 round trips from 300 on (opo_batch_set_accel)


def headline_sweep(weights, count=256, subset=12, seed=31):
    """The headline: steady points/s of the learned-warm-start FP64 batch on
    `count` points of the sPPLT family, and of the plain per-instance FP64
    cold start on `subset` of the same points (every count/subset-th)."""
    import cuopo_amd as P
    from cuopo_amd import surrogate as S
    from cuopo_amd import sweep as SW
    from cuopo_amd.train_surrogate import load
    pts = S.sample_points(count, seed)  # the network was trained on other seeds
    N, NZ, cap = FAMILY_N, FAMILY_NZ, SWEEP_CAP
    sur, meta = load(weights)
    ds = [S.derive(S.family_args(*p, N=N, NZ=NZ, NRT=cap)) for p in pts]
    # timed: network inference on the GPU + decoding into fields, then the batch
    t0 = time.perf_counter()
    warm = sur.warm_fields(pts, N, [d.Ap0 for d in ds], seed=seed)
    t_infer = time.perf_counter() - t0
    sur.net.close()
    b, _ = S.family_batch(pts, N, NZ, cap, P.FP64, warm=warm)  # upload (PCIe): outside the timed region
    log("headline: tuning")
    tuning = tune(b)
    b.set_accel(*SWEEP_ACCEL)  # round-trip extrapolation of the slowly converging points
    log("headline: warm batch")
    t1 = time.perf_counter()
    rts, st = b.run_ex(cap, check_every=50, tol=1e-6, off_ratio=1e-6)
    t_run = time.perf_counter() - t1
    power = b.power()
    dev_ms = b.last_run_ms()
    col_us, row_us = b.time_stages(N, 20)
    b.close()
    arm_s = t_infer + t_run
    steady = np.flatnonzero(st == 1)
    # plain per-instance FP64 cold starts on the subset, one after another
    log("headline: per-instance FP64 cold starts")
    idx = list(range(0, count, max(1, count // subset)))[:subset]
    cold = {"points": [[float(v) for v in pts[k]] for k in idx], "rts": [], "states": [], "wall_s": []}
    agree = []
    for j, k in enumerate(idx):
        o = SW.run_single_fp64(None, max_rt=cap, seed=1000, k=j, args=S.family_args(*pts[k], N=N, NZ=NZ, NRT=cap))
        cold["rts"].append(int(o["rts"]))
        cold["states"].append(int(o["state"]))
        cold["wall_s"].append(float(o["wall_s"]))
        if o["state"] == 1 and st[k] == 1:
            agree.append(abs(power[k] - o["power"]) / o["power"])
        log(f"headline: cold point {j + 1}/{len(idx)}: state {o['state']} after {o['rts']} round trips")
    cold_s = float(sum(cold["wall_s"]))
    cold_steady = int(sum(1 for v in cold["states"] if v == 1))
    warm_rate = float(steady.size / arm_s)
    cold_rate = cold_steady / cold_s
    sub_warm_steady = int(sum(1 for k in idx if st[k] == 1))
    # one read + one write of every field of the instances one stage launch covers (a batch of one grid
    # size runs as two groups of half the points, each launched on its own stream; time_stages times the first)
    launch_n = int(tuning[0]["count"])
    bz = 2 * 3 * N * 16 * launch_n
    return {
        "value": warm_rate,
        "per_instance_fp64_cold": {
            "points": len(idx), "steady": cold_steady, "off": int(sum(1 for v in cold["states"] if v == 2)),
            "capped": int(sum(1 for v in cold["states"] if v == 0)), "run_s": cold_s,
            "steady_points_per_s": cold_rate, "round_trips": cold["rts"],
            "mean_round_trips": float(np.mean(cold["rts"])),
            "seconds_per_point": cold_s / len(idx)},
        "warm_batch": {"points": count, "steady": int(steady.size), "off": int((st == 2).sum()),
                       "capped": int((st == 0).sum()), "diverged": int((st == -1).sum()),
                       "inference_s": t_infer, "run_s": t_run, "device_ms": dev_ms,
                       "mean_round_trips": float(np.mean(np.where(rts > 0, rts, cap))),
                       "median_round_trips_steady": float(np.median(rts[steady])) if steady.size else None,
                       "subset_steady": sub_warm_steady, "tuning": tuning,
                       "accel": {"every": SWEEP_ACCEL[0], "start": SWEEP_ACCEL[1]}},
        "speedup_vs_per_instance": (warm_rate / cold_rate) if cold_rate > 0 else None,
        "speedup_lower_bound_if_cold_none": (warm_rate * cold_s) if cold_rate == 0 else None,
        "warm_vs_cold_power_where_both_steady": {
            "points": len(agree), "median": float(np.median(agree)) if agree else None,
            "max": float(np.max(agree)) if agree else None,
            "within_1e-4": int(sum(1 for v in agree if v <= 1e-4)), "values": agree},
        "surrogate": {k: meta[k] for k in ("family", "points", "N", "NZ", "steady_states", "epochs", "val")
                      if k in meta},
        "cold_subset": cold,
        "kernel": {"col_us": col_us, "row_us": row_us, "bz": bz, "launch_instances": launch_n,
                   "col_frac": bz / (col_us * 1e-6) / HBM_PEAK, "row_frac": bz / (row_us * 1e-6) / HBM_PEAK},
    }


def sweep_kernels(weights, count=256, rts=10, seed=31):
    """The headline batch alone, for profilers: the warm-started FP64 batch
    of the sPPLT family, tuned, `rts` round trips without the convergence
    test, then each stage timed alone (20 launches) as the headline does."""
    import cuopo_amd as P
    from cuopo_amd import surrogate as S
    from cuopo_amd.train_surrogate import load
    pts = S.sample_points(count, seed)
    N, NZ = FAMILY_N, FAMILY_NZ
    sur, _ = load(weights)
    warm = sur.warm_fields(pts, N, [S.derive(S.family_args(*p, N=N, NZ=NZ)).Ap0 for p in pts], seed=seed)
    sur.net.close()
    b, _ = S.family_batch(pts, N, NZ, rts + 10, P.FP64, warm=warm)
    tuning = tune(b)
    t0 = time.perf_counter()
    b.run(rts, check_every=0, tol=0.0)
    wall = time.perf_counter() - t0
    col_us, row_us = b.time_stages(N, 20)
    b.close()
    bz = 2 * 3 * N * 16 * int(tuning[0]["count"])  # the first group's launch (see headline_sweep)
    return {"workload": f"{count}-point sPPLT sweep batch, N=2^14, 100 slices, FP64", "round_trips": rts,
            "round_trips_per_s_per_instance": rts / wall, "col_stage_us": col_us, "row_stage_us": row_us,
            "col_frac": bz / (col_us * 1e-6) / HBM_PEAK, "row_frac": bz / (row_us * 1e-6) / HBM_PEAK,
            "tuning": tuning}


def c2_section(a):
    """Single-instance round trips/s of the 2^14-sample sPPLT workload (3x
    threshold, 0.3 rad), plain FP64 cold start, with its kernel roofline."""
    import cuopo_amd as P
    args = workload_c2(NRT=a.warmup + a.steps + 1)
    inst = P.setup_instance(args, P.FP64, seed=1, noise_sigma=1e-20)
    d = inst["derived"]
    s = P.Solver(inst["params"])
    s.upload(inst["w"], inst["T"], inst["Ap_in"], inst["Ap"], inst["As"], inst["Ai"])
    log("c2: tuning")
    tuning = s.autotune(iters=20) if TUNING is None else tune(s)
    log("c2: running")
    s.run(0, max(a.warmup, 3))  # warm-up (also builds the round-trip graph)
    t0 = time.perf_counter()
    s.run(a.warmup, a.steps)  # graph replays; returns after the stream's end event
    wall = time.perf_counter() - t0
    dev_ms = s.last_run_ms()
    rts = a.steps / wall
    t1 = time.perf_counter()
    Ap, As, Ai = s.download()
    down_s = time.perf_counter() - t1
    if not (np.isfinite(As).all() and np.isfinite(Ap).all()):
        raise SystemExit("non-finite fields after the timed run")
    t1 = time.perf_counter()
    s.upload(inst["w"], inst["T"], inst["Ap_in"], inst["Ap"], inst["As"], inst["Ai"])
    up_s = time.perf_counter() - t1
    col_us, row_us = s.time_stages(200)
    s.close()
    N, nf, sc = d.N, 3, 16
    Bz = 2 * nf * N * sc
    dom, dom_us = ("col_stage", col_us) if col_us >= row_us else ("row_stage", row_us)
    copy_gbps, _ = P.copy_probe(256 << 20, 20)
    _, copy_bz_us = P.copy_probe(nf * N * sc, 200)
    return {"workload": "CW sPPLT three-equation OPO, N=2^14 complex128 samples, 100 z-slices, 20 mm crystal, "
                        "pump 3x threshold, detuning 0.3 rad, 532->1000+1137 nm, one instance",
            "round_trips_per_s": rts, "ms_per_round_trip": 1e3 * wall / a.steps, "steps": a.steps,
            "device_ms_timed_region": dev_ms, "col_stage_us": col_us, "row_stage_us": row_us,
            "roofline": {"kernel": dom, "kernel_us": dom_us, "achieved_GBps": Bz / (dom_us * 1e-6) / 1e9,
                         "frac": Bz / (dom_us * 1e-6) / HBM_PEAK, "traffic": load_traffic(f"c2_{dom}")},
            "graph_slice_us": 1e3 * dev_ms / (a.steps * args.NZ),
            "copy_probe_GBps": copy_gbps, "copy_same_bytes_us": copy_bz_us, "tuning": tuning,
            "upload_ms": 1e3 * up_s, "download_ms": 1e3 * down_s,
            "round_trips_per_s_incl_pcie": a.steps / (wall + up_s + down_s),
            "round_trip_frac_of_hbm": b_rt(N, args.NZ, nf, 2, sc) * rts / HBM_PEAK,
            "algorithmic_fp64_tflops": nf * (10 * math.log2(N) + 156) * N * args.NZ * rts / 1e12}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed round trips of the single-instance section")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-shapes", action="store_true", help="skip the N=2^18 and batched PPLN sections")
    ap.add_argument("--subset", type=int, default=12, help="points of the per-instance FP64 cold-start baseline")
    ap.add_argument("--c5-steps", type=int, default=20, help="timed round trips of the N=2^18 section")
    ap.add_argument("--tuning", default=None,
                    help="col_threads,row_threads,xcd_pack,phase_mode[,col_spt,row_spt] applied instead of autotuning "
                         "(profiling runs: the profiler perturbs the autotuner's timings)")
    ap.add_argument("--only", choices=["c5", "c3", "c3f", "c2", "sweep"], default=None,
                    help="run only the N=2^18 (c5), the batched FP64 (c3) or FP32C (c3f) PPLN section, the "
                         "single-instance 2^14 section (c2) or the headline batch's kernels (sweep) and print its "
                         "JSON (for profiling those shapes)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "cuopo_amd", "data", "surrogate.npz"))
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this solver runs one instance per GPU and is benchmarked on one GPU; "
                                   "--gpus above 1 is not supported (independent sweep points shard "
                                   "across GPUs without communication)"}))
        sys.exit(2)

    global TUNING
    if a.tuning:
        TUNING = tuple(int(v) for v in a.tuning.split(","))
        if len(TUNING) not in (4, 6):
            raise SystemExit("--tuning wants col_threads,row_threads,xcd_pack,phase_mode[,col_spt,row_spt]")
    if a.only:
        fn = {"c5": lambda: c5_section(a.c5_steps), "c3": c3_section, "c3f": lambda: c3_section(fp32c=True), "c2": lambda: c2_section(a),
              "sweep": lambda: sweep_kernels(a.weights)}
        print(json.dumps(fn[a.only](), default=jsonable))
        return

    h = headline_sweep(a.weights, subset=a.subset)
    k = h["kernel"]
    dom, dom_us = ("col_stage", k["col_us"]) if k["col_us"] >= k["row_us"] else ("row_stage", k["row_us"])
    achieved = k["bz"] / (dom_us * 1e-6)
    value = h.pop("value")
    out = {
        "metric": "sweep points/s to steady state (learned warm start + batched FP64, vs plain per-instance FP64 "
                  "cold start)",
        "value": value,
        "unit": "sweep points/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": 1e3 / value if value > 0 else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (sweep points drawn from a seeded generator; cold starts from seeded 1e-20 complex "
                "Gaussian noise; random-init-free: the network's weights are the committed trained ones)",
        "config": {"workload": "256-point CW three-equation sPPLT sweep (532 -> 1000 + 1137 nm; pump 1.2-5x "
                               "threshold, detuning -pi..pi rad, crystal 5-40 mm, GDD -500..500 fs^2, R 0.90-0.99), "
                               "N=2^14 complex128 samples, 100 z-slices, one batch from the learned warm start to "
                               "the steady test (|dP| < 1e-6 P over 50 round trips, round-trip extrapolation every "
                               f"{SWEEP_ACCEL[0]} round trips from {SWEEP_ACCEL[1]} on), cap {SWEEP_CAP} round trips",
                   "points": 256, "N": FAMILY_N, "NZ": FAMILY_NZ, "cap": SWEEP_CAP,
                   "per_instance_subset": a.subset},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": load_traffic(f"sweep_fp64_{dom}"),
                     "kernel": dom, "kernel_us": dom_us, "algorithmic_bytes_per_launch": k["bz"],
                     "launch_instances": k["launch_instances"]},
        "extra": {"sweep": h},
    }
    out["extra"]["c2_single_fp64"] = c2_section(a)
    c3_cold = None
    if not a.no_shapes:
        out["extra"]["c5_single_fp64"] = c5_section(a.c5_steps)
        out["extra"]["c3_batched_fp64"] = c3_section()
        log("c3 sweep from noise")
        out["extra"]["c3_sweep_fp32c_cold"], c3_cold = c3_sweep_cold()
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_headline(h["cold_subset"])
        log("cpu baselines: C2 one core, C1, C5, PPLN sweep on all cores")
        out["extra"]["cpu"] = cpu_extra(c3_cold=c3_cold)
        out["extra"]["cpu"]["c2_1core"] = cpu_c2_1core(workload_c2(), a.cpu_seconds)
    print(json.dumps(out, default=jsonable))


def jsonable(o):
    """numpy scalars and arrays in the output line"""
    if hasattr(o, "tolist"):
        return o.tolist()
    raise TypeError(f"{type(o).__name__} is not JSON serialisable")


if __name__ == "__main__":
    main()
