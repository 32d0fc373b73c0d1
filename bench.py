"""Benchmark of the SSFM round-trip solver on one MI355X.

Step = one cavity round trip (pump re-injection, NZ split-step slices of the
three coupled-wave equations through the crystal, cavity boundary) of one
instance, FP64 complex, fields resident in HBM.  Workload: continuous-wave
sPPLT OPO, three equations (pump 532 nm, signal 1000 nm, idler 1137 nm),
20 mm crystal, 2^14 complex128 samples per round trip, 100 z-slices per pass,
pump 3x threshold, detuning 0.3 rad, signal/idler seeded with 1e-20 complex
Gaussian noise.

Prints one JSON line: round trips/s, the roofline of the dominant kernel
(per-launch kernel time measured with start/stop events on the solver's
stream) and a CPU baseline (the FP64 oracle on this host's cores).
"""
import argparse
import json
import os
import sys
import time

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


def load_traffic(name):
    """HBM bytes per launch of `name` from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def cpu_baseline(args_c2, seconds=12.0):
    """FP64 oracle (a scalar C transcription of the reference kernels with the
    project's radix-2 FFT), one core, on round trips of the same workload."""
    from oracle import oracle as O
    cfg = O.Config(**{k: getattr(args_c2, k) for k in args_c2.__dataclass_fields__})
    p = O.setup(cfg)
    T, w = O.grids(p)
    Ap_in = O.pump_input(p)
    As = O.noise(1, 0, 1e-20, p.N)
    Ai = O.noise(1, 1, 1e-20, p.N)
    Ap = Ap_in.copy()
    n = 0
    t0 = time.perf_counter()
    while True:
        O.run(p, w, T, Ap_in, Ap, As, Ai, n, 1)
        n += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "round trips/s", "cores": 1, "kind": "port",
            "sample": f"{n} round trips of the same workload ({dt:.1f} s), FP64 oracle "
                      "(radix-2 FFT, no FFTW on the host), single thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this solver runs one instance per GPU and is benchmarked on one GPU; "
                                   "--gpus above 1 is not supported (independent sweep points shard "
                                   "across GPUs without communication)"}))
        sys.exit(2)

    import numpy as np
    import cuopo_amd as P

    args = workload_c2(NRT=a.warmup + a.steps + 1)
    inst = P.setup_instance(args, P.FP64, seed=1, noise_sigma=1e-20)
    d, prm = inst["derived"], inst["params"]
    s = P.Solver(prm)
    s.upload(inst["w"], inst["T"], inst["Ap_in"], inst["Ap"], inst["As"], inst["Ai"])
    # warm-up (also builds the round-trip graph used by the timed run)
    s.run(0, max(a.warmup, 3))
    # timed region: opo_run launches the graph replays and returns after the
    # stream's end event has completed (synchronous on both sides)
    t0 = time.perf_counter()
    s.run(a.warmup, a.steps)
    wall = time.perf_counter() - t0
    dev_ms = s.last_run_ms()
    rts = a.steps / wall
    Ap, As, Ai = s.download()
    if not (np.isfinite(As).all() and np.isfinite(Ap).all()):
        raise SystemExit("non-finite fields after the timed run")

    # roofline of the dominant kernel: per-launch times of the slice's two stages
    col_us, row_us = s.time_stages(200)
    N, nf, sc = d.N, 3, 16
    Bz = 2 * nf * N * sc  # one read + one write of every field: algorithmic bytes of one slice
    dom, dom_us = ("col_stage", col_us) if col_us >= row_us else ("row_stage", row_us)
    achieved = Bz / (dom_us * 1e-6)
    traffic = load_traffic(f"c2_{dom}")
    copy_gbps, _ = P.copy_probe(256 << 20, 20)
    out = {
        "metric": "round trips/s (plain FP64, one instance, cold start)",
        "value": rts,
        "unit": "round trips/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 1e-20 complex Gaussian noise on signal and idler, CW pump)",
        "config": {"workload": "CW sPPLT three-equation OPO, N=2^14 complex128 samples, 100 z-slices, "
                               "20 mm crystal, pump 3x threshold, detuning 0.3 rad, 532->1000+1137 nm",
                   "N": N, "NZ": args.NZ, "fields": nf, "round_trips_timed": a.steps},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": dom, "kernel_us": dom_us,
                     "algorithmic_bytes_per_launch": Bz},
        "extra": {"device_ms_timed_region": dev_ms, "col_stage_us": col_us, "row_stage_us": row_us,
                  "slice_frac_of_hbm": Bz / ((col_us + row_us) * 1e-6) / HBM_PEAK,
                  "copy_probe_GBps": copy_gbps,
                  "intracavity_power_end": float(np.sum(np.abs(As) ** 2))},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
