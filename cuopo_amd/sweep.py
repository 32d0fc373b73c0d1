"""Batched parameter sweeps: the work src/cuOPO.sh:91-175 runs as one process
per point, run here as one batch of instances on one GPU.

The sweep family (continuous wave, three equations, MgO:PPLN, 532 nm pump,
20 mm crystal, 100 z-slices per pass) draws per point: grid size N from
{2^12, 2^13, 2^14, 2^15}, pump power uniform in [1.2, 5] x threshold, cavity
detuning uniform in [-pi, pi] rad and signal reflectivity uniform in
[0.90, 0.99].  A point is finished when its intracavity signal power changes
by less than `tol` over `check_every` round trips (steady state), when it has
decayed below threshold (off: power a factor 1/off_ratio below its first
check with a settled decay rate; seeded white noise can lose up to ~N before
a growing mode dominates, so off_ratio = 1e-6 leaves margin for N <= 2^15),
or at `max_rt`.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import FP32C, FP64, Batch, Solver, derive, grids, noise, params_from, pump_input
from .surrogate import sweep_args

SWEEP_N = (1 << 12, 1 << 13, 1 << 14, 1 << 15)


def sweep_points(count=256, seed=3):
    """Rows of (N, nth_rel, delta_rad, lcr_mm, gdd_fs2, R)."""
    rng = np.random.default_rng(seed)
    N = rng.choice(SWEEP_N, count)
    nth = rng.uniform(1.2, 5.0, count)
    delta = rng.uniform(-math.pi, math.pi, count)
    R = rng.uniform(0.90, 0.99, count)
    return [(int(N[k]), float(nth[k]), float(delta[k]), 20.0, 0.0, float(R[k])) for k in range(count)]


def _setup(pt, NZ, max_rt, precision, seed, k, sigma, warm):
    N, nth, delta, lcr, gdd, R = pt
    a = sweep_args(nth, delta, lcr, gdd, R, N=N, NZ=NZ, NRT=max_rt)
    d = derive(a)
    p = params_from(d, precision)
    T, w = grids(d)
    Ap_in = pump_input(d)
    if warm is not None:
        As, Ai = warm[k]
    else:
        As = noise(seed + k, 0, sigma, N)
        Ai = noise(seed + k, 1, sigma, N)
    return p, d, (w, T, Ap_in, Ap_in[:N].copy(), As, Ai)


def make_batch(points, NZ=100, max_rt=5000, precision=FP32C, seed=1000, sigma=1e-20, warm=None, tiles=(0, 0),
               autotune=False):
    """An uploaded Batch of the sweep points (and the per-point setup); with
    autotune, the per-grid-size launch settings are timed and chosen first."""
    made = [_setup(pt, NZ, max_rt, precision, seed, k, sigma, warm) for k, pt in enumerate(points)]
    b = Batch([m[0] for m in made])
    if tiles != (0, 0):
        b.set_tiles(*tiles)
    for k, m in enumerate(made):
        b.upload(k, *m[2])
    if autotune:
        b.autotune(iters=5)
    return b, made


def run_batched(points, NZ=100, max_rt=5000, check_every=50, tol=1e-6, off_ratio=1e-6, precision=FP32C,
                seed=1000, sigma=1e-20, warm=None, tiles=(0, 0), autotune=True, accel=None):
    """All points in one batch.  Returns a dict with per-point round trips,
    states (1 steady, 2 off, 0 capped, -1 diverged), final powers, the
    device time of the run and the wall time including uploads.  accel =
    (every, start) turns on the batch's round-trip extrapolation
    (Batch.set_accel)."""
    t0 = time.perf_counter()
    b, made = make_batch(points, NZ, max_rt, precision, seed, sigma, warm, tiles, autotune)
    if accel:
        b.set_accel(*accel)
    t1 = time.perf_counter()
    rts, st = b.run_ex(max_rt, check_every=check_every, tol=tol, off_ratio=off_ratio)
    t2 = time.perf_counter()
    power = b.power()
    out = dict(rts=rts, states=st, power=power, device_ms=b.last_run_ms(), run_s=t2 - t1, total_s=t2 - t0,
               ap0=np.array([m[1].Ap0 for m in made]), batch=b)
    return out


def run_warm(points, warm, seed=1000, sigma=1e-20, **kw):
    """run_batched from warm-start fields, with the cold start as the
    fallback: every point whose warm run diverged (non-finite power) runs
    again from the noise the cold start gives it (seed + its index), in a
    second batch.  Its round trips add up (the warm run's up to the
    divergence, plus the fallback's, a capped fallback counting max_rt), its
    state and power are the fallback's, and run_s covers both runs.  Returns
    run_batched's dict plus "fallback" (the indices that fell back) and
    "fallback_batch" (the second batch, or None; the fallback points' fields
    live there, in that order)."""
    r = run_batched(points, seed=seed, sigma=sigma, warm=warm, **kw)
    bad = np.flatnonzero(r["states"] == -1)
    r["fallback"], r["fallback_batch"] = bad, None
    if bad.size:
        max_rt = kw.get("max_rt", 5000)
        cold = [(noise(seed + int(k), 0, sigma, points[k][0]), noise(seed + int(k), 1, sigma, points[k][0]))
                for k in bad]
        f = run_batched([points[k] for k in bad], seed=seed, sigma=sigma, warm=cold, **kw)
        r["rts"] = r["rts"].copy()
        # a diverged instance reports the round trip it stopped at; a capped
        # fallback reports -1 and ran max_rt
        r["rts"][bad] = r["rts"][bad] + np.where(f["rts"] > 0, f["rts"], max_rt)
        r["states"] = r["states"].copy()
        r["states"][bad] = f["states"]
        r["power"] = r["power"].copy()
        r["power"][bad] = f["power"]
        r["run_s"] += f["run_s"]
        r["total_s"] += f["total_s"]
        r["fallback_batch"] = f["batch"]
    return r


def run_single_fp64(pt, NZ=100, max_rt=5000, check_every=50, tol=1e-6, off_ratio=1e-6, seed=1000, k=0,
                    sigma=1e-20, autotune=True, args=None):
    """Plain per-instance FP64 cold start: one solver, round trips in chunks of
    check_every with the same convergence and off tests evaluated on the host
    from the power history (the way a per-point driver would).  pt is a row of
    sweep_points, or with `args` given (an Args of any family) ignored."""
    if args is None:
        p, d, up = _setup(pt, NZ, max_rt, FP64, seed, k, sigma, None)
    else:
        d = derive(args)
        p = params_from(d, FP64)
        T, w = grids(d)
        Ap_in = pump_input(d)
        up = (w, T, Ap_in, Ap_in[:d.N].copy(), noise(seed + k, 0, sigma, d.N), noise(seed + k, 1, sigma, d.N))
    s = Solver(p)
    s.upload(*up)
    if autotune:
        s.autotune(iters=10)
    t0 = time.perf_counter()
    done, state, prev, first, rprev = 0, 0, 0.0, 0.0, 0.0
    while done < max_rt:
        n = min(check_every, max_rt - done)
        s.run(done, n)  # the fields stay on the device between calls
        done += n
        P = float(s.power_history(n)[-1])
        if prev > 0 and P > 0 and abs(P - prev) <= tol * P:
            state = 1
            break
        if prev > 0 and first > 0:
            r = P / prev
            if r < 1 and rprev > 0 and abs(r - rprev) <= 1e-3 * r and P < off_ratio * first:
                state = 2
                break
        if prev > 0:
            rprev = P / prev
        if first == 0:
            first = P
        prev = P
    wall = time.perf_counter() - t0
    fields = s.download()
    s.close()
    return dict(rts=done, state=state, power=P, wall_s=wall, fields=fields)
