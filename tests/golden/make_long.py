"""Golden records of full-size workloads over their long runs, made once with
the FP64 oracle:

    python -m tests.golden.make_long c2     # about 7 minutes on one core
    python -m tests.golden.make_long c1     # about 75 seconds
    python -m tests.golden.make_long ns     # about 30 seconds

c2: the C2 grid and crystal (3 eqs, CW sPPLT, N = 2^14, 100 slices per pass,
0.3 rad detuning) at 10x threshold from 1e-6 seed noise, 1000 round trips, so
that the oscillator builds up and saturates within the run.
c1: the C1 workload as it is (2 eqs, CW PPLN, N = 2^12, 32 slices per pass,
2 W of pump, 1e-20 seed noise), its full 2000 round trips: build-up from
1e-37 to saturation.
ns: the pulsed three-equation PPLN workload with the EOM and GDD
compensation on (4x threshold, EOM depth 0.8 rad at 4 x FSR, GDD 100%) on
N = 2^12, 16 slices per pass, over NRT = 1000 round trips: every round trip
reads its own pump slot of the N x NRT pulse (the last one shifted by one
sample) and is saved.

Stored per workload: the intracavity signal power after every round trip, and
64 random projections of each final field (rows of +-1/sqrt(N) from a fixed
seed).  The relative L2 error of a field is estimated by that of its
projections (a few tens of percent spread), which is ample against a 1e-7 bar
while keeping the records small.  For ns, also the projections of the saved
pump and signal slots of the round trips in NS_SLOTS."""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "..")))

from oracle import oracle as O  # noqa: E402
from tests.helpers import c1, c2, c5, inputs  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NPROJ, PROJ_SEED = 64, 20251202
WORKLOADS = {
    "c2": dict(nrt=1000, seed=7, sigma=1e-6, file="c2_long_1000rt.json"),
    "c1": dict(nrt=2000, seed=7, sigma=1e-20, file="c1_long_2000rt.json"),
    "ns": dict(nrt=1000, seed=7, sigma=1e-3, file="ns_long_1000rt.json"),
}
# saved round trips recorded for ns: start, pulse rise, peak (slot 500 holds
# T = 0 of the pulse), fall, tail and the last (Q17-shifted) slot
NS_SLOTS = (0, 250, 480, 500, 520, 750, 999)


def config(name):
    w = WORKLOADS[name]
    if name == "c2":
        cfg = c2(NRT=w["nrt"])
        cfg.Nth = 10.0
    elif name == "ns":
        cfg = c5(N=4096, NZ=16, NRT=w["nrt"])
    else:
        cfg = c1(NRT=w["nrt"])
    return cfg


def path(name):
    return os.path.join(HERE, WORKLOADS[name]["file"])


def projector(N):
    """64 x N matrix of +-1/sqrt(N), fixed by PROJ_SEED."""
    r = np.random.default_rng(PROJ_SEED).integers(0, 2, size=(NPROJ, N))
    return (2.0 * r - 1.0) / np.sqrt(N)


def make(name):
    w = WORKLOADS[name]
    cfg = config(name)
    p, T, wg, Ap_in, Ap, As, Ai = inputs(cfg, seed=w["seed"], sigma=w["sigma"])
    sv, pw = O.run(p, wg, T, Ap_in, Ap, As, Ai, 0, w["nrt"], save=name == "ns", power=True)
    R = projector(p.N)
    out = {"workload": name, "nrt": w["nrt"], "seed": w["seed"], "sigma": w["sigma"], "nth": cfg.Nth,
           "nproj": NPROJ, "proj_seed": PROJ_SEED, "power": [float(v) for v in pw]}
    fields = (("p", Ap), ("s", As), ("i", Ai)) if Ai is not None else (("p", Ap), ("s", As))
    for k, f in fields:
        y = R @ f
        out[f"proj_{k}_re"] = [float(v) for v in y.real]
        out[f"proj_{k}_im"] = [float(v) for v in y.imag]
        out[f"norm_{k}"] = float(np.linalg.norm(f))
    if sv is not None:
        out["slots"] = list(NS_SLOTS)
        for k in ("p", "s"):
            y = sv[k][list(NS_SLOTS)] @ R.T
            out[f"slot_{k}_re"] = [[float(v) for v in row] for row in y.real]
            out[f"slot_{k}_im"] = [[float(v) for v in row] for row in y.imag]
    with open(path(name), "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"{path(name)}: P0={pw[0]:.3e} P_end={pw[-1]:.3e}")


if __name__ == "__main__":
    for name in sys.argv[1:] or list(WORKLOADS):
        make(name)
