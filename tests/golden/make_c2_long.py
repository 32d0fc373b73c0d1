"""Golden fixture for the full-size three-equation workload over 1000 round
trips, made once with the FP64 oracle (about 7 minutes on one core):

    python -m tests.golden.make_c2_long

Workload: the C2 grid and crystal (3 eqs, CW sPPLT, N = 2^14, 100 slices per
pass, 0.3 rad detuning) at 10x threshold from 1e-6 seed noise, so that the
oscillator builds up and saturates within the 1000 round trips.  Stored: the
intracavity signal power after every round trip, and 64 random projections
of each final field (rows of +-1/sqrt(N) from a fixed seed): the relative L2
error of a field is estimated by that of its projections (a few tens of
percent spread), which is ample against a 1e-7 bar while keeping the fixture
small."""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "..")))

from oracle import oracle as O  # noqa: E402
from tests.helpers import c2, inputs  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "c2_long_1000rt.json")
NRT = 1000
SEED, SIGMA, NTH = 7, 1e-6, 10.0
NPROJ, PROJ_SEED = 64, 20251202


def config():
    cfg = c2(NRT=NRT)
    cfg.Nth = NTH
    return cfg


def projector(N):
    """64 x N matrix of +-1/sqrt(N), fixed by PROJ_SEED."""
    r = np.random.default_rng(PROJ_SEED).integers(0, 2, size=(NPROJ, N))
    return (2.0 * r - 1.0) / np.sqrt(N)


def main():
    cfg = config()
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=SEED, sigma=SIGMA)
    _, pw = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, NRT, power=True)
    R = projector(p.N)
    out = {"workload": "C2 grid and crystal, 10x threshold, seed 7, noise 1e-6, 1000 round trips",
           "nrt": NRT, "seed": SEED, "sigma": SIGMA, "nth": NTH, "nproj": NPROJ, "proj_seed": PROJ_SEED,
           "power": [float(v) for v in pw]}
    for k, f in (("p", Ap), ("s", As), ("i", Ai)):
        y = R @ f
        out[f"proj_{k}_re"] = [float(v) for v in y.real]
        out[f"proj_{k}_im"] = [float(v) for v in y.imag]
        out[f"norm_{k}"] = float(np.linalg.norm(f))
    with open(OUT, "w") as fh:
        json.dump(out, fh, indent=0)
    print(f"{OUT}: P0={pw[0]:.3e} P_end={pw[-1]:.3e}")


if __name__ == "__main__":
    main()
