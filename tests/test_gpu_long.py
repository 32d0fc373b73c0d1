"""The full-size three-equation workload (C2 grid and crystal, N = 2^14, 100
slices per pass) over 1000 round trips against the FP64 oracle's record in
tests/golden/c2_long_1000rt.json (made by tests/golden/make_c2_long.py): the
power after every round trip and the final fields, through their random
projections, within 1e-7 relative; the oscillator builds up and saturates on
the way."""
import json

import numpy as np
import pytest

from tests.golden import make_c2_long as G
from tests.helpers import inputs, rel_l2, to_args

pytestmark = pytest.mark.gpu


def test_c2_grid_1000_round_trips(gpu):
    ref = json.load(open(G.OUT))
    cfg = G.config()
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=G.SEED, sigma=G.SIGMA)
    P = gpu
    s = P.Solver(P.params_from(P.derive(to_args(cfg)), 0))
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(0, G.NRT)
    gpow = s.power_history(G.NRT)
    out = s.download()
    s.close()
    opow = np.array(ref["power"])
    assert opow[-1] > 1e6 * opow[0]
    err_p = float(np.max(np.abs(gpow - opow) / opow))
    R = G.projector(p.N)
    errs = {}
    for k, g in zip("psi", out):
        yref = np.array(ref[f"proj_{k}_re"]) + 1j * np.array(ref[f"proj_{k}_im"])
        errs[k] = rel_l2(R @ g, yref)
        assert abs(np.linalg.norm(g) / ref[f"norm_{k}"] - 1) < 1e-7, k
    print(f"power max rel err {err_p:.2e}; projected field errors {errs}")
    assert err_p < 1e-7
    assert max(errs.values()) < 1e-7
