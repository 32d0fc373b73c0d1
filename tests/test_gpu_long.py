"""Full-size workloads over their long runs against the FP64 oracle's records
in tests/golden/ (made by tests/golden/make_long.py): the C2 grid and crystal
over 1000 round trips (N = 2^14, 100 slices per pass), the C1 workload over
its 2000 round trips (2 eqs, N = 2^12, 32 slices per pass), and the pulsed
(NS) three-equation workload with the EOM and GDD compensation over 1000
round trips (N = 2^12, 16 slices per pass; every pump slot read, the last one
shifted by a sample, src/headers/functions.cuh:191-213, src/cuOPO.cu:509-553).
The power after every round trip and the final fields, through their random
projections, within 1e-7 relative; the CW oscillators build up from noise and
saturate on the way.  For NS, the saved pump and signal slots of seven round
trips across the pulse as well."""
import json

import numpy as np
import pytest

from tests.golden import make_long as G
from tests.helpers import inputs, rel_l2, to_args

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["c2", "c1", "ns"])
def test_long_run_against_oracle_record(gpu, name):
    ref = json.load(open(G.path(name)))
    w = G.WORKLOADS[name]
    cfg = G.config(name)
    p, T, wg, Ap_in, Ap, As, Ai = inputs(cfg, seed=w["seed"], sigma=w["sigma"])
    P = gpu
    s = P.Solver(P.params_from(P.derive(to_args(cfg)), 0))
    s.upload(wg, T, Ap_in, Ap, As, Ai)
    if name == "ns":
        s.enable_save(True)
    s.run(0, w["nrt"])
    gpow = s.power_history(w["nrt"])
    out = s.download()
    saved = s.download_saved() if name == "ns" else None
    s.close()
    opow = np.array(ref["power"])
    if name != "ns":
        assert opow[-1] > 1e6 * opow[0]
    err_p = float(np.max(np.abs(gpow - opow) / opow))
    R = G.projector(p.N)
    errs = {}
    for k, g in zip("psi", out):
        if g is None:
            continue
        yref = np.array(ref[f"proj_{k}_re"]) + 1j * np.array(ref[f"proj_{k}_im"])
        errs[k] = rel_l2(R @ g, yref)
        assert abs(np.linalg.norm(g) / ref[f"norm_{k}"] - 1) < 1e-7, k
    if saved is not None:
        for k in ("p", "s"):
            yref = np.array(ref[f"slot_{k}_re"]) + 1j * np.array(ref[f"slot_{k}_im"])
            for j, rt in enumerate(ref["slots"]):
                errs[f"{k}@{rt}"] = rel_l2(R @ saved[k][rt], yref[j])
    print(f"{name}: power max rel err {err_p:.2e}; projected field errors {errs}")
    assert err_p < 1e-7
    assert max(errs.values()) < 1e-7
