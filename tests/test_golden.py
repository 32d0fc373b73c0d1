"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py with the FP64 oracle): the oracle must keep
reproducing them (CPU), and the HIP solver must match them (GPU)."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rel_l2, to_args

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    cfg = O.Config(**json.loads(str(z["config"])))
    return cfg, z


def case_inputs(cfg, z):
    p = O.setup(cfg)
    T, w = O.grids(p)
    Ap_in = O.pump_input(p)
    As = z["As0"].copy()
    Ai = z["Ai0"].copy() if "Ai0" in z.files else None
    return p, T, w, Ap_in, Ap_in[: p.N].copy(), As, Ai


def test_fixtures_present():
    assert len(FIXTURES) >= 3


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def test_oracle_reproduces_fixture(oracle_lib, path):
    cfg, z = load(path)
    p, T, w, Ap_in, Ap, As, Ai = case_inputs(cfg, z)
    _, pw = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, cfg.NRT, power=True)
    assert rel_l2(As, z["As_out"]) < 1e-13
    assert rel_l2(Ap, z["Ap_out"]) < 1e-13
    if Ai is not None:
        assert rel_l2(Ai, z["Ai_out"]) < 1e-13
    assert np.allclose(pw, z["power"], rtol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def test_gpu_matches_fixture(gpu, path):
    P = gpu
    cfg, z = load(path)
    p, T, w, Ap_in, Ap, As, Ai = case_inputs(cfg, z)
    s = P.Solver(P.params_from(P.derive(to_args(cfg)), P.FP64))
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.enable_save(True)
    s.run(0, cfg.NRT)
    gp, gs, gi = s.download()
    assert rel_l2(gs, z["As_out"]) < 1e-9
    assert rel_l2(gp, z["Ap_out"]) < 1e-9
    if Ai is not None:
        assert rel_l2(gi, z["Ai_out"]) < 1e-9
    assert np.allclose(s.power_history(cfg.NRT), z["power"], rtol=1e-9)
    saved = s.download_saved()["s"]
    slots = saved[: cfg.NRT] if cfg.regime == O.NS else saved[-cfg.NRT:]
    assert rel_l2(slots, z["save_s"]) < 1e-9


@pytest.mark.parametrize("name", ["c2", "c1"])
def test_long_record_is_well_formed(name):
    """The long-run records of full-size workloads (tests/golden/make_long.py;
    the GPU side is tests/test_gpu_long.py): one finite power per round trip,
    built up by many orders of magnitude, and one projection per row for
    every field."""
    from tests.golden import make_long as G
    ref = json.load(open(G.path(name)))
    nrt = G.WORKLOADS[name]["nrt"]
    pw = np.array(ref["power"])
    assert pw.shape == (nrt,) and np.isfinite(pw).all() and (pw > 0).all()
    assert pw[-1] > 1e6 * pw[0]
    for k in ("psi" if name == "c2" else "ps"):
        assert len(ref[f"proj_{k}_re"]) == len(ref[f"proj_{k}_im"]) == G.NPROJ
        assert ref[f"norm_{k}"] > 0
    # the projector is fixed by its seed: rows of +-1/sqrt(N), nearly orthonormal
    R = G.projector(1 << 14)
    assert np.allclose(np.abs(R), 2.0 ** -7)
    gram = R @ R.T
    assert np.allclose(np.diag(gram), 1.0) and np.max(np.abs(gram - np.eye(G.NPROJ))) < 0.05
