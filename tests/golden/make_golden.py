"""Writes the golden fixtures tests/golden/*.npz with the FP64 oracle.

Each fixture holds the reference argv/switches, the seeded initial fields and
the fields after NRT round trips (plus the power history).  Run from the
repository root after `make -C oracle`:  python tests/golden/make_golden.py [name ...]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from tests.helpers import c1, c2, c5, inputs  # noqa: E402

def ns_spplt_2eq(N=1024, NZ=8, NRT=5):
    """NS, sPPLT, two equations (degenerate signal, the reference's phase-matching
    scan), EOM and GDD: the -DNS_OPO -DSPPLT build without -DTHREE_EQS."""
    import math
    return O.Config(three_eqs=False, regime=O.NS, crystal=O.SPPLT, N=N, NZ=NZ, NRT=NRT, lp_nm=532, Lambda=7.97,
                    Nth=4.0, R_pct=98, gdd_pct=100, upm=1, md=0.8 / math.pi, fm=4, lcr_mm=5)


CASES = {
    "cw_ppln_2eq": (lambda: c1(N=1024, NZ=8, NRT=5), 11, 1e-3),
    "cw_spplt_3eq": (lambda: c2(N=1024, NZ=8, NRT=5), 12, 1e-3),
    "ns_ppln_3eq_eom_gdd": (lambda: c5(N=1024, NZ=8, NRT=5), 13, 1e-3),
    "ns_spplt_2eq_eom_gdd": (ns_spplt_2eq, 14, 1e-3),
}


def make(name):
    make_cfg, seed, sigma = CASES[name]
    cfg = make_cfg()
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=seed, sigma=sigma)
    As0, Ai0 = As.copy(), (None if Ai is None else Ai.copy())
    sv, pw = O.run(p, w, T, Ap_in, Ap, As, Ai, 0, cfg.NRT, save=True, power=True)
    out = dict(config=json.dumps({k: getattr(cfg, k) for k in cfg.__dataclass_fields__}),
               As0=As0, Ap_out=Ap, As_out=As, power=pw, save_s=sv["s"][: cfg.NRT] if cfg.regime == O.NS else sv["s"][-cfg.NRT:])
    if Ai is not None:
        out.update(Ai0=Ai0, Ai_out=Ai)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


if __name__ == "__main__":
    # names on the command line (default: every case); an existing fixture is
    # rewritten only when named, since the compressed file is not byte-stable
    for n in sys.argv[1:] or CASES:
        make(n)
        print("wrote", n)
