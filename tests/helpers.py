"""Shared test helpers: workload definitions and error metrics."""
import numpy as np

from oracle import oracle as O


def rel_l2(a, b):
    a = np.asarray(a); b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def oracle_config(**kw):
    return O.Config(**kw)


def c1(N=4096, NZ=32, NRT=2000):
    """2 eqs, CW, PPLN, 532 -> 1064 nm, 5 mm, 2 W pump (Nth = 2 W / Pth)."""
    cfg = O.Config(three_eqs=False, regime=O.CW, crystal=O.PPLN, N=N, NZ=NZ, NRT=NRT, lp_nm=532, Lambda=6.97,
                   Nth=1.0, R_pct=98, lcr_mm=5)
    p = O.setup(cfg)
    cfg.Nth = 2.0 / p.Pth
    return cfg


def c2(N=16384, NZ=100, NRT=5000):
    """3 eqs, CW, sPPLT, 532 nm pump, 1000 nm signal, 20 mm, 3x threshold,
    0.3 rad detuning (argv6 is in units of pi)."""
    return O.Config(three_eqs=True, regime=O.CW, crystal=O.SPPLT, N=N, NZ=NZ, NRT=NRT, lp_nm=532,
                    Lambda=C2_LAMBDA, Nth=3.0, R_pct=98, delta=0.3 / np.pi, lcr_mm=20, ls_um=1.0,
                    pm_mode=O.PM_THREE_WAVE)


# the sweep script's sPPLT grating period (src/cuOPO.sh:75); with the three-wave scan it
# phase-matches 532 -> 1000 + 1137 nm near 70 C
C2_LAMBDA = 7.97


def c5(N=1 << 18, NZ=200, NRT=1000):
    """3 eqs, NS, PPLN, 4x threshold, EOM depth 0.8 rad (argv10 in units of pi), GDD 100%."""
    return O.Config(three_eqs=True, regime=O.NS, crystal=O.PPLN, N=N, NZ=NZ, NRT=NRT, lp_nm=532, Lambda=6.97,
                    Nth=4.0, R_pct=98, gdd_pct=100, upm=1, md=0.8 / np.pi, fm=4, lcr_mm=5)


def inputs(cfg, seed=1, sigma=1e-20):
    p = O.setup(cfg)
    T, w = O.grids(p)
    Ap_in = O.pump_input(p)
    As = O.noise(seed, 0, sigma, p.N)
    Ai = O.noise(seed, 1, sigma, p.N) if cfg.three_eqs else None
    Ap = Ap_in[: p.N].copy()
    return p, T, w, Ap_in, Ap, As, Ai


def to_args(cfg):
    import cuopo_amd as P
    return P.Args(**{k: getattr(cfg, k) for k in cfg.__dataclass_fields__})


def random_fields(N, three, seed=0, amp_p=0.3, amp_s=0.1):
    rng = np.random.default_rng(seed)
    f = lambda a: (rng.standard_normal(N) + 1j * rng.standard_normal(N)) * a
    return f(amp_p), f(amp_s), (f(amp_s) if three else None)
