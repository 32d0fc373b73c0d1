"""Closed-form pins of the oracle's two-equation nonlinear half step
(src/headers/cwes2.cuh:147-166, dAdz signs :37-38).  CPU only.

The half step is the reference's four-stage chain with unscaled stage
increments (stages on A + k1/2, A + k2/2, A + k3, Q2), stage phases at z,
z + dz/4, z + dz/4, z + dz/2 and the combine A + (k1 + 2k2 + 2k3 + k4) h/6,
h = dz/2, where
    Fp = i kp As^2 E,   Fs = i ks Ap conj(As) / E,   E = e^{i dk z}.
Two results of that definition, derived by hand, pin it here:

1. Ap = 0 (empty pump): every stage stays a multiple of As = a, and the
   chain closes exactly.  With g = kp ks |a|^2, E1, E2, E3 the phases at
   z, z + dz/4, z + dz/2, sigma = -g E1 / (2 E2), mu = 1 - (g/2)(1 + conj(sigma)/2):
       Ap' = (h/6) i kp a^2 [E1 + 2 E2 + 2 (1 + sigma/2)^2 E2 + mu^2 E3]
       As' = a + (h/6) a [2 sigma - g (1 + conj(sigma)/2) - g (1 + sigma/2)^2 conj(mu) E2 / E3].
2. Small fields Ap = eps b, As = eps a: the chain's Taylor coefficients are
       eps^2:  Ap: (h/6) i kp a^2 (E1 + 4 E2 + E3)
               As: (h/6) i ks b conj(a) (1/E1 + 4/E2 + 1/E3)
       eps^3:  Ap: -(h/3) kp ks |a|^2 b (E2/E1 + 1 + E3/E2)
               As: -(h/6) ks a (kp |a|^2 - ks |b|^2) (E1/E2 + 1 + E2/E3)
   The eps^2 terms fix the stage phases, the signs and the weights; the
   eps^3 terms exist only because the stage increments are unscaled (Q2).

A mutant chain (stage phase z+dz/4 <-> z+dz/2, stage increments scaled by h,
swapped dk signs) fails at least one of these, which the last test shows.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import c1

# arithmetic-only parameters: O(1) couplings and phases so every term counts
KP, KS, DK, DZ, Z = 0.7, 1.3, 2.9, 0.8, 0.35


def params():
    p = O.setup(c1(N=256, NZ=4, NRT=2))
    p.kp, p.ks, p.dk, p.dz = KP, KS, DK, DZ
    return p


def phases(sign=1.0):
    E = lambda z: np.exp(1j * sign * DK * z)
    return E(Z), E(Z + DZ / 4), E(Z + DZ / 2)


def exact_empty_pump(a):
    h = DZ / 2
    E1, E2, E3 = phases()
    g = KP * KS * np.abs(a) ** 2
    sig = -g * E1 / (2 * E2)
    mu = 1 - 0.5 * g * (1 + 0.5 * np.conj(sig))
    ap = h / 6 * 1j * KP * a ** 2 * (E1 + 2 * E2 + 2 * (1 + 0.5 * sig) ** 2 * E2 + mu ** 2 * E3)
    as_ = a + h / 6 * a * (2 * sig - g * (1 + 0.5 * np.conj(sig)) - g * (1 + 0.5 * sig) ** 2 * np.conj(mu) * E2 / E3)
    return ap, as_


def coefficients(b, a):
    """(C2p, C2s, C3p, C3s) of the small-field expansion."""
    h = DZ / 2
    E1, E2, E3 = phases()
    c2p = h / 6 * 1j * KP * a ** 2 * (E1 + 4 * E2 + E3)
    c2s = h / 6 * 1j * KS * b * np.conj(a) * (1 / E1 + 4 / E2 + 1 / E3)
    c3p = -h / 3 * KP * KS * np.abs(a) ** 2 * b * (E2 / E1 + 1 + E3 / E2)
    c3s = -h / 6 * KS * a * (KP * np.abs(a) ** 2 - KS * np.abs(b) ** 2) * (E1 / E2 + 1 + E2 / E3)
    return c2p, c2s, c3p, c3s


def chain(Ap, As, mutation=None):
    """The stage chain with an optional planted mistake (test of the tests)."""
    h = DZ / 2
    sgn = -1.0 if mutation == "sign" else 1.0
    zs = [Z, Z + DZ / 4, Z + DZ / 4, Z + DZ / 2]
    if mutation == "phase":
        zs = [Z, Z + DZ / 2, Z + DZ / 2, Z + DZ / 4]
    inc = h if mutation == "scaled" else 1.0

    def F(ap, as_, z):
        E = np.exp(1j * sgn * DK * z)
        return 1j * KP * as_ * as_ * E, 1j * KS * ap * np.conj(as_) / E

    k1 = F(Ap, As, zs[0])
    k2 = F(Ap + 0.5 * inc * k1[0], As + 0.5 * inc * k1[1], zs[1])
    k3 = F(Ap + 0.5 * inc * k2[0], As + 0.5 * inc * k2[1], zs[2])
    k4 = F(Ap + inc * k3[0], As + inc * k3[1], zs[3])
    return (Ap + (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0]) * h / 6,
            As + (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1]) * h / 6)


def oracle_step(Ap, As):
    Ap, As = Ap.copy(), As.copy()
    O.nl_half(params(), Ap, As, None, Z)
    return Ap, As


def samples(seed, n=256, scale=1.0):
    rng = np.random.default_rng(seed)
    return scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def check_empty_pump(step):
    a = samples(1, scale=0.6)  # g = kp ks |a|^2 up to ~3: every stage term is O(1)
    ap, as_ = step(np.zeros_like(a), a)
    ep, es = exact_empty_pump(a)
    return max(np.max(np.abs(ap - ep) / np.abs(ep)), np.max(np.abs(as_ - es) / np.abs(es)))


def check_expansion(step):
    """Worst relative error of the eps^2 and eps^3 coefficients recovered by
    finite differences in eps (the eps^4 remainder is cancelled by taking
    eps and -eps: odd orders change sign, even ones do not)."""
    b, a = samples(2), samples(3)
    c2p, c2s, c3p, c3s = coefficients(b, a)
    eps = 2e-3
    up = step(eps * b, eps * a)
    dn = step(-eps * b, -eps * a)
    # f(eps) = eps x + eps^2 C2 + eps^3 C3 + eps^4 C4 + ...
    #   even part (f(eps) + f(-eps)) / 2 = eps^2 C2 + eps^4 C4
    #   odd part  (f(eps) - f(-eps)) / 2 = eps x + eps^3 C3 + eps^5 C5
    errs = []
    for f_up, f_dn, x, C2, C3 in ((up[0], dn[0], b, c2p, c3p), (up[1], dn[1], a, c2s, c3s)):
        even = (f_up + f_dn) / 2 / eps ** 2
        odd = ((f_up - f_dn) / 2 - eps * x) / eps ** 3
        errs.append(np.max(np.abs(even - C2)) / np.max(np.abs(C2)))
        errs.append(np.max(np.abs(odd - C3)) / np.max(np.abs(C3)))
    return max(errs)


def test_empty_pump_exact(oracle_lib):
    assert check_empty_pump(oracle_step) < 1e-12


def test_small_field_expansion(oracle_lib):
    # remainder eps^2 relative (C4/C2, C5/C3 are O(1) here) plus rounding
    # ~1e-16 / eps^3 ~ 1e-8
    assert check_expansion(oracle_step) < 1e-4


def test_unmutated_chain_equals_oracle(oracle_lib):
    b, a = samples(4, scale=0.8), samples(5, scale=0.8)
    cp, cs = chain(b, a)
    op, os_ = oracle_step(b, a)
    assert np.max(np.abs(cp - op)) < 1e-13 * np.max(np.abs(op))
    assert np.max(np.abs(cs - os_)) < 1e-13 * np.max(np.abs(os_))
    assert check_empty_pump(chain) < 1e-12 and check_expansion(chain) < 1e-4


@pytest.mark.parametrize("mutation", ["phase", "scaled", "sign"])
def test_each_mutation_fails_a_closed_form(mutation):
    step = lambda Ap, As: chain(Ap, As, mutation)
    assert max(check_empty_pump(step), check_expansion(step)) > 1e-2
