"""ctypes bindings for the FP64 oracle (``oracle/libopo_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, the package smoke check and
the CPU-baseline leg of ``bench.py``, to compare against.  The product path
(``cuopo_amd``) never imports it.

Fields are numpy ``complex128`` arrays (interleaved re/im, the layout of
``std::complex<double>``), passed by pointer and modified in place where the C
function does so.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopo_oracle.so")

PPLN, SPPLT = 0, 1
CW, NS = 0, 1
PM_REFERENCE, PM_THREE_WAVE = 0, 1


class Inputs(C.Structure):
    _fields_ = [
        ("three_eqs", C.c_int), ("regime", C.c_int), ("crystal", C.c_int),
        ("N", C.c_int), ("NZ", C.c_int), ("NRT", C.c_int),
        ("lp_nm", C.c_double), ("temp_arg", C.c_double), ("Lambda", C.c_double),
        ("Nth", C.c_double), ("R_pct", C.c_double), ("delta", C.c_double),
        ("gdd_pct", C.c_double), ("tod_pct", C.c_double), ("upm", C.c_int),
        ("md", C.c_double), ("fm", C.c_double), ("lcr_mm", C.c_double),
        ("ls_um", C.c_double), ("pm_mode", C.c_int),
    ]


_PARAM_FIELDS = [
    ("three_eqs", C.c_int), ("regime", C.c_int), ("crystal", C.c_int),
    ("N", C.c_int), ("NZ", C.c_int), ("NRT", C.c_int),
] + [(n, C.c_double) for n in (
    "lp ls li Lambda Temp Lcr np ns ni vp vs vi b2p b2s b2i b3p b3s b3i kp ks ki "
    "dk dkp Lcav Rs Ri alphas alphai t_rt FSR finesse lw delta epsilon GDD TODscomp TODicomp dz dT dF "
    "deff dQ alpha_crp alpha_crs alpha_cri waist spot Ith Pth Nth Inten Power Ap0 GammaL").split()] + [
    ("use_eom", C.c_int), ("mod_depth", C.c_double), ("df", C.c_double), ("fpm", C.c_double),
    ("Nrts", C.c_int), ("SIZEL", C.c_size_t), ("T_width", C.c_double), ("dFp", C.c_double),
    ("FWHM", C.c_double), ("sigmap", C.c_double),
]


class Params(C.Structure):
    _fields_ = _PARAM_FIELDS

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: build it with `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        d, i, sz, p = C.c_double, C.c_int, C.c_size_t, C.c_void_p
        sigs = {
            "orc_n": (d, [i, d, d]), "orc_resonances": (d, [i, d, d, i]),
            "orc_dndl": (d, [i, d, d]), "orc_d2ndl2": (d, [i, d, d]), "orc_d3ndl3": (d, [i, d, d]),
            "orc_GV": (d, [i, d, d]), "orc_GVD": (d, [i, d, d]), "orc_TOD": (d, [i, d, d]),
            "orc_find_pm_temp": (d, [i, d, d, d, d, d, i]),
            "orc_setup": (i, [C.POINTER(Inputs), C.POINTER(Params)]),
            "orc_linspace": (None, [p, sz, d, d]), "orc_fftshift": (None, [p, p, sz]),
            "orc_grids": (None, [C.POINTER(Params), p, p]),
            "orc_grids_long": (None, [C.POINTER(Params), p, p]),
            "orc_pump_input": (None, [C.POINTER(Params), p]),
            "orc_noise": (None, [C.c_uint64, C.c_uint64, d, p, sz]),
            "orc_fft": (None, [p, sz, i]),
            "orc_nl_half": (None, [C.POINTER(Params), p, p, p, d]),
            "orc_linear": (None, [C.POINTER(Params), p, p, p, p]),
            "orc_crystal_pass": (None, [C.POINTER(Params), p, p, p, p]),
            "orc_eom": (None, [C.POINTER(Params), p, p]),
            "orc_gdd": (None, [C.POINTER(Params), p, p]),
            "orc_mirror": (None, [C.POINTER(Params), d, p]),
            "orc_read_pump": (None, [C.POINTER(Params), p, i, p]),
            "orc_run": (None, [C.POINTER(Params), p, p, p, p, p, p, i, i, p, p, p, p]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be contiguous"
    return a.ctypes.data_as(C.c_void_p)


@dataclass
class Config:
    """The reference's argv and switches (src/cuOPO.cu:92-267, cuOPO.sh:69-88)."""
    three_eqs: bool = False
    regime: int = CW
    crystal: int = PPLN
    N: int = 1 << 14
    NZ: int = 150
    NRT: int = 10000
    lp_nm: float = 532.0
    temp_arg: float = 0.0
    Lambda: float = 6.97
    Nth: float = 2.0
    R_pct: float = 98.0
    delta: float = 0.0
    gdd_pct: float = 0.0
    tod_pct: float = 0.0
    upm: int = 0
    md: float = 0.0
    fm: float = 0.0
    lcr_mm: float = 5.0
    ls_um: float = 0.0
    pm_mode: int = PM_REFERENCE

    def to_inputs(self) -> Inputs:
        x = Inputs()
        for k, _ in Inputs._fields_:
            setattr(x, k, int(getattr(self, k)) if k in ("three_eqs", "regime", "crystal", "N", "NZ", "NRT",
                                                          "upm", "pm_mode") else float(getattr(self, k)))
        return x


def setup(cfg: Config) -> Params:
    p = Params()
    rc = lib().orc_setup(C.byref(cfg.to_inputs()), C.byref(p))
    if rc != 0:
        raise ValueError("oracle setup rejected the inputs (sqrt(Nth-1) with the EOM on needs Nth >= 1)")
    return p


def grids(p: Params):
    T = np.empty(p.N); w = np.empty(p.N)
    lib().orc_grids(C.byref(p), _ptr(T), _ptr(w))
    return T, w


def grids_long(p: Params):
    Tp = np.empty(p.SIZEL); Fp = np.empty(p.SIZEL)
    lib().orc_grids_long(C.byref(p), _ptr(Tp), _ptr(Fp))
    return Tp, Fp


def pump_input(p: Params):
    n = p.N if p.regime == CW else p.SIZEL
    A = np.empty(n, np.complex128)
    lib().orc_pump_input(C.byref(p), _ptr(A))
    return A


def noise(seed: int, stream: int, sigma: float, n: int):
    A = np.empty(n, np.complex128)
    lib().orc_noise(seed, stream, sigma, _ptr(A), n)
    return A


def fft(x, sign: int):
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft(_ptr(y), y.size, sign)
    return y


def nl_half(p, Ap, As, Ai, z):
    lib().orc_nl_half(C.byref(p), _ptr(Ap), _ptr(As), _ptr(Ai), z)


def linear(p, w, Ap, As, Ai):
    lib().orc_linear(C.byref(p), _ptr(w), _ptr(Ap), _ptr(As), _ptr(Ai))


def crystal_pass(p, w, Ap, As, Ai):
    lib().orc_crystal_pass(C.byref(p), _ptr(w), _ptr(Ap), _ptr(As), _ptr(Ai))


def eom(p, T, A):
    lib().orc_eom(C.byref(p), _ptr(T), _ptr(A))


def gdd(p, w, A):
    lib().orc_gdd(C.byref(p), _ptr(w), _ptr(A))


def mirror(p, R, A):
    lib().orc_mirror(C.byref(p), R, _ptr(A))


def read_pump(p, Ap_in, nn):
    Ap = np.empty(p.N, np.complex128)
    lib().orc_read_pump(C.byref(p), _ptr(Ap_in), nn, _ptr(Ap))
    return Ap


def run(p, w, T, Ap_in, Ap, As, Ai, rt0, nrt, save=False, power=False):
    """Round trips rt0..rt0+nrt-1 in place.  Returns (saves, power) where saves is
    a dict of (slots, N) arrays when ``save`` is set."""
    slots = p.Nrts if p.regime == CW else p.NRT
    sv = {k: np.zeros((slots, p.N), np.complex128) for k in ("p", "s", "i")} if save else None
    pw = np.zeros(nrt) if power else None
    lib().orc_run(C.byref(p), _ptr(w), _ptr(T), _ptr(Ap_in), _ptr(Ap), _ptr(As),
                  _ptr(Ai) if Ai is not None else None, rt0, nrt,
                  _ptr(sv["p"]) if sv else None, _ptr(sv["s"]) if sv else None,
                  _ptr(sv["i"]) if sv else None, _ptr(pw))
    return sv, pw
