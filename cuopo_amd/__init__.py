"""cuopo_amd -- MI355X (gfx950) split-step Fourier round-trip solver for the
three-wave-mixing coupled-wave equations, a drop-in for the cuOPO solver.

Python mirror of the C ABI in ``include/opo.h`` (ctypes over
``libcuopo_amd.so``).  The library is the product: every propagation call
runs the hand-written HIP kernels; there is no CPU fallback, and importing a
solver without the built library raises.

Reference interface mirrored (paths relative to the cuOPO source tree):
``Args`` holds the 12 positional arguments of ``src/cuOPO.cu:92-267`` plus the
``-D`` switches; ``derive`` is the setup block ``src/cuOPO.cu:92-269``;
``Solver.crystal_pass`` is ``EvolutionInCrystal`` (``src/headers/cwes3.cuh:139``);
``Solver.run`` is the round-trip loop ``src/cuOPO.cu:474-610``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcuopo_amd.so")

OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_STATE, ERR_DIVERGED = 0, -1, -2, -3, -4, -5
PPLN, SPPLT = 0, 1
CW, NS = 0, 1
FP64, FP32C = 0, 1
PM_REFERENCE, PM_THREE_WAVE = 0, 1


class OpoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"opo status {status}: {msg}")
        self.status = status


class _Args(C.Structure):
    _fields_ = [
        ("three_eqs", C.c_int), ("regime", C.c_int), ("crystal", C.c_int),
        ("N", C.c_int), ("NZ", C.c_int), ("NRT", C.c_int),
        ("lp_nm", C.c_double), ("temp_arg", C.c_double), ("Lambda", C.c_double),
        ("Nth", C.c_double), ("R_pct", C.c_double), ("delta", C.c_double),
        ("gdd_pct", C.c_double), ("tod_pct", C.c_double), ("upm", C.c_int),
        ("md", C.c_double), ("fm", C.c_double), ("lcr_mm", C.c_double),
        ("ls_um", C.c_double), ("pm_mode", C.c_int),
    ]


class Derived(C.Structure):
    _fields_ = [
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

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Params(C.Structure):
    _fields_ = [
        ("three_eqs", C.c_int), ("regime", C.c_int), ("N", C.c_int), ("NZ", C.c_int), ("NRT", C.c_int),
        ("precision", C.c_int),
    ] + [(n, C.c_double) for n in (
        "dz dk kp ks ki vp vs vi b2p b2s b2i b3p b3s b3i alpha_crp alpha_crs alpha_cri "
        "Rs Ri delta GDD TODscomp TODicomp").split()] + [
        ("use_eom", C.c_int), ("mod_depth", C.c_double), ("fpm", C.c_double),
        ("Nrts", C.c_int), ("device", C.c_int),
    ]

    def copy(self):
        p = Params()
        C.pointer(p)[0] = self
        return p


# Every symbol include/opo.h declares, with its ctypes signature.
_d, _i, _sz, _p, _u64 = C.c_double, C.c_int, C.c_size_t, C.c_void_p, C.c_uint64
SIGNATURES = {
    "opo_derive": (_i, [C.POINTER(_Args), C.POINTER(Derived)]),
    "opo_params_from_derived": (None, [C.POINTER(Derived), _i, C.POINTER(Params)]),
    "opo_grids": (None, [C.POINTER(Derived), _p, _p]),
    "opo_grids_long": (None, [C.POINTER(Derived), _p, _p]),
    "opo_pump_input": (None, [C.POINTER(Derived), _p]),
    "opo_noise": (None, [_u64, _u64, _d, _p, _sz]),
    "opo_find_pm_temp": (_d, [_i, _d, _d, _d, _d, _d, _i]),
    "opo_n": (_d, [_i, _d, _d]),
    "opo_group_velocity": (_d, [_i, _d, _d]),
    "opo_gvd": (_d, [_i, _d, _d]),
    "opo_create": (_i, [C.POINTER(Params), C.POINTER(_p)]),
    "opo_upload": (_i, [_p, _p, _p, _p, _sz, _p, _p, _p]),
    "opo_set_fields": (_i, [_p, _p, _p, _p]),
    "opo_crystal_pass": (_i, [_p]),
    "opo_linear_step": (_i, [_p]),
    "opo_nl_half": (_i, [_p, _i, _i]),
    "opo_boundary": (_i, [_p]),
    "opo_run": (_i, [_p, _i, _i]),
    "opo_run_noalloc": (_i, [C.POINTER(Params), _p, _p, _p, C.c_size_t, _p, _i, _i, _p, _p]),
    "opo_enable_save": (_i, [_p, _i]),
    "opo_download": (_i, [_p, _p, _p, _p]),
    "opo_download_saved": (_i, [_p, _p, _p, _p]),
    "opo_power_history": (_i, [_p, _p, _i]),
    "opo_set_tiles": (_i, [_p, _i, _i]),
    "opo_use_graph": (_i, [_p, _i]),
    "opo_linear_factors": (_i, [_p, _p, _i, _i, _i, _p]),
    "opo_set_xcd_pack": (_i, [_p, _i]),
    "opo_autotune": (_i, [_p, _i]),
    "opo_get_tuning": (_i, [_p, _p]),
    "opo_set_tuning": (_i, [_p, _p]),
    "opo_batch_autotune": (_i, [_p, _i]),
    "opo_batch_enable_save": (_i, [_p, _i]),
    "opo_batch_download_saved": (_i, [_p, _i, _p, _p, _p]),
    "opo_batch_get_tuning": (_i, [_p, _p, _i]),
    "opo_batch_set_tuning": (_i, [_p, _p]),
    "opo_batch_time_stages": (_i, [_p, _i, _i, C.POINTER(_d), C.POINTER(_d)]),
    "opo_last_run_ms": (_i, [_p, C.POINTER(_d)]),
    "opo_time_stages": (_i, [_p, _i, C.POINTER(_d), C.POINTER(_d)]),
    "opo_copy_probe": (_i, [_sz, _i, _i, C.POINTER(_d), C.POINTER(_d)]),
    "opo_destroy": (_i, [_p]),
    "opo_batch_create": (_i, [C.POINTER(Params), _i, C.POINTER(_p)]),
    "opo_batch_upload": (_i, [_p, _i, _p, _p, _p, _sz, _p, _p, _p]),
    "opo_batch_run": (_i, [_p, _i, _i, _d, _p]),
    "opo_batch_run_ex": (_i, [_p, _i, _i, _d, _d, _p, _p]),
    "opo_batch_download": (_i, [_p, _i, _p, _p, _p]),
    "opo_batch_power": (_i, [_p, _p]),
    "opo_batch_power_history": (_i, [_p, _i, _p, _i]),
    "opo_batch_last_run_ms": (_i, [_p, C.POINTER(_d)]),
    "opo_batch_set_accel": (_i, [_p, _i, _i]),
    "opo_batch_set_tiles": (_i, [_p, _i, _i]),
    "opo_batch_destroy": (_i, [_p]),
    "opo_mlp_create": (_i, [_i, _i, _i, _i, _i, _u64, _i, C.POINTER(_p)]),
    "opo_mlp_param_count": (_i, [_p, C.POINTER(_sz)]),
    "opo_mlp_set_weights": (_i, [_p, _p, _sz]),
    "opo_mlp_get_weights": (_i, [_p, _p, _sz]),
    "opo_mlp_forward": (_i, [_p, _p, _i, _p]),
    "opo_mlp_gradients": (_i, [_p, _p, _p, _i, _p, _p]),
    "opo_mlp_train_step": (_i, [_p, _p, _p, _i, C.c_float, _p]),
    "opo_mlp_destroy": (_i, [_p]),
    "opo_strerror": (C.c_char_p, [_i]),
    "opo_last_error": (C.c_char_p, []),
    "opo_device_count": (_i, [C.POINTER(_i)]),
}

_lib = None


class Tuning(C.Structure):
    """opo_tuning (include/opo.h)."""
    _fields_ = [(n, C.c_int) for n in ("N", "count", "col_threads", "row_threads", "xcd_pack", "phase_mode",
                                       "col_spt", "row_spt", "col_tile", "row_tile", "factored")] + \
        [("slice_us", C.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def lib():
    """Load libcuopo_amd.so.  Raises if it has not been built: there is no
    fallback path."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {HERE}/csrc`; "
                               "cuopo_amd has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != OK:
        msg = lib().opo_last_error().decode(errors="replace")
        raise OpoError(rc, msg or lib().opo_strerror(rc).decode())


def _ptr(a):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)


def _f64(a, n):
    """A real grid (w or T) of exactly n values as contiguous float64."""
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != n:
        raise ValueError(f"expected a grid of {n} values, got {a.size}")
    return a


def _c128(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if n is not None and a.size < n:
        raise ValueError(f"expected at least {n} samples, got {a.size}")
    return a


@dataclass
class Args:
    """Positional arguments and switches of the cuOPO executable."""
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

    def _c(self):
        a = _Args()
        for f in fields(self):
            v = getattr(self, f.name)
            setattr(a, f.name, int(v) if isinstance(getattr(a, f.name), int) else float(v))
        return a


def derive(args: Args) -> Derived:
    d = Derived()
    _check(lib().opo_derive(C.byref(args._c()), C.byref(d)))
    return d


def params_from(d: Derived, precision=FP64, device=0) -> Params:
    p = Params()
    lib().opo_params_from_derived(C.byref(d), precision, C.byref(p))
    p.device = device
    return p


def grids(d: Derived):
    T = np.empty(d.N); w = np.empty(d.N)
    lib().opo_grids(C.byref(d), _ptr(T), _ptr(w))
    return T, w


def grids_long(d: Derived):
    Tp = np.empty(d.SIZEL); Fp = np.empty(d.SIZEL)
    lib().opo_grids_long(C.byref(d), _ptr(Tp), _ptr(Fp))
    return Tp, Fp


def pump_input(d: Derived):
    n = d.N if d.regime == CW else d.SIZEL
    A = np.empty(n, np.complex128)
    lib().opo_pump_input(C.byref(d), _ptr(A))
    return A


def noise(seed: int, stream: int, sigma: float, n: int):
    A = np.empty(n, np.complex128)
    lib().opo_noise(seed, stream, sigma, _ptr(A), n)
    return A


def find_pm_temp(crystal, lp, ls, Lambda, Tmin=10.0, Tmax=150.0, pm_mode=PM_REFERENCE):
    return lib().opo_find_pm_temp(crystal, lp, ls, Lambda, Tmin, Tmax, pm_mode)


def device_count():
    n = C.c_int(0)
    _check(lib().opo_device_count(C.byref(n)))
    return n.value


class Solver:
    """One OPO instance on one GPU (opo_solver)."""

    def __init__(self, params: Params):
        self.params = params.copy()
        self.N = params.N
        self.three = bool(params.three_eqs)
        h = C.c_void_p()
        _check(lib().opo_create(C.byref(self.params), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().opo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, w, T, Ap_in, Ap, As, Ai=None):
        w, T = _f64(w, self.N), _f64(T, self.N)
        Ap_in = _c128(Ap_in)
        Ap, As, Ai = _c128(Ap, self.N), _c128(As, self.N), _c128(Ai, self.N)
        _check(lib().opo_upload(self._h, _ptr(w), _ptr(T), _ptr(Ap_in), Ap_in.size, _ptr(Ap), _ptr(As), _ptr(Ai)))

    def set_fields(self, Ap, As, Ai=None):
        Ap, As, Ai = _c128(Ap, self.N), _c128(As, self.N), _c128(Ai, self.N)
        _check(lib().opo_set_fields(self._h, _ptr(Ap), _ptr(As), _ptr(Ai)))

    def crystal_pass(self):
        _check(lib().opo_crystal_pass(self._h))

    def linear_step(self):
        _check(lib().opo_linear_step(self._h))

    def nl_half(self, slice_index, n_times=1):
        _check(lib().opo_nl_half(self._h, slice_index, n_times))

    def boundary(self):
        _check(lib().opo_boundary(self._h))

    def enable_save(self, on=True):
        _check(lib().opo_enable_save(self._h, 1 if on else 0))

    def run(self, rt0, n_rt):
        _check(lib().opo_run(self._h, rt0, n_rt))

    def download(self):
        Ap = np.empty(self.N, np.complex128); As = np.empty(self.N, np.complex128)
        Ai = np.empty(self.N, np.complex128) if self.three else None
        _check(lib().opo_download(self._h, _ptr(Ap), _ptr(As), _ptr(Ai)))
        return Ap, As, Ai

    def download_saved(self):
        slots = self.params.Nrts if self.params.regime == CW else self.params.NRT
        out = {k: np.empty((slots, self.N), np.complex128) for k in ("p", "s")}
        if self.three:
            out["i"] = np.empty((slots, self.N), np.complex128)
        _check(lib().opo_download_saved(self._h, _ptr(out["p"]), _ptr(out["s"]), _ptr(out.get("i"))))
        return out

    def power_history(self, n):
        out = np.zeros(n)
        _check(lib().opo_power_history(self._h, _ptr(out), n))
        return out

    def set_tiles(self, col_threads=0, row_threads=0):
        _check(lib().opo_set_tiles(self._h, col_threads, row_threads))

    def use_graph(self, on=True):
        _check(lib().opo_use_graph(self._h, 1 if on else 0))

    def set_xcd_pack(self, on=True):
        _check(lib().opo_set_xcd_pack(self._h, 1 if on else 0))

    def autotune(self, iters=10):
        """Pick tiles / XCD packing / phase source by timing them on the
        uploaded instance (fields are restored afterwards)."""
        _check(lib().opo_autotune(self._h, iters))
        return self.get_tuning()

    def get_tuning(self):
        t = Tuning()
        _check(lib().opo_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, col_threads=0, row_threads=0, xcd_pack=-1, phase_mode=-1, col_spt=0, row_spt=0):
        t = Tuning(col_threads=col_threads, row_threads=row_threads, xcd_pack=xcd_pack, phase_mode=phase_mode,
                   col_spt=col_spt, row_spt=row_spt)
        _check(lib().opo_set_tuning(self._h, C.byref(t)))

    def last_run_ms(self):
        v = C.c_double()
        _check(lib().opo_last_run_ms(self._h, C.byref(v)))
        return v.value

    def time_stages(self, iters=50):
        """(column-stage us, row-stage us) per launch, each kernel timed alone."""
        a, b = C.c_double(), C.c_double()
        _check(lib().opo_time_stages(self._h, iters, C.byref(a), C.byref(b)))
        return a.value, b.value


def copy_probe(nbytes, iters=20, device=0):
    """Measured streaming-copy bandwidth (GB/s, read+write) and kernel time (us)."""
    g, u = C.c_double(), C.c_double()
    _check(lib().opo_copy_probe(nbytes, iters, device, C.byref(g), C.byref(u)))
    return g.value, u.value


class Batch:
    """Many independent instances (possibly different N) iterated together."""

    def __init__(self, params_list):
        n = len(params_list)
        arr = (Params * n)()
        for k, p in enumerate(params_list):
            arr[k] = p
        self.params = [p.copy() for p in params_list]
        self.count = n
        h = C.c_void_p()
        _check(lib().opo_batch_create(arr, n, C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().opo_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, idx, w, T, Ap_in, Ap, As, Ai=None):
        N = self.params[idx].N
        w, T = _f64(w, N), _f64(T, N)
        Ap_in = _c128(Ap_in)
        Ap, As, Ai = _c128(Ap, N), _c128(As, N), _c128(Ai, N)
        _check(lib().opo_batch_upload(self._h, idx, _ptr(w), _ptr(T), _ptr(Ap_in), Ap_in.size,
                                      _ptr(Ap), _ptr(As), _ptr(Ai)))

    def run(self, max_rt, check_every=50, tol=1e-6):
        rts = np.zeros(self.count, np.int32)
        _check(lib().opo_batch_run(self._h, max_rt, check_every, tol, _ptr(rts)))
        return rts

    def run_ex(self, max_rt, check_every=50, tol=1e-6, off_ratio=1e-3):
        """Returns (round trips run, state) per instance; state 1 steady,
        2 off (decayed below threshold), 0 not converged, -1 diverged."""
        rts = np.zeros(self.count, np.int32)
        st = np.zeros(self.count, np.int32)
        rc = lib().opo_batch_run_ex(self._h, max_rt, check_every, tol, off_ratio, _ptr(rts), _ptr(st))
        if rc not in (OK, ERR_DIVERGED):
            _check(rc)
        return rts, st

    def download(self, idx):
        N = self.params[idx].N
        three = bool(self.params[idx].three_eqs)
        Ap = np.empty(N, np.complex128); As = np.empty(N, np.complex128)
        Ai = np.empty(N, np.complex128) if three else None
        _check(lib().opo_batch_download(self._h, idx, _ptr(Ap), _ptr(As), _ptr(Ai)))
        return Ap, As, Ai

    def power(self):
        out = np.empty(self.count)
        _check(lib().opo_batch_power(self._h, _ptr(out)))
        return out

    def power_history(self, idx, n):
        out = np.zeros(n)
        _check(lib().opo_batch_power_history(self._h, idx, _ptr(out), n))
        return out

    def set_tiles(self, col_threads=0, row_threads=0):
        _check(lib().opo_batch_set_tiles(self._h, col_threads, row_threads))

    def set_accel(self, every=400, start=1000):
        """Round-trip extrapolation of slowly converging instances in the
        convergence-tested runs (every = 0 turns it off)."""
        _check(lib().opo_batch_set_accel(self._h, every, start))

    def last_run_ms(self):
        v = C.c_double()
        _check(lib().opo_batch_last_run_ms(self._h, C.byref(v)))
        return v.value

    def autotune(self, iters=10):
        """Pick tiles / XCD packing / phase source per grid size by timing
        them on the uploaded instances (fields are restored afterwards)."""
        _check(lib().opo_batch_autotune(self._h, iters))
        return self.get_tuning()

    def get_tuning(self):
        n = lib().opo_batch_get_tuning(self._h, None, 0)
        if n < 0:
            _check(n)
        arr = (Tuning * max(n, 1))()
        lib().opo_batch_get_tuning(self._h, C.cast(arr, C.c_void_p), n)
        return [arr[k].as_dict() for k in range(n)]

    def set_tuning(self, N, col_threads=0, row_threads=0, xcd_pack=-1, phase_mode=-1, col_spt=0, row_spt=0):
        t = Tuning(N=N, col_threads=col_threads, row_threads=row_threads, xcd_pack=xcd_pack,
                   phase_mode=phase_mode, col_spt=col_spt, row_spt=row_spt)
        _check(lib().opo_batch_set_tuning(self._h, C.byref(t)))

    def time_stages(self, N, iters=20):
        a, b = C.c_double(), C.c_double()
        _check(lib().opo_batch_time_stages(self._h, N, iters, C.byref(a), C.byref(b)))
        return a.value, b.value


def setup_instance(args: Args, precision=FP64, seed=1, noise_sigma=1e-20, device=0):
    """Everything a run needs for one argv point: derived scalars, solver
    params, grids, pump and seeded noise (signal stream 0, idler stream 1)."""
    d = derive(args)
    p = params_from(d, precision, device)
    T, w = grids(d)
    Ap_in = pump_input(d)
    As = noise(seed, 0, noise_sigma, d.N)
    Ai = noise(seed, 1, noise_sigma, d.N) if args.three_eqs else None
    Ap = Ap_in[: d.N].copy()
    return dict(derived=d, params=p, T=T, w=w, Ap_in=Ap_in, Ap=Ap, As=As, Ai=Ai)
