"""opo_run_noalloc (the entry point that allocates nothing: static per-GPU
scratch, caller-owned device fields and pump) against opo_run of a solver
with the same parameters, through the C ABI: two calls in a row, the second
on a smaller grid, then a request larger than the scratch holds.  GPU test;
the device buffers come from the HIP runtime the library itself links."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import c1, c5, inputs, rel_l2, to_args

pytestmark = pytest.mark.gpu


def hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy):
        f.restype = C.c_int
    return h


H2D, D2H = 1, 2


class Dev:
    """A device buffer holding a copy of a host complex128 array."""

    def __init__(self, h, host):
        self.h, self.n = h, host.nbytes
        self.p = C.c_void_p()
        assert h.hipMalloc(C.byref(self.p), self.n) == 0
        assert h.hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), self.n, H2D) == 0

    def get(self, n):
        out = np.empty(n, np.complex128)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def compare(P, cfg, seed, rt0, n_rt):
    """Run rt0 .. rt0+n_rt-1 through opo_run (a fresh solver, fields uploaded
    as of round trip rt0) and through opo_run_noalloc; return the largest
    relative difference of the fields and of the power history."""
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=seed, sigma=1e-3)
    prm = P.params_from(P.derive(to_args(cfg)))
    s = P.Solver(prm)
    s.upload(w, T, Ap_in, Ap, As, Ai)
    s.run(rt0, n_rt)
    want = s.download()
    want_pw = s.power_history(n_rt)
    s.close()
    h = hip()
    fields = np.concatenate([Ap, As] + ([Ai] if cfg.three_eqs else []))
    df, dp = Dev(h, fields), Dev(h, np.ascontiguousarray(Ap_in))
    pw = np.zeros(n_rt)
    try:
        rc = P.lib().opo_run_noalloc(C.byref(prm), w.ctypes.data_as(C.c_void_p), T.ctypes.data_as(C.c_void_p),
                                     dp.p, Ap_in.size, df.p, rt0, n_rt, pw.ctypes.data_as(C.c_void_p), None)
        P._check(rc)
        got = df.get(fields.size)
    finally:
        df.free()
        dp.free()
    N = cfg.N
    errs = [rel_l2(got[k * N:(k + 1) * N], want[k]) for k in range(3 if cfg.three_eqs else 2)]
    errs.append(float(np.max(np.abs(pw - want_pw) / want_pw)))
    return errs, all(np.array_equal(got[k * N:(k + 1) * N], want[k]) for k in range(len(errs) - 1))


def test_noalloc_equals_opo_run_twice_then_rejects_oversize(gpu):
    P = gpu
    # first call: pulsed, three equations, EOM and GDD tables, N = 4096
    big = c5(N=4096, NZ=8, NRT=6)
    e1, exact1 = compare(P, big, 31, 0, 6)
    print("noalloc vs opo_run, NS N=4096:", e1, "bit-exact" if exact1 else "")
    assert max(e1) < 1e-13
    # second call, smaller grid: CW two equations, N = 1024, round trips 3..7
    small = c1(N=1024, NZ=4, NRT=10)
    e2, exact2 = compare(P, small, 32, 3, 5)
    print("noalloc vs opo_run, CW N=1024:", e2, "bit-exact" if exact2 else "")
    assert max(e2) < 1e-13
    # a grid whose tables exceed the scratch: rejected before any round trip runs
    huge = c5(N=1 << 18, NZ=4, NRT=2)
    prm = P.params_from(P.derive(to_args(huge)))
    p, T, w, Ap_in, Ap, As, Ai = inputs(huge, seed=1, sigma=1e-3)
    h = hip()
    df, dp = Dev(h, np.concatenate([Ap, As, Ai])), Dev(h, np.ascontiguousarray(Ap_in))
    try:
        rc = P.lib().opo_run_noalloc(C.byref(prm), w.ctypes.data_as(C.c_void_p), T.ctypes.data_as(C.c_void_p),
                                     dp.p, Ap_in.size, df.p, 0, 1, None, None)
        assert rc == P.ERR_INVALID
        assert "scratch" in P.lib().opo_last_error().decode()
        # the rejected call left the caller's fields as they were
        assert np.array_equal(df.get(3 * huge.N), np.concatenate([Ap, As, Ai]))
    finally:
        df.free()
        dp.free()
