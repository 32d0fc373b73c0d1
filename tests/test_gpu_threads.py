"""Distinct handles driven from different host threads at the same time (GPU,
C ABI).  Each handle owns its stream and device memory and the library keeps
no other mutable state, so concurrent runs must equal the same runs made one
after the other, bit for bit, and each thread must see its own last error."""
import threading

import numpy as np
import pytest

from tests.helpers import c1, c2, inputs, to_args

pytestmark = pytest.mark.gpu

CASES = [(c1(N=4096, NZ=8, NRT=40), 7), (c2(N=8192, NZ=8, NRT=40), 8)]


def make(P, cfg, seed):
    p, T, w, Ap_in, Ap, As, Ai = inputs(cfg, seed=seed, sigma=1e-3)
    s = P.Solver(P.params_from(P.derive(to_args(cfg)), P.FP64))
    s.upload(w, T, Ap_in, Ap, As, Ai)
    return s


def three_runs(s):
    for k in range(3):
        s.run(10 * k, 10)


def join_all(threads):
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)


def test_two_solvers_on_two_threads_equal_sequential_runs(gpu):
    P = gpu
    ref = []
    for cfg, seed in CASES:
        s = make(P, cfg, seed)
        three_runs(s)
        ref.append(s.download())
        s.close()
    solvers = [make(P, cfg, seed) for cfg, seed in CASES]
    errors = []

    def work(s):
        try:
            three_runs(s)
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    join_all([threading.Thread(target=work, args=(s,)) for s in solvers])
    assert not errors
    for s, r in zip(solvers, ref):
        for a, b in zip(s.download(), r):
            if b is not None:
                assert np.array_equal(a, b)
        s.close()


def test_last_error_is_per_thread(gpu):
    """One thread fails on purpose (argument errors, rejected on the host
    before any launch) while another runs round trips: every failure carries
    its own message and the running thread sees none."""
    P = gpu
    bad, good = make(P, *CASES[0]), make(P, *CASES[1])
    seen, errors = [], []

    def fail_many():
        for _ in range(200):
            with pytest.raises(P.OpoError) as e:
                bad.run(-1, 1)
            seen.append(str(e.value))

    def run_good():
        try:
            three_runs(good)
        except Exception as e:
            errors.append(e)

    join_all([threading.Thread(target=fail_many), threading.Thread(target=run_good)])
    assert not errors
    assert len(seen) == 200 and all("negative round-trip index" in m for m in seen)
    bad.close()
    good.close()
