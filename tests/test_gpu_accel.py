"""Round-trip extrapolation of slowly converging batch instances (GPU).

The headline sweep's points, warm-started by the committed surrogate on the
grid it was trained on (4096 samples, 25 slices), run to the steady-state
test twice: by plain iteration, and with opo_batch_set_accel (every 400 round
trips from 1000 on).  Extrapolation must save round trips and reach the same
steady states: the same dominant signal mode and, on those points, the same
steady power to within the steady test's own resolution (an instance that
passes |dP| <= 1e-6 P over 50 round trips while relaxing at 1e-3 per round
trip can still be ~2e-5 from its fixed point).  Instances that never jumped
restart from their own snapshots unchanged, so most powers agree bit for
bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cuopo_amd", "data",
                       "surrogate.npz")


def test_extrapolation_reaches_the_same_steady_states_in_fewer_round_trips(gpu):
    P = gpu
    from cuopo_amd import surrogate as S
    from cuopo_amd.train_surrogate import load
    sur, meta = load(WEIGHTS, max_batch=256)
    N, NZ, cap = meta["N"], meta["NZ"], 20000
    pts = S.sample_points(256, 31)
    warm = sur.warm_fields(pts, N, [S.derive(S.family_args(*p, N=N, NZ=NZ)).Ap0 for p in pts], seed=31, NZ=NZ)
    sur.net.close()

    def run(accel):
        b, _ = S.family_batch(pts, N, NZ, cap, P.FP64, warm=warm)
        if accel:
            b.set_accel(400, 1000)
        rts, st = b.run_ex(cap, check_every=50, tol=1e-6, off_ratio=1e-6)
        pw = b.power()
        modes = np.array([int(np.argmax(np.abs(np.fft.fft(b.download(k)[1])))) for k in range(len(pts))])
        b.close()
        return np.where(rts > 0, rts, cap), st, pw, modes

    r0, s0, p0, m0 = run(False)
    r1, s1, p1, m1 = run(True)
    both = (s0 == 1) & (s1 == 1)
    same = both & (m0 == m1)
    rel = np.abs(p1 - p0) / p0
    print(f"round trips {r0.sum()} -> {r1.sum()}; steady {int((s0 == 1).sum())} -> {int((s1 == 1).sum())}; "
          f"same mode {int(same.sum())} of {int(both.sum())}; max rel power {rel[same].max():.2e}; "
          f"bit-equal {int((p0 == p1).sum())}")
    assert r1.sum() < r0.sum()
    assert (s1 == 1).sum() >= (s0 == 1).sum() - 2
    assert same.sum() >= 0.97 * both.sum()
    assert rel[same].max() < 1e-4
    assert (p0 == p1).sum() >= len(pts) // 2
