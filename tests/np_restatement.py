"""Independent NumPy restatement of the reference's round trip, written from
src/cuOPO.cu:474-610 and src/headers/{cwes2,cwes3,functions}.cuh without
reference to the C oracle.  Vectorised over samples, numpy.fft for the
transforms.  Used to cross-check the C oracle on small grids (test only).
"""
import numpy as np


def to_freq(x):
    """cufftExecC2C(CUFFT_INVERSE) followed by CUFFTscale: sum x e^{+2 pi i jk/N} / N."""
    return np.fft.ifft(x)


def to_time(X):
    """cufftExecC2C(CUFFT_FORWARD): sum X e^{-2 pi i jk/N}."""
    return np.fft.fft(X)


def dadz(p, Ap, As, Ai, z):
    ex = lambda a: np.exp(1j * a)
    if p.three_eqs:  # cwes3.cuh:37-39
        return (1j * p.kp * As * Ai * ex(-p.dk * z),
                1j * p.ks * Ap * np.conj(Ai) * ex(p.dk * z),
                1j * p.ki * Ap * np.conj(As) * ex(p.dk * z))
    # cwes2.cuh:37-38
    return (1j * p.kp * As * As * ex(p.dk * z), 1j * p.ks * Ap * np.conj(As) * ex(-p.dk * z), None)


def nl_half(p, A, z):
    Ap, As, Ai = A
    dz = p.dz
    k1 = dadz(p, Ap, As, Ai, z)
    if p.three_eqs:  # stages on the unchanged fields (LinealCombination results unused)
        k2 = dadz(p, Ap, As, Ai, z + dz / 4)
        k3 = dadz(p, Ap, As, Ai, z + dz / 4)
        k4 = dadz(p, Ap, As, Ai, z + dz / 2)
    else:  # stages on A + 0.5 k1, A + 0.5 k2, A + k3 (increments not scaled by the step)
        k2 = dadz(p, Ap + 0.5 * k1[0], As + 0.5 * k1[1], None, z + dz / 4)
        k3 = dadz(p, Ap + 0.5 * k2[0], As + 0.5 * k2[1], None, z + dz / 4)
        k4 = dadz(p, Ap + k3[0], As + k3[1], None, z + dz / 2)
    h = dz / 2
    out = []
    for f, A0 in enumerate((Ap, As, Ai)):
        if A0 is None:
            out.append(None)
            continue
        out.append(A0 + (k1[f] + 2 * k2[f] + 2 * k3[f] + k4[f]) * h / 6)
    return out


def linear(p, w, A):
    Ap, As, Ai = A
    z = p.dz

    def step(X, D, b2, b3, alpha):
        Xw = to_freq(X)
        Xw = Xw * np.exp(1j * z * w * (D + 0.5 * w * b2 + w * w * b3 / 6)) * np.exp(-0.5 * alpha * z)
        return to_time(Xw)

    if p.three_eqs:  # cwes3.cuh:117-119, pump walk-off against the idler
        return (step(Ap, 1 / p.vp - 1 / p.vi, p.b2p, p.b3p, p.alpha_crp),
                step(As, 0.0, p.b2s, p.b3s, p.alpha_crs),
                step(Ai, 1 / p.vi - 1 / p.vs, p.b2i, p.b3i, p.alpha_cri))
    return (step(Ap, 1 / p.vp - 1 / p.vs, p.b2p, p.b3p, p.alpha_crp),
            step(As, 0.0, p.b2s, p.b3s, p.alpha_crs), None)


def crystal_pass(p, w, A):
    z = 0.0
    for _ in range(p.NZ):
        A = nl_half(p, A, z)
        A = linear(p, w, A)
        A = nl_half(p, A, z)
        z += p.dz
    return A


def pump_slot(p, Ap_in, nn):
    N = p.N
    if p.regime == 0:
        return Ap_in[:N].copy()
    if nn == 0:
        off = 0
    elif 0 < nn < p.NRT - 1:
        off = nn * N
    else:
        off = (N * p.NRT - 1) - N
    return Ap_in[off:off + N].copy()


def run(p, w, T, Ap_in, As, Ai, nrt):
    A = [None, As.copy(), None if Ai is None else Ai.copy()]
    powers = []
    for nn in range(nrt):
        A[0] = pump_slot(p, Ap_in, nn)
        A = list(crystal_pass(p, w, A))
        if p.use_eom:
            ph = np.exp(1j * p.mod_depth * np.sin(2 * np.pi * p.fpm * T))
            A[1] = A[1] * ph
            if A[2] is not None:
                A[2] = A[2] * ph
        if p.GDD != 0:
            for f in (1, 2):
                if A[f] is not None:
                    A[f] = to_time(to_freq(A[f]) * np.exp(1j * 0.5 * w * w * p.GDD))
        A[1] = np.sqrt(p.Rs) * A[1] * np.exp(1j * np.pi * p.delta)
        if A[2] is not None:
            A[2] = np.sqrt(p.Ri) * A[2] * np.exp(1j * np.pi * p.delta)
        powers.append(float(np.sum(np.abs(A[1]) ** 2)))
    return A, np.array(powers)
