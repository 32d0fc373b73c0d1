"""Learned warm start for parameter sweeps.

A small MLP (5 inputs -> 4 x 256 GELU -> 4096 outputs; BF16 weights and
activations, FP32 accumulation, trained and evaluated by the library's own
MFMA kernels through opo_mlp_*) maps a sweep point's (pump power over
threshold, cavity detuning, crystal length, GDD, mirror reflectivity) to its
steady state in a gauge-free form (the signal and idler mode phases are free;
only their sum is fixed by the pump):
  - the amplitude shape: |spectrum| of signal and idler on 1024 bins
    (free-spectral-range spacing, centred on the carrier), each normalised
    to unit norm -- the real halves of the 2 x 1024 real/imag output;
  - the power: log10(P / Ap0^2) / 10 of each field, replicated over that
    field's imaginary half so the loss weights it like the shape.
An off point (decayed below threshold) has a zero shape and the power floor.
Warm fields take the predicted shape at the predicted power with random mode
phases.  The solver starts from that instead of noise and iterates to its own
convergence test; the cold start stays the fallback and the yardstick of the
prediction's error.

Data flow: `generate_dataset` runs batched sweeps (opo_batch_*) to steady
state and encodes the fields of the points that finished; `Surrogate.fit`
trains; `Surrogate.warm_fields` decodes predictions into time-domain fields on
any grid N >= 1024.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import (CW, FP32C, FP64, PPLN, Args, Batch, OpoError, _check, _ptr, derive, grids, lib,
               noise, params_from, pump_input)

NBINS = 1024
N_OUT = 4 * NBINS
# feature ranges: pump power / threshold, detuning [rad], crystal length [mm],
# GDD [fs^2], signal mirror reflectivity
RANGES = np.array([[1.2, 5.0], [-math.pi, math.pi], [5.0, 40.0], [-500.0, 500.0], [0.90, 0.99]])


class MLP:
    """opo_mlp: BF16/FP32 MLP on the GPU (no CPU path)."""

    def __init__(self, n_in=5, width=256, depth=4, n_out=N_OUT, max_batch=256, seed=0, device=0):
        h = C.c_void_p()
        _check(lib().opo_mlp_create(n_in, width, depth, n_out, max_batch, seed, device, C.byref(h)))
        self._h = h
        self.shape = (n_in, width, depth, n_out)
        self.max_batch = max_batch
        n = C.c_size_t()
        _check(lib().opo_mlp_param_count(self._h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "_h", None):
            lib().opo_mlp_destroy(self._h)
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

    def get_weights(self):
        w = np.empty(self.nparams, np.float32)
        _check(lib().opo_mlp_get_weights(self._h, _ptr(w), w.size))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _check(lib().opo_mlp_set_weights(self._h, _ptr(w), w.size))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        out = np.empty((x.shape[0], self.shape[3]), np.float32)
        for s in range(0, x.shape[0], self.max_batch):
            xb = np.ascontiguousarray(x[s:s + self.max_batch])
            yb = np.empty((xb.shape[0], self.shape[3]), np.float32)
            _check(lib().opo_mlp_forward(self._h, _ptr(xb), xb.shape[0], _ptr(yb)))
            out[s:s + xb.shape[0]] = yb
        return out

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_float()
        _check(lib().opo_mlp_gradients(self._h, _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
        return loss.value, g

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        _check(lib().opo_mlp_train_step(self._h, _ptr(x), _ptr(t), x.shape[0], float(lr), C.byref(loss)))
        return loss.value


# ------------------------------------------------------------ encodings

def encode_features(nth_rel, delta_rad, lcr_mm, gdd_fs2, refl):
    v = np.array([nth_rel, delta_rad, lcr_mm, gdd_fs2, refl], dtype=np.float64)
    return (2 * (v - RANGES[:, 0]) / (RANGES[:, 1] - RANGES[:, 0]) - 1).astype(np.float32)


def coarse_spectrum(A, nb=NBINS):
    """Centred nb-bin spectrum of a time-domain field (bin spacing = FSR)."""
    N = A.size
    X = np.fft.fftshift(np.fft.fft(A)) / N
    out = np.zeros(nb, complex)
    lo = N // 2 - nb // 2
    src_lo, src_hi = max(lo, 0), min(lo + nb, N)
    out[src_lo - lo:src_hi - lo] = X[src_lo:src_hi]
    return out


def field_from_coarse(S, N):
    """Inverse of coarse_spectrum onto an N-point grid (bins outside are zero)."""
    nb = S.size
    X = np.zeros(N, complex)
    lo = N // 2 - nb // 2
    dst_lo, dst_hi = max(lo, 0), min(lo + nb, N)
    X[dst_lo:dst_hi] = S[dst_lo - lo:dst_hi - lo]
    return np.fft.ifft(np.fft.ifftshift(X)) * N


def gauge_fix(S, I):
    """Signal and idler phases are free up to their sum (the pump fixes
    phi_s + phi_i): rotate so the strongest signal bin is real positive."""
    k = int(np.argmax(np.abs(S)))
    ph = np.exp(-1j * np.angle(S[k])) if np.abs(S[k]) > 0 else 1.0
    return S * ph, I * np.conj(ph)


LOG_FLOOR = -20.0  # log10(P / Ap0^2) of an off point
TARGET_KIND = "shape+logpower"


def _log_power(P, scale):
    return max(LOG_FLOOR, math.log10(P / scale ** 2)) if P > 0 else LOG_FLOOR


def encode_target(As, Ai, scale):
    """Training target of one steady state (As, Ai time-domain fields, scale =
    the pump amplitude Ap0): [shape_s, logP_s, shape_i, logP_i], 1024 each."""
    out = []
    for A in (As, Ai):
        a = np.abs(coarse_spectrum(A))
        P = float(np.sum(a * a))
        lp = _log_power(P, scale)
        shape = a / math.sqrt(P) if lp > LOG_FLOOR else np.zeros_like(a)
        out += [shape, np.full(a.size, lp / 10.0)]
    return np.concatenate(out).astype(np.float32)


def decode_target(y, scale):
    """Amplitude spectra (signal, idler) on the coarse grid from one output."""
    y = np.asarray(y, np.float64)
    nb = y.size // 4
    out = []
    for k in range(2):
        shape = np.maximum(y[2 * k * nb:(2 * k + 1) * nb], 0.0)
        lp = 10.0 * float(np.mean(y[(2 * k + 1) * nb:(2 * k + 2) * nb]))
        n = float(np.linalg.norm(shape))
        if lp <= LOG_FLOOR + 1.0 or n == 0:
            out.append(np.zeros(nb))
            continue
        out.append(shape / n * scale * 10.0 ** (lp / 2))
    return out[0], out[1]


# ------------------------------------------------------------ sweep points

def sweep_args(nth_rel, delta_rad, lcr_mm, gdd_fs2, refl, N=4096, NZ=25, NRT=3000, crystal=PPLN):
    """A continuous-wave three-equation point of the sweep family (PPLN,
    532 nm pump, degenerate signal/idler phase-matched by the reference scan,
    Lambda = 6.97 um).  GDD [fs^2] is set through the reference's GDD
    compensation percentage (GDD = -epsilon * beta2_s * Lcr, cuOPO.cu:142)."""
    a = Args(three_eqs=True, regime=CW, crystal=crystal, N=N, NZ=NZ, NRT=NRT, lp_nm=532.0, Lambda=6.97,
             Nth=nth_rel, R_pct=100.0 * refl, delta=delta_rad / math.pi, lcr_mm=lcr_mm)
    d = derive(a)
    if gdd_fs2 != 0:
        a.gdd_pct = -100.0 * (gdd_fs2 * 1e-6) / (d.b2s * d.Lcr)
    return a


def sample_points(n, seed=0):
    rng = np.random.default_rng(seed)
    lo, hi = RANGES[:, 0], RANGES[:, 1]
    return lo + (hi - lo) * rng.random((n, 5))


def _instance(a, precision, As, Ai):
    d = derive(a)
    p = params_from(d, precision)
    T, w = grids(d)
    Ap_in = pump_input(d)
    return p, d, (w, T, Ap_in, Ap_in[: d.N].copy(), As, Ai)


def run_sweep(points, N=4096, NZ=25, max_rt=3000, check_every=50, tol=1e-6, precision=FP32C, seed=0,
              warm=None, noise_sigma=1e-20, tiles=(0, 0)):
    """Run sweep points (rows of [nth_rel, delta, lcr_mm, gdd_fs2, R]) in one
    batch to convergence.  warm: optional list of (As, Ai) initial fields.
    Returns (fields, rts, final power, device ms, derived scalars, states
    (1 steady, 2 off, 0 capped, -1 diverged))."""
    made = []
    for k, pt in enumerate(points):
        a = sweep_args(*pt, N=N, NZ=NZ, NRT=max_rt)
        if warm is not None:
            As, Ai = warm[k]
        else:
            As = noise(seed + k, 0, noise_sigma, N)
            Ai = noise(seed + k, 1, noise_sigma, N)
        made.append(_instance(a, precision, As, Ai))
    b = Batch([m[0] for m in made])
    if tiles != (0, 0):
        b.set_tiles(*tiles)
    for k, m in enumerate(made):
        b.upload(k, *m[2])
    try:
        rts, states = b.run_ex(max_rt, check_every=check_every, tol=tol, off_ratio=1e-6)
    except OpoError as e:
        if e.status != -5:
            raise
        rts, states = None, None
    fields = [b.download(k) for k in range(len(points))]
    power = b.power()
    ms = b.last_run_ms()
    b.close()
    return fields, rts, power, ms, [m[1] for m in made], states


def generate_dataset(n, seed=0, chunk=256, **kw):
    """Encoded finished points (steady or off) of n random sweep points:
    (features [m,5], targets [m,4096], scales [m], rts [m], states [m])."""
    pts = sample_points(n, seed)
    X, Y, S, R, ST = [], [], [], [], []
    for s in range(0, n, chunk):
        part = pts[s:s + chunk]
        fields, rts, power, ms, ds, states = run_sweep(part, seed=seed * 100003 + s, **kw)
        if states is None:
            continue
        for k, pt in enumerate(part):
            _, As, Ai = fields[k]
            if states[k] not in (1, 2) or not (np.isfinite(As).all() and np.isfinite(Ai).all()):
                continue
            scale = ds[k].Ap0
            X.append(encode_features(*pt))
            Y.append(encode_target(As, Ai, scale) if states[k] == 1 else
                     encode_target(np.zeros_like(As), np.zeros_like(Ai), scale))
            S.append(scale)
            R.append(int(rts[k]))
            ST.append(int(states[k]))
    return np.array(X), np.array(Y), np.array(S), np.array(R), np.array(ST)


class Surrogate:
    """Trained MLP + the decoding onto solver grids.  `mean_y` (the mean
    training target) gives the input-independent seed used as the ablation
    of the network."""

    def __init__(self, seed=0, device=0, max_batch=256):
        self.net = MLP(max_batch=max_batch, seed=seed, device=device)
        self.mean_y = None

    def fit(self, X, Y, epochs=200, batch=256, lr=1e-3, seed=0, log=None):
        rng = np.random.default_rng(seed)
        n = X.shape[0]
        self.mean_y = Y.mean(axis=0).astype(np.float32)
        hist = []
        for ep in range(epochs):
            order = rng.permutation(n)
            tot = 0.0
            for s in range(0, n - batch + 1 if n >= batch else 1, batch):
                idx = order[s:s + batch]
                tot += self.net.train_step(X[idx], Y[idx], lr) * len(idx)
            hist.append(tot / max(1, (n // batch) * batch if n >= batch else n))
            if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                log(ep, hist[-1])
        return np.array(hist)

    def predict(self, X):
        return self.net.forward(X)

    def warm_fields_mixed(self, points, Ns, scales, floor_sigma=1e-12, seed=0, use_mean=False):
        """warm_fields for points on different grids (Ns[k] for point k)."""
        out = []
        for k, (pt, N) in enumerate(zip(points, Ns)):
            out.append(self.warm_fields([pt], int(N), [scales[k]], floor_sigma, seed + 7919 * k, use_mean)[0])
        return out

    def warm_fields(self, points, N, scales, floor_sigma=1e-12, seed=0, use_mean=False):
        """Predicted (As, Ai) time-domain fields for sweep points on grid N
        (use_mean: the mean training target for every point instead)."""
        if use_mean:
            if self.mean_y is None:
                raise ValueError("no mean target: fit or load a trained surrogate first")
            Y = np.tile(self.mean_y, (len(points), 1))
        else:
            Y = self.predict(np.array([encode_features(*p) for p in points]))
        rng = np.random.default_rng(seed)
        out = []
        for k in range(len(points)):
            aS, aI = decode_target(Y[k], scales[k])
            # predicted amplitudes, free mode phases drawn at random
            S = aS * np.exp(2j * np.pi * rng.random(aS.size))
            I = aI * np.exp(2j * np.pi * rng.random(aI.size))
            As = field_from_coarse(S, N) + noise(seed + k, 0, floor_sigma, N)
            Ai = field_from_coarse(I, N) + noise(seed + k, 1, floor_sigma, N)
            out.append((As, Ai))
        return out
