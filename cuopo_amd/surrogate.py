"""Learned warm start for parameter sweeps.

A small MLP (5 inputs -> 4 x 256 GELU -> 4096 outputs; BF16 weights and
activations, FP32 accumulation, trained and evaluated by the library's own
MFMA kernels through opo_mlp_*) maps a sweep point's (pump power over
threshold, cavity detuning, crystal length, GDD, mirror reflectivity) to its
steady intracavity signal and idler spectrum on a coarse grid: 1024 bins per
field, one per free spectral range, centred on the carrier (modes -512..511).
The 2 x 1024 complex outputs hold, per field, the gauge-free part of the
state: the real half the amplitude spectrum normalised to unit norm, the
imaginary half log10(P / Ap0^2) / 10 of that field (replicated over the bins so
the loss weights the power like the shape).  Signal and idler phases are free
up to their sum, which the pump fixes, so they are not a target.  The bins are
counted from the cavity-resonance mode pair (m0, -m0): the pair whose idler
walk-off phase per pass, (1/v_i - 1/v_s) L_cr w_m, cancels the mirrors'
detuning phase 2 delta (mod 2 pi, smallest |m|).  The lasing mode is a steep
sawtooth of detuning and crystal length (up to ~600 modes per radian), which
the network cannot resolve from its five inputs on the carrier's grid; in the
resonance frame its target is the small residual mode offset and the power.

The family is the one the three-equation sPPLT workload generates (CW,
532 nm pump, 1000 nm signal, 1137 nm idler phase-matched by the three-wave
scan of the Lambda = 7.97 um grating; src/cuOPO.sh:73-75 gives the grating):
pump 1.2-5 x threshold, detuning -pi..pi rad, crystal 5-40 mm, GDD -500..500
fs^2, reflectivity 0.90-0.99.  Its steady states carry almost all their power
in one signal/idler mode pair (m, -m), and a single pair at the right mode
converges in a few hundred round trips where the cold start from noise needs
thousands.  The warm start is therefore the predicted spectrum's dominant
mode pair at the predicted power (random gauge phase); the solver iterates it
to its own convergence test, and the cold start stays the fallback and the
reference for the warm state's power and field.

Data flow: `generate_dataset` runs batched sweeps (opo_batch_*) from seeded
noise to the convergence test and keeps the points that reached a steady
state (compact: top modes per field); `targets` encodes them;
`Surrogate.fit` trains on the GPU; `Surrogate.warm_fields` decodes
predictions into time-domain fields on any grid N >= 1024.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import (CW, FP32C, PM_THREE_WAVE, PPLN, SPPLT, Args, Batch, _check, _ptr, derive, grids, lib,
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


# ------------------------------------------------------------ the family

def family_args(nth_rel, delta_rad, lcr_mm, gdd_fs2, refl, N=1 << 14, NZ=100, NRT=5000):
    """One point of the three-equation sPPLT family (the 2^14-sample,
    100-slice workload's generator): argv of src/cuOPO.cu:92-267 with the
    opt-in 1000 nm signal and three-wave phase matching.  GDD [fs^2] is set
    through the reference's GDD compensation percentage (GDD = -epsilon *
    beta2_s * Lcr, src/cuOPO.cu:142); argv6 is detuning / pi (Q11)."""
    a = Args(three_eqs=True, regime=CW, crystal=SPPLT, N=N, NZ=NZ, NRT=NRT, lp_nm=532.0, Lambda=7.97,
             Nth=nth_rel, R_pct=100.0 * refl, delta=delta_rad / math.pi, lcr_mm=lcr_mm, ls_um=1.0,
             pm_mode=PM_THREE_WAVE)
    if gdd_fs2 != 0:
        d = derive(a)
        a.gdd_pct = -100.0 * (gdd_fs2 * 1e-6) / (d.b2s * d.Lcr)
    return a


def sweep_args(nth_rel, delta_rad, lcr_mm, gdd_fs2, refl, N=4096, NZ=25, NRT=3000, crystal=PPLN):
    """A continuous-wave three-equation point of the degenerate PPLN sweep
    family (532 nm pump, signal and idler at 1064 nm phase-matched by the
    reference scan, Lambda = 6.97 um).  GDD [fs^2] as in family_args."""
    a = Args(three_eqs=True, regime=CW, crystal=crystal, N=N, NZ=NZ, NRT=NRT, lp_nm=532.0, Lambda=6.97,
             Nth=nth_rel, R_pct=100.0 * refl, delta=delta_rad / math.pi, lcr_mm=lcr_mm)
    if gdd_fs2 != 0:
        d = derive(a)
        a.gdd_pct = -100.0 * (gdd_fs2 * 1e-6) / (d.b2s * d.Lcr)
    return a


def resonance_mode(d, delta_rad):
    """The mode m0 (in free spectral ranges, on the grid of d) of the signal
    whose pair (m0, -m0) is resonant: the idler's walk-off phase per pass
    (1/v_i - 1/v_s) L_cr w_m0 (w stretched by N/(N-1), Q7) cancels the two
    mirror phases 2 delta modulo 2 pi, on the branch of smallest |m0|."""
    w1 = 2.0 * math.pi / d.t_rt * d.N / (d.N - 1)
    c = (1.0 / d.vi - 1.0 / d.vs) * d.Lcr * w1
    phase = (-2.0 * delta_rad + math.pi) % (2.0 * math.pi) - math.pi
    return int(round(phase / c)) if c != 0 else 0


def sample_points(n, seed=0):
    """n rows of (nth_rel, delta_rad, lcr_mm, gdd_fs2, R), uniform in RANGES."""
    rng = np.random.default_rng(seed)
    lo, hi = RANGES[:, 0], RANGES[:, 1]
    return lo + (hi - lo) * rng.random((n, 5))


# ------------------------------------------------------------ encodings

def encode_features(nth_rel, delta_rad, lcr_mm, gdd_fs2, refl):
    v = np.array([nth_rel, delta_rad, lcr_mm, gdd_fs2, refl], dtype=np.float64)
    return (2 * (v - RANGES[:, 0]) / (RANGES[:, 1] - RANGES[:, 0]) - 1).astype(np.float32)


def coarse_spectrum(A, nb=NBINS):
    """Centred nb-bin spectrum of a time-domain field: bin nb/2 + m holds the
    amplitude of mode m (field component e^{2 pi i m j / N}); bin spacing = FSR."""
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


LOG_FLOOR = -20.0  # log10(P / Ap0^2) floor
TARGET_KIND = "shape+logpower/resonance-frame"
TOPK = 8  # modes kept per field in the compact dataset


def top_modes(A, k=TOPK, nb=NBINS):
    """The k strongest coarse bins of a field: (bins [k] int, amplitudes [k])
    and the field's P = sum |X|^2 over all N modes (mean |A|^2)."""
    S = np.abs(coarse_spectrum(A, nb))
    order = np.argsort(S)[::-1][:k]
    X = np.fft.fft(A) / A.size
    return order.astype(np.int32), S[order], float(np.sum(np.abs(X) ** 2))


def _log_power(P, scale):
    return max(LOG_FLOOR, math.log10(P / scale ** 2)) if P > 0 else LOG_FLOOR


def encode_target(bins_s, amp_s, P_s, bins_i, amp_i, P_i, scale, m0=0, nb=NBINS):
    """Training target of one steady state from its compact form (carrier
    bins): [shape_s, logP_s, shape_i, logP_i], nb values each, the shapes in
    the resonance frame (signal bins shifted by -m0, idler bins by +m0; bins
    that leave the window are dropped)."""
    out = []
    for bins, amp, P, sh in ((bins_s, amp_s, P_s, -m0), (bins_i, amp_i, P_i, m0)):
        shape = np.zeros(nb)
        b = np.asarray(bins) + sh
        ok = (b >= 0) & (b < nb)
        shape[b[ok]] = np.asarray(amp)[ok]
        n = np.linalg.norm(shape)
        out += [shape / n if n > 0 else shape, np.full(nb, _log_power(P, scale) / 10.0)]
    return np.concatenate(out).astype(np.float32)


def decode_target(y, scale, m0=0):
    """(carrier bin, amplitude) of signal and idler from one output: the
    dominant bin of each predicted shape, moved back from the resonance frame,
    at the predicted power."""
    y = np.asarray(y, np.float64)
    nb = y.size // 4
    out = []
    for k, sh in ((0, m0), (1, -m0)):
        shape = y[2 * k * nb:(2 * k + 1) * nb]
        lp = 10.0 * float(np.mean(y[(2 * k + 1) * nb:(2 * k + 2) * nb]))
        out.append((int(np.argmax(shape)) + sh, scale * 10.0 ** (lp / 2)))
    return out


# ------------------------------------------------------------ data

def family_batch(points, N, NZ, max_rt, precision=FP32C, seed=0, noise_sigma=1e-20, warm=None):
    """An uploaded Batch of family points (rows of RANGES order) from seeded
    noise (seed + k, streams 0 / 1) or from warm (As, Ai) fields; also the
    derived scalars of each point."""
    made = []
    for k, pt in enumerate(points):
        d = derive(family_args(*pt, N=N, NZ=NZ, NRT=max_rt))
        T, w = grids(d)
        Ap_in = pump_input(d)
        if warm is not None:
            As, Ai = warm[k]
        else:
            As, Ai = noise(seed + k, 0, noise_sigma, N), noise(seed + k, 1, noise_sigma, N)
        made.append((params_from(d, precision), d, (w, T, Ap_in, Ap_in[:N].copy(), As, Ai)))
    b = Batch([m[0] for m in made])
    for k, m in enumerate(made):
        b.upload(k, *m[2])
    return b, [m[1] for m in made]


def run_family(points, N=4096, NZ=25, max_rt=20000, check_every=50, tol=1e-6, off_ratio=1e-6, precision=FP32C,
               seed=0, noise_sigma=1e-20, warm=None, autotune=True, fields=True, accel=None):
    """Run family points in one batch to the convergence test.  Returns a dict:
    rts, states (1 steady, 2 off, 0 capped, -1 diverged), power, device ms,
    the derived scalars and (fields=True) the final (Ap, As, Ai) per point."""
    b, ds = family_batch(points, N, NZ, max_rt, precision, seed, noise_sigma, warm)
    try:
        if autotune:
            b.autotune(iters=3)
        if accel:
            b.set_accel(*accel)  # (every, start): round-trip extrapolation, see Batch.set_accel
        rts, st = b.run_ex(max_rt, check_every=check_every, tol=tol, off_ratio=off_ratio)
        out = dict(rts=rts, states=st, power=b.power(), device_ms=b.last_run_ms(), derived=ds)
        if fields:
            out["fields"] = [b.download(k) for k in range(len(points))]
    finally:
        b.close()
    return out


def generate_dataset(n, seed=0, chunk=512, N=4096, NZ=25, max_rt=20000, noise_sigma=1e-20, precision=FP32C,
                     log=None):
    """Steady states of n random family points, cold-started from seeded
    noise.  Returns a dict of arrays over the points that reached the steady
    test: X features [m, 5], points [m, 5], scale (Ap0) [m], rts [m], and per
    field f in (s, i): bins_f [m, TOPK], amp_f [m, TOPK], P_f [m]; plus counts
    of every finish state over all n points."""
    pts = sample_points(n, seed)
    keep = {k: [] for k in ("X", "points", "scale", "rts", "bins_s", "amp_s", "P_s", "bins_i", "amp_i", "P_i")}
    counts = {"steady": 0, "off": 0, "capped": 0, "diverged": 0}
    for s0 in range(0, n, chunk):
        part = pts[s0:s0 + chunk]
        r = run_family(part, N=N, NZ=NZ, max_rt=max_rt, precision=precision, seed=seed * 1000003 + s0,
                       noise_sigma=noise_sigma)
        for k, pt in enumerate(part):
            st = int(r["states"][k])
            counts[{1: "steady", 2: "off", 0: "capped", -1: "diverged"}[st]] += 1
            _, As, Ai = r["fields"][k]
            if st != 1 or not (np.isfinite(As).all() and np.isfinite(Ai).all()):
                continue
            keep["X"].append(encode_features(*pt))
            keep["points"].append(pt)
            keep["scale"].append(r["derived"][k].Ap0)
            keep["rts"].append(int(r["rts"][k]))
            for f, A in (("s", As), ("i", Ai)):
                b_, a_, P = top_modes(A)
                keep["bins_" + f].append(b_)
                keep["amp_" + f].append(a_)
                keep["P_" + f].append(P)
        if log:
            log(s0 + len(part), dict(counts))
    out = {k: np.array(v) for k, v in keep.items()}
    out.update({k: np.array(v) for k, v in counts.items()})
    return out


def frames(points, N, NZ):
    """Resonance-frame offsets m0 of family points on grid N."""
    return [resonance_mode(derive(family_args(*p, N=N, NZ=NZ)), p[1]) for p in points]


def targets(ds, N=4096, NZ=25):
    """Training targets [m, 4096] of a dataset from generate_dataset (grid N
    x NZ, the grid the dataset was made on)."""
    m0 = frames(ds["points"], N, NZ)
    return np.array([encode_target(ds["bins_s"][j], ds["amp_s"][j], ds["P_s"][j], ds["bins_i"][j], ds["amp_i"][j],
                                   ds["P_i"][j], ds["scale"][j], m0[j]) for j in range(len(ds["scale"]))])


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

    def modes(self, points, scales, m0, use_mean=False):
        """Per point: ((signal bin, amplitude), (idler bin, amplitude)),
        carrier bins; m0 the points' resonance frames."""
        if use_mean:
            if self.mean_y is None:
                raise ValueError("no mean target: fit or load a trained surrogate first")
            Y = np.tile(self.mean_y, (len(points), 1))
        else:
            Y = self.predict(np.array([encode_features(*p) for p in points]))
        return [decode_target(Y[k], scales[k], m0[k]) for k in range(len(points))]

    def warm_fields(self, points, N, scales, seed=0, use_mean=False, NZ=100):
        """Warm-start (As, Ai) time-domain fields on grid N: the predicted
        dominant signal and idler modes at the predicted amplitudes, each with
        a random phase (use_mean: the mean training target for every point,
        i.e. the resonance pair at the mean power)."""
        rng = np.random.default_rng(seed)
        out = []
        m0 = frames(points, N, NZ)
        for (bs, a_s), (bi, a_i) in self.modes(points, scales, m0, use_mean):
            S = np.zeros(N, complex)
            I = np.zeros(N, complex)
            # carrier bin b = NBINS/2 + m holds mode m: on the N-point grid
            # (fftshifted) that is index N/2 + m
            S[(N // 2 + bs - NBINS // 2) % N] = a_s * np.exp(2j * np.pi * rng.random())
            I[(N // 2 + bi - NBINS // 2) % N] = a_i * np.exp(2j * np.pi * rng.random())
            out.append((np.fft.ifft(np.fft.ifftshift(S)) * N, np.fft.ifft(np.fft.ifftshift(I)) * N))
        return out
