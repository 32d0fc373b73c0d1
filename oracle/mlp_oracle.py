"""FP64 CPU oracle for the warm-start MLP (TEST INFRASTRUCTURE ONLY).

Same network as libcuopo_amd's opo_mlp_*: n_in -> width (GELU) x depth -> n_out,
flat parameters laid out per layer as W [dout x din] row-major then b [dout].
`quantize=True` rounds inputs, weights and hidden activations to BF16 at the
points where the GPU stores them, and evaluates everything else in FP64, so
the GPU's forward pass and gradients are checked against FP64 arithmetic on
identical BF16 operands.  Its own gradient is pinned by finite differences
(tests/test_mlp_oracle.py).  The reference (cuOPO) has no learned component.
"""
import math

import numpy as np

_erf = np.vectorize(math.erf)


def bf16(x):
    """Round float values to BF16 (round to nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def layers(n_in, width, depth, n_out):
    din = [n_in] + [width] * depth
    dout = [width] * depth + [n_out]
    offs, o = [], 0
    for a, b in zip(din, dout):
        offs.append((o, o + a * b))
        o += a * b + b
    return din, dout, offs, o


def unpack(theta, shape):
    din, dout, offs, _ = layers(*shape)
    out = []
    for a, b, (wo, bo) in zip(din, dout, offs):
        out.append((theta[wo:wo + a * b].reshape(b, a), theta[bo:bo + b]))
    return out


def gelu(z):
    return 0.5 * z * (1 + _erf(z / math.sqrt(2)))


def gelu_grad(z):
    return 0.5 * (1 + _erf(z / math.sqrt(2))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(theta, shape, x, quantize=True):
    q = bf16 if quantize else (lambda v: np.asarray(v, np.float64))
    Ws = unpack(np.asarray(theta, np.float64), shape)
    h = q(x)
    hs, zs = [h], []
    for l, (W, b) in enumerate(Ws):
        z = h @ q(W).T + b
        zs.append(z)
        if l < len(Ws) - 1:
            h = q(gelu(z))
            hs.append(h)
        else:
            return z, hs, zs


def loss_and_grad(theta, shape, x, t, quantize=True):
    """Mean-squared error over all outputs and its gradient (flat, FP64)."""
    q = bf16 if quantize else (lambda v: np.asarray(v, np.float64))
    Ws = unpack(np.asarray(theta, np.float64), shape)
    y, hs, zs = forward(theta, shape, x, quantize)
    n = y.size
    loss = float(np.sum((y - t) ** 2) / n)
    dz = 2 * (y - t) / n
    grads = [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        W, _ = Ws[l]
        grads[l] = (dz.T @ hs[l], dz.sum(axis=0))
        if l > 0:
            dz = (dz @ q(W)) * gelu_grad(zs[l - 1])
    flat = np.concatenate([np.concatenate([gw.reshape(-1), gb]) for gw, gb in grads])
    return loss, flat
