"""Float64 reference of the surrogate network's forward and backward pass -- TEST
INFRASTRUCTURE ONLY (the checker of the GPU kernels in cudaprob3_amd/csrc/mlp.hip).

Network: 9 inputs -> 5 hidden layers of 256 with GELU (tanh form) -> 9 linear outputs;
loss = mean over samples and outputs of (y - t)^2.  Parameters use the library's flat
layout: for each layer, W [outputs][inputs] row-major, then b.

With ``bf16=True`` the values the GPU stores in bfloat16 are rounded the same way (round to
nearest even): the weights and each back-propagated dZ are rounded to bf16; the scaled inputs
and each hidden layer's GELU output enter the forward products as two bf16 terms, hi = bf16(a)
and lo = bf16(a - hi), and the weight gradients as hi alone; everything else stays float64.
Without it this is the exact math, which ``tests/test_surrogate_oracle.py`` pins against
finite differences.
"""
from __future__ import annotations

import math

import numpy as np

SHAPES = [(256, 9), (256, 256), (256, 256), (256, 256), (256, 256), (9, 256)]


def param_count() -> int:
    return sum(n * k + n for n, k in SHAPES)


def unpack(params):
    out, off = [], 0
    for n, k in SHAPES:
        W = np.asarray(params[off:off + n * k], np.float64).reshape(n, k)
        off += n * k
        b = np.asarray(params[off:off + n], np.float64)
        off += n
        out.append((W, b))
    return out


def pack(layers) -> np.ndarray:
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def bf16_round(x):
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


GELU_C = 0.044715
GELU_K = math.sqrt(2.0 / math.pi)


def gelu(z):
    """GELU, tanh form: 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))."""
    return 0.5 * z * (1.0 + np.tanh(GELU_K * (z + GELU_C * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_K * (z + GELU_C * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_K * (1.0 + 3.0 * GELU_C * z * z)


def two_term(a):
    """hi + lo with hi = bf16(a), lo = bf16(a - hi), both from float32 values."""
    a = np.asarray(a, np.float32)
    hi = bf16_round(a)
    return hi + bf16_round(a - hi.astype(np.float32))


def forward_backward(params, x, t, shift, scale, bf16=False):
    """Returns (loss, flat gradients, outputs) for raw inputs x [M][9] and targets t [M][9]."""
    ident = (lambda a: np.asarray(a, np.float64))
    rnd = bf16_round if bf16 else ident
    fwd = two_term if bf16 else ident
    layers = [(rnd(W), b) for W, b in unpack(params)]
    M = x.shape[0]
    a = (np.asarray(x, np.float32) - np.asarray(shift, np.float32)) * np.asarray(scale, np.float32) \
        if bf16 else (np.asarray(x, np.float64) - shift) * scale
    acts, grad_acts, zs = [fwd(a)], [rnd(a)], []
    for i, (W, b) in enumerate(layers):
        z = acts[-1] @ W.T + b
        zs.append(z)
        if i < len(layers) - 1:
            g = gelu(z)
            acts.append(fwd(g))
            grad_acts.append(rnd(g))
    y = zs[-1]
    e = y - t
    loss = float((e * e).sum() / (9 * M))
    dz = rnd(2.0 * e / (9 * M))
    grads = [None] * len(layers)
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        grads[i] = (dz.T @ grad_acts[i], dz.sum(axis=0))
        if i > 0:
            dz = rnd((dz @ W) * gelu_grad(zs[i - 1]))
    return loss, pack(grads), y
