"""FP64 numpy restatement of the surrogate's forward and backward passes
(test infrastructure).  With bf16=True the GEMM operands are rounded to BF16
exactly as the device does (round to nearest even), the products accumulated
in FP64; tests/test_mlp_reference.py pins the backward pass against central
finite differences."""
from __future__ import annotations

import numpy as np


def to_bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(flat, dims):
    Ws, bs, o = [], [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        Ws.append(np.asarray(flat[o:o + a * b], np.float64).reshape(b, a))
        o += a * b
        bs.append(np.asarray(flat[o:o + b], np.float64))
        o += b
    return Ws, bs


def _rounder(bf16):
    if bf16:
        return lambda v: to_bf16(np.asarray(v, np.float64).astype(np.float32))
    return lambda v: np.asarray(v, np.float64)


def forward(flat, dims, x, bf16=False):
    Ws, bs = unpack(flat, dims)
    rnd = _rounder(bf16)
    H, Z = [np.asarray(x, np.float64)], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = rnd(H[-1]) @ rnd(W).T + b
        Z.append(z)
        H.append(z if l == len(Ws) - 1 else np.maximum(z, 0.0))
    return H, Z


def loss_grad(flat, dims, x, t, w, bf16=False):
    Ws, bs = unpack(flat, dims)
    rnd = _rounder(bf16)
    H, Z = forward(flat, dims, x, bf16)
    rows = x.shape[0]
    d = H[-1] - t
    loss = float(np.sum(w[None, :] * d * d)) / rows
    dz = 2.0 * w[None, :] * d / rows
    grads = []
    for l in range(len(Ws) - 1, -1, -1):
        gW = rnd(dz).T @ rnd(H[l])
        gb = dz.sum(axis=0)
        grads.append((gW, gb))
        if l > 0:
            dh = rnd(dz) @ rnd(Ws[l])
            dz = dh * (Z[l - 1] > 0)
    flat_g = []
    for gW, gb in reversed(grads):
        flat_g += [gW.ravel(), gb]
    return loss, np.concatenate(flat_g)
