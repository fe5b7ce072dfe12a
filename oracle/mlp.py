"""CPU oracle of the learned-closure MLP (test infrastructure: only tests/ use it).

The network of littleharsh_amd/csrc/lh_mlp.hip restated in numpy: layers
nin -> nhid (tanh) x nlayers -> nout (linear), loss = mean squared error over the outputs,
parameters flattened per layer as W (out x in, row-major) then b (out).  There is no
reference counterpart: the closure is project-defined.

`loss_grad(..., bf16=True)` rounds the operands of every matrix product to bfloat16
(round to nearest even) as the GPU's matrix cores see them -- the layer inputs, the
weights and, in the backward pass, the gradients of the pre-activations -- and
accumulates in float64; `bf16=False` is exact float64 backpropagation, which
tests/test_closure_oracle.py pins with central finite differences.
"""
from __future__ import annotations

import numpy as np


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def shapes(nin: int, nhid: int, nlayers: int, nout: int):
    """(out, in) of every weight layer."""
    dims = [nin] + [nhid] * nlayers + [nout]
    return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]


def unpack(params: np.ndarray, shp):
    out, o = [], 0
    for n, k in shp:
        W = params[o:o + n * k].reshape(n, k)
        o += n * k
        b = params[o:o + n]
        o += n
        out.append((W, b))
    assert o == params.size
    return out


def loss_grad(params, x, y, shp, bf16_ops: bool = False):
    """Loss and gradient (flattened like params) of one batch."""
    rnd = bf16 if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(np.asarray(params, dtype=np.float64), shp)
    L = len(layers)
    X = [rnd(x)]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = X[-1] @ rnd(W).T + b
        Z.append(z)
        if l < L - 1:
            X.append(rnd(np.tanh(z)))
    yp = Z[-1]
    B, nout = y.shape
    d = yp - y
    loss = float(np.sum(d * d) / (B * nout))
    dH = 2.0 * d / (B * nout)
    grads = [None] * L
    for l in range(L - 1, -1, -1):
        W, b = layers[l]
        dZ = dH if l == L - 1 else dH * (1.0 - np.tanh(Z[l]) ** 2)
        dZr = rnd(dZ)
        grads[l] = (dZr.T @ X[l], dZ.sum(axis=0))
        if l > 0:
            dH = dZr @ rnd(W)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, flat


def predict(params, x, shp):
    layers = unpack(np.asarray(params, dtype=np.float64), shp)
    h = np.asarray(x, dtype=np.float64)
    for l, (W, b) in enumerate(layers):
        h = h @ W.T + b
        if l < len(layers) - 1:
            h = np.tanh(h)
    return h
