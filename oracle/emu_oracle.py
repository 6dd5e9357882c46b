"""FP64 numpy restatement of the emulator's MLP (the network of include/opensps_emu.h): forward, loss and the
gradient of the loss with respect to the flat parameter vector.

TEST INFRASTRUCTURE ONLY: the tests use it as the checker of the HIP forward / backward / Adam kernels; the product
never imports it. No reference counterpart (the reference has no trained network on this path); the mathematics it
restates: y = W_L g(... g(W_0 x' + b_0) ...) + b_L with g the exact GELU x Phi(x), x' = (x - in_mu) / in_sd,
L = sum_i sum_o w_o (y_io - t'_io)^2 / n with t' = (t - out_mu) / out_sd, and the chain rule. With bf16=True the
matrix operands are rounded to bfloat16 (round to nearest even) where the kernels round them: the scaled inputs, the
weights, the activations and the pre-activation gradients; sums stay FP64.
"""
from __future__ import annotations

import math

import numpy as np

_ERF = np.vectorize(math.erf, otypes=[np.float64])


def bf16(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _ERF(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _ERF(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def unflatten(p, n_in, n_out, hidden, layers):
    """Flat vector -> [(W, b)] in layer order (W: n_out_l x n_in_l)."""
    out, at = [], 0
    dims = [n_in] + [hidden] * layers + [n_out]
    for l in range(layers + 1):
        k, n = dims[l], dims[l + 1]
        W = np.asarray(p[at:at + n * k], dtype=np.float64).reshape(n, k)
        at += n * k
        b = np.asarray(p[at:at + n], dtype=np.float64)
        at += n
        out.append((W, b))
    assert at == len(p)
    return out


def flatten(layers_wb):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers_wb])


def forward(p, x, dims, scale, round_bf16=False):
    """Network outputs in physical units (y' out_sd + out_mu) and the saved activations."""
    n_in, n_out, hidden, layers = dims
    in_mu, in_sd, out_mu, out_sd = scale
    R = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    L = unflatten(p, n_in, n_out, hidden, layers)
    h = R((np.asarray(x, np.float64) - in_mu) / in_sd)
    acts, pres = [h], []
    for l, (W, b) in enumerate(L):
        z = h @ R(W).T + b
        pres.append(z)
        if l < layers:
            h = R(gelu(z))
            acts.append(h)
    y = pres[-1]
    return y * out_sd + out_mu, (acts, pres, L)


def loss_and_grad(p, x, t, dims, scale, w=None, round_bf16=False):
    n_in, n_out, hidden, layers = dims
    in_mu, in_sd, out_mu, out_sd = scale
    R = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    _, (acts, pres, L) = forward(p, x, dims, scale, round_bf16)
    n = x.shape[0]
    w = np.ones(n_out) if w is None else np.asarray(w, np.float64)
    d = pres[-1] - (np.asarray(t, np.float64) - out_mu) / out_sd
    loss = float(np.sum(w * d * d) / n)
    g = R(2.0 * w * d / n)
    grads = [None] * (layers + 1)
    for l in range(layers, -1, -1):
        W, b = L[l]
        grads[l] = (g.T @ acts[l], g.sum(axis=0))
        if l == 0:
            break
        g = R((g @ R(W)) * gelu_grad(pres[l - 1]))
    return loss, flatten(grads)
