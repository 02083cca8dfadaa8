"""ORACLE -- test infrastructure only (used by tests/ as the checker, never by pyprop_amd).

Float64 numpy restatement of the learned preconditioner's network, its loss and its
gradient, and of one Adam step.  No reference counterpart: the network is an addition on
top of the reference's Cayley solve.  With round_bf16=True it rounds to BF16 (round to
nearest even, from float32) at the same points as the device code -- the inputs x, the
three weight matrices, the activations h1, h2 and the backward signals that enter a
matrix product -- so that the device result differs only by FP32 summation order.

Pinned in tests/test_oracle.py by central finite differences of the loss (round_bf16 off)
for every parameter block, and by the closed-form derivative of GELU.

Network (in = 2 n^2, shared by all channel blocks):
  z1 = W1 x + b1 + emb[chan]  h1 = GELU(z1)
  z2 = W2 h1 + b2             h2 = GELU(z2)
  y  = W3 h2 + b3             loss = sum |y - t / sigma|^2 / rows
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

NAMES = ("W1", "b1", "emb", "W2", "b2", "W3", "b3")


def layout(inp: int, hidden: int, nchan: int):
    sizes = [("W1", (hidden, inp)), ("b1", (hidden,)), ("emb", (nchan, hidden)), ("W2", (hidden, hidden)),
             ("b2", (hidden,)), ("W3", (inp, hidden)), ("b3", (inp,))]
    out, o = {}, 0
    for name, shp in sizes:
        k = int(np.prod(shp))
        out[name] = (o, shp)
        o += k
    return out, o


def unpack(flat: np.ndarray, inp: int, hidden: int, nchan: int):
    lay, _ = layout(inp, hidden, nchan)
    return {k: np.asarray(flat[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for k, (o, s) in lay.items()}


def pack(parts: dict, inp: int, hidden: int, nchan: int) -> np.ndarray:
    lay, total = layout(inp, hidden, nchan)
    out = np.zeros(total)
    for k, (o, s) in lay.items():
        out[o:o + int(np.prod(s))] = np.asarray(parts[k]).ravel()
    return out


def bf16(a) -> np.ndarray:
    """Round to BF16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(p: dict, x, chan, round_bf16: bool = True):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    xb = r(x)
    W1, W2, W3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    z1 = xb @ W1.T + p["b1"] + p["emb"][np.asarray(chan)]
    h1 = r(gelu(z1))
    z2 = h1 @ W2.T + p["b2"]
    h2 = r(gelu(z2))
    y = h2 @ W3.T + p["b3"]
    return y, dict(xb=xb, W1=W1, W2=W2, W3=W3, z1=z1, h1=h1, z2=z2, h2=h2)


def loss_and_grad(p: dict, x, chan, t, sigma: float = 1.0, round_bf16: bool = True):
    """loss = sum |y - t/sigma|^2 / rows and its gradient (straight through the BF16
    roundings, as the device computes it)."""
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    chan = np.asarray(chan)
    rows = x.shape[0]
    y, c = forward(p, x, chan, round_bf16)
    d = y - np.asarray(t, dtype=np.float64) / sigma
    loss = float(np.sum(d * d) / rows)
    g = 2.0 * d / rows
    gb = r(g)
    grads = {}
    grads["W3"] = gb.T @ c["h2"]
    grads["b3"] = g.sum(0)
    dz2 = (gb @ c["W3"]) * gelu_grad(c["z2"])
    dz2b = r(dz2)
    grads["b2"] = dz2.sum(0)
    grads["W2"] = dz2b.T @ c["h1"]
    dz1 = (dz2b @ c["W2"]) * gelu_grad(c["z1"])
    dz1b = r(dz1)
    grads["b1"] = dz1.sum(0)
    emb = np.zeros_like(p["emb"])
    np.add.at(emb, chan, dz1)
    grads["emb"] = emb
    grads["W1"] = dz1b.T @ c["xb"]
    return loss, grads


def adam_step(params: np.ndarray, grads: np.ndarray, m: np.ndarray, v: np.ndarray, step: int, lr: float,
              b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
    """One Adam update (Kingma & Ba, Algorithm 1) in place; step counts from 1."""
    m[:] = b1 * m + (1 - b1) * grads
    v[:] = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    params -= lr * mh / (np.sqrt(vh) + eps)
    return params
