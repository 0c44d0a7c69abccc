"""CPU restatement of the learned column correction MLP (forward, backward, Adam).

TEST INFRASTRUCTURE ONLY: the CPU reference for the engine's MLP kernels; only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it.

The network has no counterpart in the SpeedyWeather reference: it is the learned per-column
tendency correction (32 -> 128 -> 128 -> 32, GELU) that a dry primitive-equation step adds
to its spectral tendencies.  The function it defines:

    h0 = bf16(x)
    z1 = bf16(W1) h0 + b1,  h1 = bf16(gelu(z1))
    z2 = bf16(W2) h1 + b2,  h2 = bf16(gelu(z2))
    y  = bf16(W3) h2 + b3
    L  = mean((y - t)^2) over all B x 32 values

with ``gelu(z) = z * Phi(z)`` (the exact, erf-based form) and ``bf16`` the round-to-nearest-
even cut of an FP32 value to 8 mantissa bits.  The gradient is the derivative of that
function with the roundings treated as identity (straight-through), the convention the
engine's backward kernel follows.  Products accumulate in FP64 here (the engine: FP32).

Parameter vector layout (float32, 24 864 values): W1[128][32] b1[128] W2[128][128] b2[128]
W3[32][128] b3[32], W[o][i] maps input i to output o.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_HID, N_OUT = 32, 128, 32
SHAPES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_OUT, N_HID)), ("b3", (N_OUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    """Split the flat parameter vector into the named arrays (views)."""
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float64.

    NaN and infinities are not expected here (the network's inputs are finite)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


_erf = np.vectorize(math.erf, otypes=[np.float64])


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(params, x, rounding=True):
    """Returns y (B x 32) and the cache (pre-activations, rounded layer inputs)."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    # the engine computes gelu and the pre-activations in FP32 before rounding
    f32 = (lambda v: np.asarray(v, dtype=np.float32)) if rounding else (lambda v: v)
    P = unpack(np.asarray(params, dtype=np.float64))
    W = {k: r(P[k]) for k in ("W1", "W2", "W3")}
    h0 = r(x)
    z1 = np.asarray(f32(h0 @ W["W1"].T + P["b1"]), dtype=np.float64)
    h1 = r(f32(gelu(z1)))
    z2 = np.asarray(f32(h1 @ W["W2"].T + P["b2"]), dtype=np.float64)
    h2 = r(f32(gelu(z2)))
    y = h2 @ W["W3"].T + P["b3"]
    return y, dict(W=W, h0=h0, z1=z1, h1=h1, z2=z2, h2=h2)


def loss_and_gradient(params, x, t, rounding=True):
    """Mean squared error and its gradient w.r.t. the flat parameter vector."""
    y, c = forward(params, x, rounding)
    B = x.shape[0]
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / (B * N_OUT)
    g = {"W3": dy.T @ c["h2"], "b3": dy.sum(0)}
    dz2 = (dy @ c["W"]["W3"]) * gelu_grad(c["z2"])
    g["W2"], g["b2"] = dz2.T @ c["h1"], dz2.sum(0)
    dz1 = (dz2 @ c["W"]["W2"]) * gelu_grad(c["z1"])
    g["W1"], g["b1"] = dz1.T @ c["h0"], dz1.sum(0)
    return loss, pack(g)


class Adam:
    """Adam (beta1 0.9, beta2 0.999, eps 1e-8) with bias correction, as the engine applies it."""

    def __init__(self, params, b1=0.9, b2=0.999, eps=1e-8):
        self.p = np.asarray(params, dtype=np.float64).copy()
        self.m = np.zeros_like(self.p)
        self.v = np.zeros_like(self.p)
        self.b1, self.b2, self.eps, self.step = b1, b2, eps, 0

    def update(self, g, lr):
        self.step += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        c1, c2 = 1 - self.b1 ** self.step, 1 - self.b2 ** self.step
        self.p = self.p - lr * (self.m / c1) / (np.sqrt(self.v / c2) + self.eps)
        return self.p


def init_params(seed, scale=1.0):
    """He-style random initial parameters (float32), biases zero."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.standard_normal(shape) * scale * math.sqrt(2.0 / shape[1])
        else:
            d[name] = np.zeros(shape)
    return pack(d).astype(np.float32)
