"""CPU FP64 oracle of the scatter-surrogate MLP (12 -> 256 -> 256 -> 256 -> 256 -> 1, SiLU).

TEST INFRASTRUCTURE ONLY (tests/, bench.py's CPU leg).  The surrogate has no
counterpart in the reference; this module states the network the GPU kernels
compute, in float64, and optionally rounds to bfloat16 at the points where the
GPU feeds the matrix cores (weights, layer inputs, back-propagated dZ), so
that GPU-vs-oracle differences reduce to FP32-vs-FP64 accumulation.

Definition (the GPU kernels follow it):
    a0 = bf16(x padded to 32 with zeros)
    z_l = bf16(W_l) @ a_{l-1} + b_l          l = 1..4  (FP32 accumulate on the GPU)
    h_l = silu(z_l),  a_l = bf16(h_l)
    y   = w5 . h4 + b5                        (h4 unrounded)
Backward (straight-through for the bf16 roundings):
    dh4 = dy w5;  dz_l = bf16(dh_l * silu'(z_l));  dh_{l-1} = bf16(W_l)^T dz_l
    dW_l = dz_l a_{l-1}^T;  db_l = sum dz_l;  dw5 = dy h4;  db5 = sum dy

Flat parameter layout (float32 on the GPU), offsets in floats:
    W1 [256][32] @0, b1 @8192, W2 [256][256] @8448, b2 @73984, W3 @74240,
    b3 @139776, W4 @140032, b4 @205568, w5 [256] @205824, b5 @206080;
    206144 floats in all (tail zero).  Columns 12..31 of W1 stay zero.
"""
from __future__ import annotations

import numpy as np

N_IN = 12
K1 = 32
H = 256
OFF = {"W1": 0, "b1": 8192, "W2": 8448, "b2": 73984, "W3": 74240, "b3": 139776, "W4": 140032, "b4": 205568,
       "w5": 205824, "b5": 206080}
N_PARAMS = 206144


def bf16(a: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (round to nearest even), returned as float64."""
    f = np.ascontiguousarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p, np.float64)
    return {"W1": p[0:8192].reshape(H, K1), "b1": p[8192:8448], "W2": p[8448:73984].reshape(H, H),
            "b2": p[73984:74240], "W3": p[74240:139776].reshape(H, H), "b3": p[139776:140032],
            "W4": p[140032:205568].reshape(H, H), "b4": p[205568:205824], "w5": p[205824:206080],
            "b5": p[206080:206081]}


def pack(d: dict) -> np.ndarray:
    p = np.zeros(N_PARAMS, np.float64)
    for k, v in d.items():
        p[OFF[k]:OFF[k] + np.size(v)] = np.ravel(v)
    return p


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, x, emulate_bf16: bool = True):
    """x: (B, 12) -> y (B,), plus the cache for backward."""
    P = unpack(params)
    rnd = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    B = x.shape[0]
    a = np.zeros((B, K1))
    a[:, :N_IN] = x
    a = rnd(a)
    cache = {"a0": a}
    h = None
    for l in range(1, 5):
        W = rnd(P[f"W{l}"])
        z = a @ W.T + P[f"b{l}"]
        h = silu(z)
        cache[f"z{l}"] = z
        cache[f"h{l}"] = h
        a = rnd(h)
        cache[f"a{l}"] = a
    y = h @ P["w5"] + P["b5"][0]
    return y, cache


def backward(params, cache, dy, emulate_bf16: bool = True) -> np.ndarray:
    """Gradient of sum(dy * y) w.r.t. the flat parameters."""
    P = unpack(params)
    rnd = bf16 if emulate_bf16 else (lambda a: np.asarray(a, np.float64))
    g = {}
    g["w5"] = dy @ cache["h4"]
    g["b5"] = np.array([dy.sum()])
    dh = np.outer(dy, P["w5"])
    for l in range(4, 0, -1):
        dz = rnd(dh * dsilu(cache[f"z{l}"]))
        g[f"W{l}"] = dz.T @ cache[f"a{l - 1}"]
        g[f"b{l}"] = dz.sum(axis=0)
        if l > 1:
            dh = dz @ rnd(P[f"W{l}"])
    g["W1"][:, N_IN:] = 0.0
    return pack(g)


def init(seed: int) -> np.ndarray:
    """Reference initialisation scheme (uniform +-sqrt(6/fan_in) scaled for SiLU); the GPU uses its own RNG."""
    rng = np.random.default_rng(seed)
    d = {}
    for l, fan in ((1, N_IN), (2, H), (3, H), (4, H)):
        lim = np.sqrt(6.0 / fan)
        W = rng.uniform(-lim, lim, (H, K1 if l == 1 else H))
        if l == 1:
            W[:, N_IN:] = 0.0
        d[f"W{l}"] = W
        d[f"b{l}"] = np.zeros(H)
    d["w5"] = rng.uniform(-1, 1, H) / np.sqrt(H)
    d["b5"] = np.zeros(1)
    return pack(d)


def mse_loss_grad(y, t, w=None):
    """Loss = sum(w (y - t)^2) / sum(w) (w = 1 by default) and dL/dy."""
    w = np.ones_like(y) if w is None else np.asarray(w, np.float64)
    sw = w.sum()
    r = y - t
    return float((w * r * r).sum() / sw), 2.0 * w * r / sw
