"""CPU oracle for the warm-start MLP (test infrastructure only).

The MLP is not part of the reference code: it is the network of the learned
warm start, 6 -> 256 x 4 (SiLU) -> 4, with a flat FP32 parameter vector laid
out as W1 (6x256, row = input), b1, W2..W4 (256x256, row = input unit), b2..b4,
W5 (256x4), b5.  Loss = mean over rows and the 4 outputs of (y - t)^2.

`forward` / `loss_grad` compute in FP64.  With bf16=True they round to BF16 at
the same points as the device kernels do: the hidden activations H1..H4, the
weights W2..W4 and the hidden-layer output gradients dZ2..dZ4 that feed the
matrix-core products; everything else stays unrounded.  The plain FP64 path is
pinned by central finite differences (tests/test_mlp_oracle.py).  This module
is imported only by tests, never by the library.
"""
from __future__ import annotations

import numpy as np

K_IN, K_H, K_OUT = 6, 256, 4
OFF = {}
_o = 0
for _name, _n in (("W1", K_IN * K_H), ("b1", K_H), ("W2", K_H * K_H), ("b2", K_H), ("W3", K_H * K_H), ("b3", K_H),
                  ("W4", K_H * K_H), ("b4", K_H), ("W5", K_H * K_OUT), ("b5", K_OUT)):
    OFF[_name] = (_o, _n)
    _o += _n
N_PARAMS = _o  # 200196
SHAPES = {"W1": (K_IN, K_H), "W2": (K_H, K_H), "W3": (K_H, K_H), "W4": (K_H, K_H), "W5": (K_H, K_OUT)}


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest, ties to even), returned as float64."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray):
    out = {}
    for k, (o, n) in OFF.items():
        v = np.asarray(p[o:o + n], np.float64)
        out[k] = v.reshape(SHAPES[k]) if k in SHAPES else v
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


def forward(p: np.ndarray, x: np.ndarray, bf16_points: bool = False, keep: bool = False):
    P = unpack(p)
    rnd = bf16 if bf16_points else (lambda a: np.asarray(a, np.float64))
    x = np.asarray(x, np.float64)
    zs, hs = [], []
    z = x @ P["W1"] + P["b1"]
    zs.append(z)
    hs.append(rnd(_silu(z)))
    for l in (2, 3, 4):
        z = hs[-1] @ rnd(P[f"W{l}"]) + P[f"b{l}"]
        zs.append(z)
        hs.append(rnd(_silu(z)))
    y = hs[-1] @ P["W5"] + P["b5"]
    if keep:
        return y, zs, hs, P
    return y


def loss_grad(p: np.ndarray, x: np.ndarray, t: np.ndarray, bf16_points: bool = False):
    rnd = bf16 if bf16_points else (lambda a: np.asarray(a, np.float64))
    y, zs, hs, P = forward(p, x, bf16_points, keep=True)
    t = np.asarray(t, np.float64)
    B = x.shape[0]
    d = y - t
    loss = float((d * d).sum() / (B * K_OUT))
    dy = 2.0 * d / (B * K_OUT)
    g = {"W5": hs[3].T @ dy, "b5": dy.sum(0)}
    dh = dy @ P["W5"].T
    for l in (4, 3, 2):
        dz = dh * _dsilu(zs[l - 1])
        dzr = rnd(dz)
        g[f"W{l}"] = hs[l - 2].T @ dzr
        g[f"b{l}"] = dz.sum(0)
        dh = dzr @ rnd(P[f"W{l}"]).T
    dz = dh * _dsilu(zs[0])
    g["W1"] = np.asarray(x, np.float64).T @ dz
    g["b1"] = dz.sum(0)
    flat = np.zeros(N_PARAMS)
    for k, (o, n) in OFF.items():
        flat[o:o + n] = g[k].ravel()
    return loss, flat
