"""CPU oracle of the learned proposal's point-set MLP (test infrastructure only:
imported by tests/, never by the package).

Two modes:
* ``bf16=True``: the GPU path's arithmetic -- weights W2..W5 and the inputs of
  every matrix product rounded to bf16 (nearest even), products accumulated
  in FP32 (numpy float32), layer 1 in FP32 with the operation order
  ((w0 x + w1 y) + w2 z) + b;
* ``bf16=False``: the same network in float64 with no rounding, whose
  backward is pinned by central differences.

The pool is a max over each scene's points with the first argmax taking the
gradient, the convention of the GPU kernel.  There is no reference
counterpart for this network (the reference initialises blocks by k-means).
"""
from __future__ import annotations

import numpy as np

H1, H2, H3, H4, OUT = 64, 128, 256, 128, 192


def bf16_round(x: np.ndarray) -> np.ndarray:
    """FP32 -> bf16 -> FP32, round to nearest even (NaN kept)."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    return (bf16_round(x).view(np.uint32) >> 16).astype(np.uint16)


class Oracle:
    def __init__(self, params: dict, bf16: bool = True):
        self.bf16 = bf16
        dt = np.float32 if bf16 else np.float64
        self.p = {k: np.asarray(v, dtype=dt) for k, v in params.items()}

    def _r(self, x):
        return bf16_round(x) if self.bf16 else x

    def _w(self, k):
        return self._r(self.p[k]) if k in ("W2", "W3", "W4", "W5") else self.p[k]

    def layer1(self, X):
        W1, b1 = self.p["W1"], self.p["b1"]
        h = ((X[:, 0:1] * W1[:, 0] + X[:, 1:2] * W1[:, 1]) + X[:, 2:3] * W1[:, 2]) + b1
        return np.maximum(h, 0)

    def shared(self, X):
        """h1, h2, h3 (post-ReLU) for points X (R x 3)."""
        h1 = self.layer1(X)
        h2 = np.maximum(self._r(h1) @ self._w("W2").T + self.p["b2"], 0)
        h3 = np.maximum(self._r(h2) @ self._w("W3").T + self.p["b3"], 0)
        return h1, h2, h3

    def forward(self, pts):
        """out (B x 192) and the cache; pts B x N x 3."""
        pts = np.asarray(pts, dtype=np.float32 if self.bf16 else np.float64)
        B, N, _ = pts.shape
        g = np.zeros((B, H3), dtype=pts.dtype)
        idx = np.zeros((B, H3), dtype=np.int64)
        for b in range(B):
            _, _, h3 = self.shared(pts[b])
            idx[b] = np.argmax(h3, axis=0)  # first maximum
            g[b] = h3[idx[b], np.arange(H3)]
        z1 = np.maximum(self._r(g) @ self._w("W4").T + self.p["b4"], 0)
        out = self._r(z1) @ self._w("W5").T + self.p["b5"]
        return out, dict(pts=pts, g=g, idx=idx, z1=z1)

    def backward(self, dout, c):
        """Parameter gradients for dL/d(out) (B x 192)."""
        dout = np.asarray(dout, dtype=c["g"].dtype)
        r = self._r
        G = {}
        G["W5"] = r(dout).T @ r(c["z1"])
        G["b5"] = dout.sum(0)
        dz1 = (r(dout) @ self._w("W5")) * (c["z1"] > 0)
        G["W4"] = r(dz1).T @ r(c["g"])
        G["b4"] = dz1.sum(0)
        dg = (r(dz1) @ self._w("W4")) * (c["g"] > 0)
        B = dg.shape[0]
        pts = c["pts"]
        X = np.concatenate([pts[b][c["idx"][b]] for b in range(B)])  # (B*256) x 3, row b*256 + ch
        h1, h2, _ = self.shared(X)
        d3 = np.zeros((X.shape[0], H3), dtype=dg.dtype)
        d3[np.arange(X.shape[0]), np.tile(np.arange(H3), B)] = dg.reshape(-1)
        G["W3"] = r(d3).T @ r(h2)
        G["b3"] = dg.sum(0)
        dh2 = (r(d3) @ self._w("W3")) * (h2 > 0)
        G["W2"] = r(dh2).T @ r(h1)
        G["b2"] = dh2.sum(0)
        dh1 = (r(dh2) @ self._w("W2")) * (h1 > 0)
        G["W1"] = dh1.T @ X
        G["b1"] = dh1.sum(0)
        return G


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step in FP32 (the kernel's formula)."""
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    m = np.float32(b1) * m + np.float32(1 - b1) * g
    v = np.float32(b2) * v + np.float32(1 - b2) * (g * g)
    c1 = np.float32(1 - b1 ** step)
    c2 = np.float32(1 - b2 ** step)
    p = p - np.float32(lr) * (m / c1) / (np.sqrt(v / c2) + np.float32(eps))
    return p, m, v
