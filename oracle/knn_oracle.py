"""CPU oracle of simple-knn's distCUDA2 (submodules/simple-knn/spatial.cu:
16-27, simple_knn.cu:127-183).

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

For every point q: the three smallest FP32 squared distances to the other
points, each evaluated as the reference's updateKBest does (simple_knn.cu:
129-133): d = c - q per axis in FP32, then (dx dx + dy dy) + dz dz with every
product and sum rounded to FP32; kept in ascending order with FLT_MAX
standing in for missing neighbours (:150,162), and the mean
((b0 + b1) + b2) / 3 in FP32 (:176).  The reference's box search is exact
(a box is skipped only when its lower bound exceeds a value the third
neighbour cannot exceed), so this is its output for any order of the points.

* ``dist2_brute``: all pairs, O(P^2), for small P.
* ``dist2_kdtree``: scipy's cKDTree proposes the 16 nearest in FP64; the FP32
  distances of those candidates are recomputed and the three smallest kept.
  The true three nearest in FP32 are among the 16 nearest in FP64 unless
  more than 13 points lie within FP32 rounding of the third distance.

Pinned (tests/test_knn.py): brute force against the KD-tree path, a regular
grid whose answer is the spacing squared, duplicates (0), and P = 1, 2, 3.
"""
from __future__ import annotations

import numpy as np

FLT_MAX = np.float32(np.finfo(np.float32).max)


def _sqdist32(q: np.ndarray, c: np.ndarray) -> np.ndarray:
    d = (c.astype(np.float32) - q.astype(np.float32)).astype(np.float32)
    return ((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]).astype(np.float32)


def _mean3(best: np.ndarray) -> np.ndarray:
    """best (N, k >= 3) FP32 candidate distances, FLT_MAX padded."""
    b = np.sort(best, axis=1)[:, :3].astype(np.float32)
    with np.errstate(over="ignore"):
        s = (b[:, 0] + b[:, 1]).astype(np.float32)
        s = (s + b[:, 2]).astype(np.float32)
    return (s / np.float32(3.0)).astype(np.float32)


def dist2_brute(points: np.ndarray) -> np.ndarray:
    p = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    P = p.shape[0]
    if P == 0:
        return np.zeros(0, np.float32)
    out = np.empty(P, np.float32)
    for i0 in range(0, P, 512):
        q = p[i0:i0 + 512]
        d = _sqdist32(q[:, None, :], p[None, :, :])  # (n, P)
        d[np.arange(q.shape[0]), np.arange(i0, i0 + q.shape[0])] = FLT_MAX  # not itself
        pad = np.full((q.shape[0], 3), FLT_MAX, np.float32)
        out[i0:i0 + q.shape[0]] = _mean3(np.concatenate([d, pad], axis=1))
    return out


def dist2_kdtree(points: np.ndarray, k: int = 16) -> np.ndarray:
    from scipy.spatial import cKDTree

    p = np.asarray(points, dtype=np.float32).reshape(-1, 3)
    P = p.shape[0]
    if P <= k:
        return dist2_brute(p)
    tree = cKDTree(p.astype(np.float64))
    _, idx = tree.query(p.astype(np.float64), k=k + 1)
    cand = _sqdist32(p[:, None, :], p[idx])
    cand[idx == np.arange(P)[:, None]] = FLT_MAX  # the point itself (duplicates keep their 0)
    return _mean3(cand)
