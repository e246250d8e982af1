"""distCUDA2 of simple-knn on the gfx950 HIP library.

``distCUDA2(points)`` (submodules/simple-knn/spatial.cu:16-27): for a (P, 3)
float32 CUDA tensor, the (P,) float32 mean of each point's three smallest
squared distances to the other points, as SimpleKNN::knn
(simple_knn.cu:185-220) returns it; the initial surfel scales come from it
(scene/gaussian_model.py:237).  The search is exact, so the result depends on
the point set alone.  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def distCUDA2(points: torch.Tensor) -> torch.Tensor:
    if not isinstance(points, torch.Tensor) or points.device.type != "cuda":
        raise RuntimeError("points must be a CUDA tensor")
    if points.dim() != 2 or points.size(1) != 3:
        raise RuntimeError("points must have dimensions (num_points, 3)")
    lib = _lib.load()
    pts = points.float().contiguous()
    P = int(pts.size(0))
    means = torch.zeros((P,), dtype=torch.float32, device=pts.device)
    if P == 0:
        return means
    ws = torch.empty(int(lib.sk_dist2_workspace_bytes(P)), dtype=torch.uint8, device=pts.device)
    with torch.cuda.device(pts.device):
        rc = lib.sk_dist2(P, pts.data_ptr(), means.data_ptr(), ws.data_ptr(), C.c_size_t(ws.numel()),
                          _lib.stream_handle(pts.device))
    _lib.check(rc)
    return means
