"""Superquadric block layer on the GPU: block parameters -> rasteriser inputs.

``generate_surfels(params, occupancy, grid)`` returns (means3D, scales,
rotations, opacities) exactly as BlockGaussianModel exposes them to the
renderer (get_xyz / get_scaling / get_rotation / get_opacity,
games/block_mesh_splatting/scene/block_gaussian_model.py:103-121,178-256),
differentiable with respect to the 12 block parameters
(sq_s(3), sq_r(4), sq_t(3), sq_eps(2)) and the occupancy logits sq_occ.
The work runs in the HIP kernels sb_generate_fwd / sb_generate_bwd.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _lib


class SbBlocks(C.Structure):
    _fields_ = [
        ("B", C.c_int), ("n_eta", C.c_int), ("n_omega", C.c_int), ("K", C.c_int),
        ("eta", C.c_void_p), ("omega", C.c_void_p), ("bary", C.c_void_p),
        ("ratio", C.c_float), ("scale_min", C.c_float), ("surfel_scale", C.c_float),
        ("params", C.c_void_p), ("occupancy", C.c_void_p),
    ]


@dataclass
class BlockGrid:
    """(eta, omega) sampling of every block: n_eta x n_omega faces, K surfels each."""

    eta: torch.Tensor           # n_eta + 1
    omega: torch.Tensor         # n_omega
    bary: torch.Tensor | None   # K x 3, None = one surfel at the centroid
    ratio: float = 0.25         # ratio_block_scene
    scale_min: float = 0.2      # scale_block_min
    surfel_scale: float = 1.0   # _scale

    @property
    def n_eta(self) -> int:
        return int(self.eta.numel()) - 1

    @property
    def n_omega(self) -> int:
        return int(self.omega.numel())

    @property
    def K(self) -> int:
        return 1 if self.bary is None else int(self.bary.shape[0])

    def surfels_per_block(self) -> int:
        return self.n_eta * self.n_omega * self.K

    def to(self, device) -> "BlockGrid":
        return BlockGrid(self.eta.to(device), self.omega.to(device),
                         None if self.bary is None else self.bary.to(device), self.ratio, self.scale_min,
                         self.surfel_scale)


def make_grid(n_eta: int = 64, n_omega: int = 64, device="cuda", **kw) -> BlockGrid:
    """Cell-centred latitudes (poles excluded) and half-cell-offset longitudes."""
    eta = torch.tensor([-math.pi / 2 + (i + 0.5) * math.pi / (n_eta + 1) for i in range(n_eta + 1)],
                       dtype=torch.float32, device=device)
    omega = torch.tensor([-math.pi + (j + 0.5) * 2 * math.pi / n_omega for j in range(n_omega)],
                         dtype=torch.float32, device=device)
    return BlockGrid(eta, omega, None, **kw)


def _desc(grid: BlockGrid, params: torch.Tensor, occ: torch.Tensor) -> SbBlocks:
    d = SbBlocks()
    d.B = int(params.shape[0])
    d.n_eta, d.n_omega, d.K = grid.n_eta, grid.n_omega, grid.K
    d.eta, d.omega = grid.eta.data_ptr(), grid.omega.data_ptr()
    d.bary = None if grid.bary is None else grid.bary.data_ptr()
    d.ratio, d.scale_min, d.surfel_scale = grid.ratio, grid.scale_min, grid.surfel_scale
    d.params, d.occupancy = params.data_ptr(), occ.data_ptr()
    return d


def _check_inputs(params, occ, grid):
    for name, t in (("params", params), ("occupancy", occ), ("eta", grid.eta), ("omega", grid.omega)):
        if t.device.type != "cuda" or t.dtype != torch.float32:
            raise RuntimeError(f"{name} must be a float32 CUDA tensor")
    if params.dim() != 2 or params.shape[1] != 12:
        raise RuntimeError("params must be B x 12 (sq_s, sq_r, sq_t, sq_eps)")
    if occ.numel() != params.shape[0]:
        raise RuntimeError("occupancy must hold one logit per block")


class _GenerateSurfels(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, occ, grid):
        lib = _lib.load()
        params = params.contiguous()
        occ = occ.contiguous().reshape(-1)
        _check_inputs(params, occ, grid)
        P = params.shape[0] * grid.surfels_per_block()
        dev = params.device
        means = torch.empty((P, 3), dtype=torch.float32, device=dev)
        scales = torch.empty((P, 2), dtype=torch.float32, device=dev)
        rots = torch.empty((P, 4), dtype=torch.float32, device=dev)
        opac = torch.empty((P, 1), dtype=torch.float32, device=dev)
        d = _desc(grid, params, occ)
        with torch.cuda.device(dev):
            _lib.check(lib.sb_generate_fwd(C.byref(d), means.data_ptr(), scales.data_ptr(), rots.data_ptr(),
                                           opac.data_ptr(), _lib.stream_handle(dev)))
        ctx.grid = grid
        ctx.save_for_backward(params, occ)
        return means, scales, rots, opac

    @staticmethod
    def backward(ctx, g_means, g_scales, g_rots, g_opac):
        lib = _lib.load()
        params, occ = ctx.saved_tensors
        grid = ctx.grid
        dev = params.device
        P = params.shape[0] * grid.surfels_per_block()
        z = lambda shape: torch.zeros(shape, dtype=torch.float32, device=dev)  # noqa: E731
        g_means = z((P, 3)) if g_means is None else g_means.contiguous()
        g_scales = z((P, 2)) if g_scales is None else g_scales.contiguous()
        g_rots = z((P, 4)) if g_rots is None else g_rots.contiguous()
        g_opac = z((P, 1)) if g_opac is None else g_opac.contiguous()
        d_params = torch.empty_like(params)
        d_occ = torch.empty_like(occ)
        d = _desc(grid, params, occ)
        ws = torch.empty(int(lib.sb_bwd_workspace_bytes(C.byref(d))), dtype=torch.uint8, device=dev)
        with torch.cuda.device(dev):
            _lib.check(lib.sb_generate_bwd(C.byref(d), g_means.data_ptr(), g_scales.data_ptr(), g_rots.data_ptr(),
                                           g_opac.data_ptr(), d_params.data_ptr(), d_occ.data_ptr(),
                                           ws.data_ptr(), ws.numel(), _lib.stream_handle(dev)))
        return d_params, d_occ, None


def generate_surfels(params: torch.Tensor, occupancy: torch.Tensor, grid: BlockGrid):
    """(means3D P x 3, scales P x 2, rotations P x 4, opacities P x 1)."""
    return _GenerateSurfels.apply(params, occupancy, grid)


def pack_params(sq_s, sq_r, sq_t, sq_eps) -> torch.Tensor:
    """B x 12 parameter block from the reference's separate parameters."""
    return torch.cat([sq_s, sq_r, sq_t, sq_eps], dim=1).contiguous()
