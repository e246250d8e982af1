"""The _part rasteriser variant on the gfx950 HIP library.

Same surface as the reference package ``diff_surfel_rasterization_part``
(submodules/diff-surfel-rasterization_part/diff_surfel_rasterization_part/
__init__.py:21-231), the one renderer/gaussian_renderer_2d/__init__.py:86-96
calls:

* ``GaussianRasterizationSettings`` (the same fields as the plain variant);
* ``GaussianRasterizer.forward(means3D, means2D, opacities, semantics,
  shs=None, colors_precomp=None, scales=None, rotations=None,
  cov3D_precomp=None) -> (color, semantic, radii, allmap)`` with an
  8-channel allmap (depth, alpha, normal xyz, median depth, distortion,
  median weight) and ``markVisible``;
* ``rasterize_gaussians`` and ``_RasterizeGaussians`` (gradients in the
  reference's order, semantics included);
* ``rasterize_gaussians_fwd`` / ``_bwd`` with the signatures of the
  reference's ``_C`` (dsp/rasterize_points.h).

Every call runs on the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import torch
import torch.nn as nn

from . import _lib
from .rasterizer import _Arena, _f32, _settings, _surfels, mark_visible

MAX_SEMANTIC_TYPES = 16


def _semantics(semantics, P, dev):
    """(P, S) float32 semantics and S = semantics.size(1), as the reference
    reads it (dsp/rasterize_points.cu:77); None or a 1-D empty tensor is S = 0."""
    if semantics is None or (semantics.ndimension() != 2 and semantics.numel() == 0):
        return torch.empty((P, 0), dtype=torch.float32, device=dev), 0
    if semantics.ndimension() != 2 or semantics.size(0) != P:
        raise RuntimeError("semantics must have dimensions (num_points, semantic_types)")
    S = int(semantics.size(1))
    if S > MAX_SEMANTIC_TYPES:
        raise RuntimeError(f"at most {MAX_SEMANTIC_TYPES} semantic types (got {S})")
    return _f32(semantics, "semantics"), S


def rasterize_gaussians_fwd(bg, means3D, colors, opacity, semantics, scales, rotations, scale_modifier,
                            transMat_precomp, viewmatrix, projmatrix, tan_fovx, tan_fovy, image_height, image_width,
                            sh, degree, campos, prefiltered, debug, tile: int = 16):
    """RasterizeGaussiansCUDA of the _part variant (dsp/rasterize_points.cu:40-146).

    Returns (num_rendered, color (3,H,W), semantic (S,H,W), others (8,H,W),
    radii (P,) int32, geomBuffer, binningBuffer, imgBuffer)."""
    if means3D.ndimension() != 2 or means3D.size(1) != 3:
        raise RuntimeError("means3D must have dimensions (num_points, 3)")
    lib = _lib.load()
    means3D = _f32(means3D, "means3D")
    dev = means3D.device
    P = means3D.size(0)
    if scales.numel() and (scales.ndimension() != 2 or scales.size(1) != 2):
        raise RuntimeError("scales must have dimensions (num_points, 2)")
    if rotations.numel() and (rotations.ndimension() != 2 or rotations.size(1) != 4):
        raise RuntimeError("rotations must have dimensions (num_points, 4)")
    args = {"bg": bg, "colors": colors, "opacity": opacity, "scales": scales, "rotations": rotations,
            "transMat_precomp": transMat_precomp, "viewmatrix": viewmatrix, "projmatrix": projmatrix, "sh": sh,
            "campos": campos}
    a = {k: _f32(v, k) for k, v in args.items()}
    sem, S = _semantics(semantics, P, dev)
    H, W = int(image_height), int(image_width)
    color = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    semantic = torch.empty((S, H, W), dtype=torch.float32, device=dev)
    others = torch.empty((8, H, W), dtype=torch.float32, device=dev)
    radii = torch.empty((P,), dtype=torch.int32, device=dev)
    s = _settings(a["bg"], scale_modifier, a["viewmatrix"], a["projmatrix"], tan_fovx, tan_fovy, H, W, degree,
                  a["campos"], prefiltered, debug, tile)
    g = _surfels(P, means3D, a["sh"], a["colors"], a["opacity"], a["scales"], a["rotations"], a["transMat_precomp"])
    arena = _Arena(dev)
    nr = C.c_int(0)
    with torch.cuda.device(dev):
        rc = lib.sr_forward_part(C.byref(s), C.byref(g), S, _lib.ptr(sem), arena.fn, None, color.data_ptr(),
                                 _lib.ptr(semantic), others.data_ptr(), _lib.ptr(radii), C.byref(nr),
                                 _lib.stream_handle(dev))
    if rc == _lib.SR_ERR_ALLOC and arena.error is not None:
        raise arena.error
    _lib.check(rc)
    return nr.value, color, semantic, others, radii, arena.get(0), arena.get(1), arena.get(2)


def rasterize_gaussians_bwd(bg, means3D, radii, colors, semantics, scales, rotations, scale_modifier,
                            transMat_precomp, viewmatrix, projmatrix, tan_fovx, tan_fovy, dL_dout_color,
                            dL_dout_semantic, dL_dout_others, sh, degree, campos, geomBuffer, R, binningBuffer,
                            imageBuffer, debug, tile: int = 16):
    """RasterizeGaussiansBackwardCUDA of the _part variant (dsp/rasterize_points.cu:149-252).

    Returns (dL_dmeans2D, dL_dcolors, dL_dsemantics, dL_dopacity,
    dL_dmeans3D, dL_dtransMat, dL_dsh, dL_dscales, dL_drotations)."""
    lib = _lib.load()
    means3D = _f32(means3D, "means3D")
    dev = means3D.device
    P = means3D.size(0)
    args = {"bg": bg, "colors": colors, "scales": scales, "rotations": rotations,
            "transMat_precomp": transMat_precomp, "viewmatrix": viewmatrix, "projmatrix": projmatrix, "sh": sh,
            "campos": campos, "dL_dout_color": dL_dout_color, "dL_dout_others": dL_dout_others}
    a = {k: _f32(v, k) for k, v in args.items()}
    sem, S = _semantics(semantics, P, dev)
    H, W = a["dL_dout_color"].size(1), a["dL_dout_color"].size(2)
    M = int(a["sh"].size(1)) if a["sh"].numel() != 0 and a["sh"].size(0) != 0 else 0
    opts = dict(dtype=torch.float32, device=dev)
    out = {"dL_dmeans2D": torch.empty((P, 3), **opts), "dL_dcolors": torch.empty((P, 3), **opts),
           "dL_dsemantics": torch.empty((P, S), **opts), "dL_dopacity": torch.empty((P, 1), **opts),
           "dL_dmeans3D": torch.empty((P, 3), **opts), "dL_dtransMat": torch.empty((P, 9), **opts),
           "dL_dsh": torch.empty((P, M, 3), **opts), "dL_dscales": torch.empty((P, 2), **opts),
           "dL_drotations": torch.empty((P, 4), **opts)}
    if P != 0:
        dsem = _f32(dL_dout_semantic, "dL_dout_semantic") if S > 0 else None
        s = _settings(a["bg"], scale_modifier, a["viewmatrix"], a["projmatrix"], tan_fovx, tan_fovy, H, W, degree,
                      a["campos"], False, debug, tile)
        g = _surfels(P, means3D, a["sh"], a["colors"], torch.empty(0, **opts), a["scales"], a["rotations"],
                     a["transMat_precomp"])
        gr = _lib.SrGrads()
        for k in ("dL_dmeans2D", "dL_dcolors", "dL_dopacity", "dL_dmeans3D", "dL_dtransMat", "dL_dscales",
                  "dL_drotations"):
            setattr(gr, k, out[k].data_ptr())
        gr.dL_dsh = _lib.ptr(out["dL_dsh"])
        radii = radii.contiguous()
        with torch.cuda.device(dev):
            rc = lib.sr_backward_part(C.byref(s), C.byref(g), S, _lib.ptr(sem), int(R), radii.data_ptr(),
                                      _lib.ptr(geomBuffer), _lib.ptr(binningBuffer), _lib.ptr(imageBuffer),
                                      a["dL_dout_color"].data_ptr(), _lib.ptr(dsem), a["dL_dout_others"].data_ptr(),
                                      C.byref(gr), _lib.ptr(out["dL_dsemantics"]), _lib.stream_handle(dev))
        _lib.check(rc)
    return (out["dL_dmeans2D"], out["dL_dcolors"], out["dL_dsemantics"], out["dL_dopacity"], out["dL_dmeans3D"],
            out["dL_dtransMat"], out["dL_dsh"], out["dL_dscales"], out["dL_drotations"])


class _RasterizeGaussians(torch.autograd.Function):
    """Autograd wrapper, dsp/diff_surfel_rasterization_part/__init__.py:46-151."""

    @staticmethod
    def forward(ctx, means3D, means2D, sh, colors_precomp, opacities, semantics, scales, rotations, cov3Ds_precomp,
                raster_settings):
        tile = getattr(raster_settings, "tile", 16)
        num_rendered, color, semantic, depth, radii, geomBuffer, binningBuffer, imgBuffer = rasterize_gaussians_fwd(
            raster_settings.bg, means3D, colors_precomp, opacities, semantics, scales, rotations,
            raster_settings.scale_modifier, cov3Ds_precomp, raster_settings.viewmatrix, raster_settings.projmatrix,
            raster_settings.tanfovx, raster_settings.tanfovy, raster_settings.image_height,
            raster_settings.image_width, sh, raster_settings.sh_degree, raster_settings.campos,
            raster_settings.prefiltered, raster_settings.debug, tile)
        ctx.raster_settings = raster_settings
        ctx.num_rendered = num_rendered
        ctx.save_for_backward(colors_precomp, semantics, means3D, scales, rotations, cov3Ds_precomp, radii, sh,
                              geomBuffer, binningBuffer, imgBuffer)
        ctx.mark_non_differentiable(radii)
        ctx.set_materialize_grads(False)
        return color, semantic, radii, depth

    @staticmethod
    def backward(ctx, grad_out_color, grad_out_semantic, grad_radii, grad_depth):
        rs = ctx.raster_settings
        (colors_precomp, semantics, means3D, scales, rotations, cov3Ds_precomp, radii, sh, geomBuffer, binningBuffer,
         imgBuffer) = ctx.saved_tensors
        H, W = rs.image_height, rs.image_width
        dev = means3D.device
        S = semantics.size(1) if semantics.ndimension() == 2 else 0
        if grad_out_color is None:
            grad_out_color = torch.zeros((3, H, W), device=dev)
        if grad_out_semantic is None:
            grad_out_semantic = torch.zeros((S, H, W), device=dev)
        if grad_depth is None:
            grad_depth = torch.zeros((8, H, W), device=dev)
        (grad_means2D, grad_colors_precomp, grad_semantics, grad_opacities, grad_means3D, grad_cov3Ds_precomp, grad_sh,
         grad_scales, grad_rotations) = rasterize_gaussians_bwd(
            rs.bg, means3D, radii, colors_precomp, semantics, scales, rotations, rs.scale_modifier, cov3Ds_precomp,
            rs.viewmatrix, rs.projmatrix, rs.tanfovx, rs.tanfovy, grad_out_color, grad_out_semantic, grad_depth, sh,
            rs.sh_degree, rs.campos, geomBuffer, ctx.num_rendered, binningBuffer, imgBuffer, rs.debug,
            getattr(rs, "tile", 16))
        if semantics.ndimension() != 2:
            grad_semantics = None
        return (grad_means3D, grad_means2D, grad_sh, grad_colors_precomp, grad_opacities, grad_semantics, grad_scales,
                grad_rotations, grad_cov3Ds_precomp, None)


def rasterize_gaussians(means3D, means2D, sh, colors_precomp, opacities, semantics, scales, rotations,
                        cov3Ds_precomp, raster_settings):
    return _RasterizeGaussians.apply(means3D, means2D, sh, colors_precomp, opacities, semantics, scales, rotations,
                                     cov3Ds_precomp, raster_settings)


class GaussianRasterizationSettings(NamedTuple):
    image_height: int
    image_width: int
    tanfovx: float
    tanfovy: float
    bg: torch.Tensor
    scale_modifier: float
    viewmatrix: torch.Tensor
    projmatrix: torch.Tensor
    sh_degree: int
    campos: torch.Tensor
    prefiltered: bool
    debug: bool
    tile: int = 16


class GaussianRasterizer(nn.Module):
    """dsp/diff_surfel_rasterization_part/__init__.py:172-231."""

    def __init__(self, raster_settings):
        super().__init__()
        self.raster_settings = raster_settings

    def markVisible(self, positions):
        with torch.no_grad():
            rs = self.raster_settings
            visible = mark_visible(positions, rs.viewmatrix, rs.projmatrix)
        return visible

    def forward(self, means3D, means2D, opacities, semantics, shs=None, colors_precomp=None, scales=None,
                rotations=None, cov3D_precomp=None):
        rs = self.raster_settings
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise Exception("Please provide excatly one of either SHs or precomputed colors!")
        if ((scales is None or rotations is None) and cov3D_precomp is None) or (
                (scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise Exception("Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!")
        empty = torch.empty(0, dtype=torch.float32, device=means3D.device)
        shs = empty if shs is None else shs
        colors_precomp = empty if colors_precomp is None else colors_precomp
        scales = empty if scales is None else scales
        rotations = empty if rotations is None else rotations
        cov3D_precomp = empty if cov3D_precomp is None else cov3D_precomp
        return rasterize_gaussians(means3D, means2D, shs, colors_precomp, opacities, semantics, scales, rotations,
                                   cov3D_precomp, rs)
