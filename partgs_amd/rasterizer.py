"""Drop-in surfel rasteriser API on the gfx950 HIP library.

Same surface as the reference package ``diff_surfel_rasterization``
(submodules/diff-surfel-rasterization/diff_surfel_rasterization/__init__.py):

* ``GaussianRasterizationSettings`` (NamedTuple, __init__.py:158-170), plus an
  optional ``tile`` field (screen tile edge, 16 = the reference);
* ``GaussianRasterizer(nn.Module)`` with ``forward(means3D, means2D,
  opacities, shs, colors_precomp, scales, rotations, cov3D_precomp) ->
  (color, radii, allmap)`` and ``markVisible`` (__init__.py:172-222);
* ``rasterize_gaussians`` and the autograd function ``_RasterizeGaussians``
  (__init__.py:21-156), returning gradients in the reference's order;
* the low-level functions ``rasterize_gaussians_fwd`` / ``_bwd`` /
  ``mark_visible`` with the signatures of the reference's ``_C`` module
  (rasterize_points.h:18-68).

Every call runs on the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from typing import NamedTuple

import torch
import torch.nn as nn

from . import _lib


def _f32(t: torch.Tensor, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a tensor")
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if t.dtype != torch.float32:
        raise RuntimeError(f"{name} must be float32 (got {t.dtype})")
    return t.contiguous()


class _Arena:
    """Holds the device buffers the library asks for during one forward."""

    def __init__(self, device):
        self.device = device
        self.bufs = {0: None, 1: None, 2: None}
        self.error = None

        def _alloc(_ctx, nbytes, which):
            try:
                t = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
                self.bufs[int(which)] = t
                return t.data_ptr()
            except Exception as ex:  # reported through SR_ERR_ALLOC
                self.error = ex
                return None

        self.fn = _lib.ALLOC_FN(_alloc)

    def get(self, which: int) -> torch.Tensor:
        t = self.bufs[which]
        if t is None:
            t = torch.empty(0, dtype=torch.uint8, device=self.device)
        return t


def _settings(bg, scale_modifier, viewmatrix, projmatrix, tan_fovx, tan_fovy, image_height, image_width,
              degree, campos, prefiltered, debug, tile, pair_store=0, capacity=0, pair_capacity=0, slot_capacity=0):
    s = _lib.SrSettings()
    s.image_height = int(image_height)
    s.image_width = int(image_width)
    s.tanfovx = float(tan_fovx)
    s.tanfovy = float(tan_fovy)
    s.bg = _lib.ptr(bg)
    s.scale_modifier = float(scale_modifier)
    s.viewmatrix = _lib.ptr(viewmatrix)
    s.projmatrix = _lib.ptr(projmatrix)
    s.sh_degree = int(degree)
    s.campos = _lib.ptr(campos)
    s.prefiltered = 1 if prefiltered else 0
    s.debug = 1 if debug else 0
    s.tile = int(tile)
    s.pair_store = int(pair_store)
    s.capacity = int(capacity)
    s.pair_capacity = int(pair_capacity)
    s.slot_capacity = int(slot_capacity)
    return s


def _surfels(P, means3D, sh, colors, opacity, scales, rotations, transMat_precomp):
    g = _lib.SrSurfels()
    g.P = P
    g.M = int(sh.size(1)) if sh.numel() != 0 and sh.size(0) != 0 else 0
    g.means3D = _lib.ptr(means3D)
    g.shs = _lib.ptr(sh)
    g.colors_precomp = _lib.ptr(colors)
    g.opacities = _lib.ptr(opacity)
    g.scales = _lib.ptr(scales)
    g.rotations = _lib.ptr(rotations)
    g.transMat_precomp = _lib.ptr(transMat_precomp)
    return g


def rasterize_gaussians_fwd(bg, means3D, colors, opacity, scales, rotations, scale_modifier, transMat_precomp,
                            viewmatrix, projmatrix, tan_fovx, tan_fovy, image_height, image_width, sh, degree,
                            campos, prefiltered, debug, tile: int = 16, pair_store: int = 0, capacity: int = 0,
                            pair_capacity: int = 0, slot_capacity: int = 0):
    """RasterizeGaussiansCUDA (rasterize_points.cu:40-134).

    Returns (num_rendered, color (3,H,W), others (7,H,W), radii (P,) int32,
    geomBuffer, binningBuffer, imgBuffer).  capacity > 0 sizes the binning
    buffer for that many tile instances up front, so the stream is not
    synchronised mid-forward; more instances raise CapacityError (whose
    `needed` is the count).  pair_capacity > 0 sizes the accepted-pair store
    in pairs (0: pair_store per instance); pairs beyond it are recomputed,
    with the same results.  slot_capacity > 0 (with capacity > 0) sizes the
    partial-gradient slots (0: one per instance and quadrant); more raise
    CapacityError with slots set.  The backward must be given the same
    capacity and pair_capacity.
    """
    if means3D.ndimension() != 2 or means3D.size(1) != 3:
        raise RuntimeError("means3D must have dimensions (num_points, 3)")
    lib = _lib.load()
    means3D = _f32(means3D, "means3D")
    dev = means3D.device
    args = {
        "bg": bg, "colors": colors, "opacity": opacity, "scales": scales, "rotations": rotations,
        "transMat_precomp": transMat_precomp, "viewmatrix": viewmatrix, "projmatrix": projmatrix, "sh": sh,
        "campos": campos,
    }
    a = {k: _f32(v, k) for k, v in args.items()}
    P = means3D.size(0)
    H, W = int(image_height), int(image_width)
    color = torch.empty((3, H, W), dtype=torch.float32, device=dev)
    others = torch.empty((7, H, W), dtype=torch.float32, device=dev)
    radii = torch.empty((P,), dtype=torch.int32, device=dev)  # the preprocess writes every entry
    s = _settings(a["bg"], scale_modifier, a["viewmatrix"], a["projmatrix"], tan_fovx, tan_fovy, H, W, degree,
                  a["campos"], prefiltered, debug, tile, pair_store, capacity, pair_capacity, slot_capacity)
    g = _surfels(P, means3D, a["sh"], a["colors"], a["opacity"], a["scales"], a["rotations"], a["transMat_precomp"])
    arena = _Arena(dev)
    nr = C.c_int(0)
    with torch.cuda.device(dev):
        rc = lib.sr_forward(C.byref(s), C.byref(g), arena.fn, None, color.data_ptr(), others.data_ptr(),
                            _lib.ptr(radii), C.byref(nr), _lib.stream_handle(dev))
    if rc == _lib.SR_ERR_ALLOC and arena.error is not None:
        raise arena.error
    if rc == _lib.SR_ERR_CAPACITY:
        raise _lib.CapacityError(nr.value, capacity, slots=nr.value <= capacity)
    _lib.check(rc)
    return nr.value, color, others, radii, arena.get(0), arena.get(1), arena.get(2)


def rasterize_gaussians_bwd(bg, means3D, radii, colors, scales, rotations, scale_modifier, transMat_precomp,
                            viewmatrix, projmatrix, tan_fovx, tan_fovy, dL_dout_color, dL_dout_others, sh, degree,
                            campos, geomBuffer, R, binningBuffer, imageBuffer, debug, tile: int = 16,
                            pair_store: int = 0, capacity: int = 0, pair_capacity: int = 0, slot_capacity: int = 0):
    """RasterizeGaussiansBackwardCUDA (rasterize_points.cu:137-233).

    Returns (dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dtransMat,
    dL_dsh, dL_dscales, dL_drotations).
    """
    lib = _lib.load()
    means3D = _f32(means3D, "means3D")
    dev = means3D.device
    args = {
        "bg": bg, "colors": colors, "scales": scales, "rotations": rotations,
        "transMat_precomp": transMat_precomp, "viewmatrix": viewmatrix, "projmatrix": projmatrix, "sh": sh,
        "campos": campos, "dL_dout_color": dL_dout_color, "dL_dout_others": dL_dout_others,
    }
    a = {k: _f32(v, k) for k, v in args.items()}
    P = means3D.size(0)
    H, W = a["dL_dout_color"].size(1), a["dL_dout_color"].size(2)
    M = int(a["sh"].size(1)) if a["sh"].numel() != 0 and a["sh"].size(0) != 0 else 0
    opts = dict(dtype=torch.float32, device=dev)
    # the library clears every gradient array before accumulating into it
    dL_dmeans3D = torch.empty((P, 3), **opts)
    dL_dmeans2D = torch.empty((P, 3), **opts)
    dL_dcolors = torch.empty((P, 3), **opts)
    dL_dopacity = torch.empty((P, 1), **opts)
    dL_dtransMat = torch.empty((P, 9), **opts)
    dL_dsh = torch.empty((P, M, 3), **opts)
    dL_dscales = torch.empty((P, 2), **opts)
    dL_drotations = torch.empty((P, 4), **opts)
    if P != 0:
        s = _settings(a["bg"], scale_modifier, a["viewmatrix"], a["projmatrix"], tan_fovx, tan_fovy, H, W, degree,
                      a["campos"], False, debug, tile, pair_store, capacity, pair_capacity, slot_capacity)
        g = _surfels(P, means3D, a["sh"], a["colors"], torch.empty(0, **opts), a["scales"], a["rotations"],
                     a["transMat_precomp"])
        gr = _lib.SrGrads()
        gr.dL_dmeans2D = dL_dmeans2D.data_ptr()
        gr.dL_dcolors = dL_dcolors.data_ptr()
        gr.dL_dopacity = dL_dopacity.data_ptr()
        gr.dL_dmeans3D = dL_dmeans3D.data_ptr()
        gr.dL_dtransMat = dL_dtransMat.data_ptr()
        gr.dL_dsh = _lib.ptr(dL_dsh)
        gr.dL_dscales = dL_dscales.data_ptr()
        gr.dL_drotations = dL_drotations.data_ptr()
        radii = radii.contiguous()
        with torch.cuda.device(dev):
            rc = lib.sr_backward(C.byref(s), C.byref(g), int(R), radii.data_ptr(), _lib.ptr(geomBuffer),
                                 _lib.ptr(binningBuffer), _lib.ptr(imageBuffer), a["dL_dout_color"].data_ptr(),
                                 a["dL_dout_others"].data_ptr(), C.byref(gr), _lib.stream_handle(dev))
        _lib.check(rc)
    return dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dtransMat, dL_dsh, dL_dscales, dL_drotations


def mark_visible(means3D, viewmatrix, projmatrix):
    """markVisible (rasterize_points.cu:235-254): bool[P], view z > 0.2."""
    lib = _lib.load()
    means3D = _f32(means3D, "means3D")
    viewmatrix = _f32(viewmatrix, "viewmatrix")
    projmatrix = _f32(projmatrix, "projmatrix")
    P = means3D.size(0)
    present = torch.zeros((P,), dtype=torch.bool, device=means3D.device)
    if P != 0:
        with torch.cuda.device(means3D.device):
            rc = lib.sr_mark_visible(P, means3D.data_ptr(), viewmatrix.data_ptr(), projmatrix.data_ptr(),
                                     present.data_ptr(), _lib.stream_handle(means3D.device))
        _lib.check(rc)
    return present


# Binning capacity per (device, tile, width, height): the last instance
# count with an eighth of headroom.  The first forward of a key sizes its
# binning buffer from the exact count (one stream synchronisation
# mid-forward, as the reference); later ones enqueue the whole forward at
# once and only fall back to a larger capacity when the count outgrows it.
# The accepted-pair store follows the pairs per instance of the latest
# completed forward on the device with the same tile (read back without
# waiting) times the instance capacity, so it shares the capacity's eighth
# of headroom; pairs beyond the store are recomputed with the same results.
_capacity_hint = {}
_pair_ratio = {}
_slot_ratio = {}  # partial-gradient slots per instance, likewise
_PAIRS_PER_INSTANCE = 8.0  # before any count is known


def _grown(n: int) -> int:
    return n + n // 8 + 1024


class _RasterizeGaussians(torch.autograd.Function):
    """Autograd wrapper, __init__.py:44-156."""

    @staticmethod
    def forward(ctx, means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                raster_settings):
        tile = getattr(raster_settings, "tile", 16)
        key = (means3D.device, tile, int(raster_settings.image_width), int(raster_settings.image_height))
        hint = 0 if raster_settings.debug else _capacity_hint.get(key, 0)
        cap = hint
        rkey = (means3D.device, tile)
        with torch.cuda.device(means3D.device):
            seen = _lib.pair_count()
            seen_slots = _lib.slot_count()
        if seen is not None and seen[1] > 0 and rkey in _pair_ratio:
            _pair_ratio[rkey] = seen[0] / seen[1]
        if seen_slots is not None and seen_slots[1] > 0 and rkey in _slot_ratio:
            _slot_ratio[rkey] = seen_slots[0] / seen_slots[1]
        while True:
            ratio = _pair_ratio.get(rkey) or _PAIRS_PER_INSTANCE
            # the capacity already holds an eighth more instances than the
            # last count: pairs and slots per instance scale with it
            pcap = min(int(ratio * cap) + 4096, 0x7FFFFFFF) if cap > 0 else 0
            sratio = _slot_ratio.get(rkey)
            scap = min(int(sratio * cap) + 1024, 0x7FFFFFFF) if cap > 0 and sratio else 0
            try:
                num_rendered, color, depth, radii, geomBuffer, binningBuffer, imgBuffer = rasterize_gaussians_fwd(
                    raster_settings.bg, means3D, colors_precomp, opacities, scales, rotations,
                    raster_settings.scale_modifier, cov3Ds_precomp, raster_settings.viewmatrix,
                    raster_settings.projmatrix, raster_settings.tanfovx, raster_settings.tanfovy,
                    raster_settings.image_height, raster_settings.image_width, sh, raster_settings.sh_degree,
                    raster_settings.campos, raster_settings.prefiltered, raster_settings.debug, tile,
                    capacity=cap, pair_capacity=pcap, slot_capacity=scap)
                break
            except _lib.CapacityError as e:
                if e.slots:
                    _slot_ratio[rkey] = 0.0  # this forward: every (instance, quadrant)
                else:
                    cap = _grown(e.needed)
        # keep the hint tracking the count: set by an exact or a retried
        # forward, shrunk when the count fell far below it
        if cap != hint or cap == 0 or _grown(num_rendered) * 4 < cap:
            _capacity_hint[key] = _grown(num_rendered)
        # from now on the pair store and the slots follow the measured counts
        _pair_ratio.setdefault(rkey, 0.0)
        _slot_ratio.setdefault(rkey, 0.0)
        ctx.pair_capacity = pcap
        ctx.slot_capacity = scap
        ctx.raster_settings = raster_settings
        ctx.capacity = cap
        ctx.num_rendered = num_rendered
        ctx.save_for_backward(colors_precomp, means3D, scales, rotations, cov3Ds_precomp, radii, sh, geomBuffer,
                              binningBuffer, imgBuffer)
        ctx.mark_non_differentiable(radii)
        # outputs without an incoming gradient arrive as None (the backward
        # fills zeros where it needs them) instead of zero tensors made here
        ctx.set_materialize_grads(False)
        return color, radii, depth

    @staticmethod
    def backward(ctx, grad_out_color, grad_radii, grad_depth):
        num_rendered = ctx.num_rendered
        rs = ctx.raster_settings
        colors_precomp, means3D, scales, rotations, cov3Ds_precomp, radii, sh, geomBuffer, binningBuffer, imgBuffer = (
            ctx.saved_tensors)
        if grad_out_color is None:
            grad_out_color = torch.zeros((3, rs.image_height, rs.image_width), device=means3D.device)
        if grad_depth is None:
            grad_depth = torch.zeros((7, rs.image_height, rs.image_width), device=means3D.device)
        (grad_means2D, grad_colors_precomp, grad_opacities, grad_means3D, grad_cov3Ds_precomp, grad_sh, grad_scales,
         grad_rotations) = rasterize_gaussians_bwd(
            rs.bg, means3D, radii, colors_precomp, scales, rotations, rs.scale_modifier, cov3Ds_precomp,
            rs.viewmatrix, rs.projmatrix, rs.tanfovx, rs.tanfovy, grad_out_color, grad_depth, sh, rs.sh_degree,
            rs.campos, geomBuffer, num_rendered, binningBuffer, imgBuffer, rs.debug, getattr(rs, "tile", 16),
            capacity=ctx.capacity, pair_capacity=ctx.pair_capacity, slot_capacity=ctx.slot_capacity)
        return (grad_means3D, grad_means2D, grad_sh, grad_colors_precomp, grad_opacities, grad_scales,
                grad_rotations, grad_cov3Ds_precomp, None)


def rasterize_gaussians(means3D, means2D, sh, colors_precomp, opacities, scales, rotations, cov3Ds_precomp,
                        raster_settings):
    return _RasterizeGaussians.apply(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
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
    """__init__.py:172-222."""

    def __init__(self, raster_settings):
        super().__init__()
        self.raster_settings = raster_settings

    def markVisible(self, positions):
        with torch.no_grad():
            rs = self.raster_settings
            visible = mark_visible(positions, rs.viewmatrix, rs.projmatrix)
        return visible

    def forward(self, means3D, means2D, opacities, shs=None, colors_precomp=None, scales=None, rotations=None,
                cov3D_precomp=None):
        rs = self.raster_settings
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise Exception("Please provide excatly one of either SHs or precomputed colors!")
        if ((scales is None or rotations is None) and cov3D_precomp is None) or (
                (scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise Exception("Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!")
        dev = means3D.device
        empty = torch.empty(0, dtype=torch.float32, device=dev)
        if shs is None:
            shs = empty
        if colors_precomp is None:
            colors_precomp = empty
        if scales is None:
            scales = empty
        if rotations is None:
            rotations = empty
        if cov3D_precomp is None:
            cov3D_precomp = empty
        return rasterize_gaussians(means3D, means2D, shs, colors_precomp, opacities, scales, rotations,
                                   cov3D_precomp, rs)
