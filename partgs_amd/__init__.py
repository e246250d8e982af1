"""MI355X-native surfel rasteriser (2D Gaussian splatting) for PartGS.

Public API mirrors the reference package ``diff_surfel_rasterization``:
``GaussianRasterizationSettings``, ``GaussianRasterizer``, ``rasterize_gaussians``.
The compute runs in ``lib/libsurfel_rast.so`` (hand-written HIP for gfx950).
"""
from .rasterizer import (GaussianRasterizationSettings, GaussianRasterizer, mark_visible,  # noqa: F401
                         rasterize_gaussians, rasterize_gaussians_bwd, rasterize_gaussians_fwd)

__all__ = [
    "GaussianRasterizationSettings",
    "GaussianRasterizer",
    "rasterize_gaussians",
    "rasterize_gaussians_fwd",
    "rasterize_gaussians_bwd",
    "mark_visible",
]
