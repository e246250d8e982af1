"""The three functions of the reference's compiled ``_C`` module for the _part
variant (dsp/ext.cpp:15-19), served by the HIP library through ctypes.

Argument order, return tuples and error behaviour are those of
RasterizeGaussiansCUDA, RasterizeGaussiansBackwardCUDA and markVisible
(dsp/rasterize_points.h), so the reference's own
``diff_surfel_rasterization_part/__init__.py`` runs unchanged with
``from . import _C`` resolving to this file.
"""
from partgs_amd.rasterizer import mark_visible  # noqa: F401
from partgs_amd.rasterizer_part import rasterize_gaussians_bwd as rasterize_gaussians_backward  # noqa: F401
from partgs_amd.rasterizer_part import rasterize_gaussians_fwd as rasterize_gaussians  # noqa: F401
