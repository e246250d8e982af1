"""Import name of the reference's _part rasteriser package, served by partgs_amd.

``from diff_surfel_rasterization_part import GaussianRasterizationSettings,
GaussianRasterizer`` (as in renderer/gaussian_renderer_2d/__init__.py) resolves
here when this repository is on sys.path.
"""
from partgs_amd.rasterizer_part import (GaussianRasterizationSettings, GaussianRasterizer,  # noqa: F401
                                        _RasterizeGaussians, rasterize_gaussians)
