"""Import name of the reference rasteriser package, served by partgs_amd.

``from diff_surfel_rasterization import GaussianRasterizationSettings,
GaussianRasterizer`` (as in renderer/gaussian_renderer/__init__.py:5) resolves
here when this repository is on sys.path.
"""
from partgs_amd.rasterizer import (GaussianRasterizationSettings, GaussianRasterizer,  # noqa: F401
                                   _RasterizeGaussians, rasterize_gaussians)
