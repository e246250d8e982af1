"""The caller-side ``render()`` of the reference and its depth-to-normal helper.

Restates renderer/gaussian_renderer/__init__.py:9-149 (rasteriser set-up and
the post-processing of the 7-channel ``allmap``) and utils/point_utils.py:4-33
(``depths_to_points`` / ``depth_to_normal``) in torch, on the device of the
inputs.  The rasteriser call goes through :class:`GaussianRasterizer`, i.e.
the HIP library; the post-processing is a handful of elementwise maps and one
(H*W, 3) x (3, 3) product, which torch runs at memory speed.

``pc`` is any object with the reference GaussianModel's read-only
properties (``get_xyz``, ``get_opacity``, ``get_scaling``, ``get_rotation``,
``get_features``, ``active_sh_degree``; ``get_covariance`` for the
``compute_cov3D_python`` branch).  ``pipe`` needs ``compute_cov3D_python`` and
``depth_ratio``.  ``viewpoint_camera`` needs the fields of
:class:`partgs_amd.camera.Camera`.
"""
from __future__ import annotations

import math

import torch

from . import rasterizer_part
from .rasterizer import GaussianRasterizationSettings, GaussianRasterizer


def depths_to_points(view, depthmap: torch.Tensor) -> torch.Tensor:
    """World points of every pixel centre at the given depth (utils/point_utils.py:4-20).

    ``depthmap`` is (1, H, W) or (H, W); returns (H*W, 3).  The pixel grid is
    the reference's: x in [0, W), y in [0, H), principal point at (W/2, H/2).
    """
    dev = depthmap.device
    c2w = view.world_view_transform.to(dev).T.inverse()
    W, H = view.image_width, view.image_height
    ndc2pix = torch.tensor([[W / 2, 0, 0, W / 2],
                            [0, H / 2, 0, H / 2],
                            [0, 0, 0, 1]], dtype=torch.float32, device=dev).T
    projection_matrix = c2w.T @ view.full_proj_transform.to(dev)
    intrins = (projection_matrix @ ndc2pix)[:3, :3].T
    grid_x, grid_y = torch.meshgrid(torch.arange(W, device=dev).float(), torch.arange(H, device=dev).float(),
                                    indexing="xy")
    points = torch.stack([grid_x, grid_y, torch.ones_like(grid_x)], dim=-1).reshape(-1, 3)
    rays_d = points @ intrins.inverse().T @ c2w[:3, :3].T
    rays_o = c2w[:3, 3]
    return depthmap.reshape(-1, 1) * rays_d + rays_o


def depth_to_normal(view, depth: torch.Tensor) -> torch.Tensor:
    """Pseudo surface normal from a (1, H, W) depth map (utils/point_utils.py:22-33).

    Central differences of the back-projected points along y (rows) and x
    (columns); n = normalize(dy_rows x dx_cols).  The one-pixel border is 0.
    Returns (H, W, 3).
    """
    points = depths_to_points(view, depth).reshape(*depth.shape[1:], 3)
    output = torch.zeros_like(points)
    dx = points[2:, 1:-1] - points[:-2, 1:-1]
    dy = points[1:-1, 2:] - points[1:-1, :-2]
    normal_map = torch.nn.functional.normalize(torch.cross(dx, dy, dim=-1), dim=-1)
    output[1:-1, 1:-1, :] = normal_map
    return output


def _cov3d_precomp(viewpoint_camera, pc, scaling_modifier: float) -> torch.Tensor:
    """The ``compute_cov3D_python`` branch (renderer/gaussian_renderer/__init__.py:55-66):
    the (P, 9) column-major splat-to-pixel homography from ``pc.get_covariance``."""
    splat2world = pc.get_covariance(scaling_modifier)
    dev = splat2world.device
    W, H = viewpoint_camera.image_width, viewpoint_camera.image_height
    near, far = viewpoint_camera.znear, viewpoint_camera.zfar
    ndc2pix = torch.tensor([[W / 2, 0, 0, (W - 1) / 2],
                            [0, H / 2, 0, (H - 1) / 2],
                            [0, 0, far - near, near],
                            [0, 0, 0, 1]], dtype=torch.float32, device=dev).T
    world2pix = viewpoint_camera.full_proj_transform.to(dev) @ ndc2pix
    return (splat2world[:, [0, 1, 3]] @ world2pix[:, [0, 1, 3]]).permute(0, 2, 1).reshape(-1, 9)


def render(viewpoint_camera, pc, pipe, bg_color: torch.Tensor, scaling_modifier: float = 1.0,
           override_color=None, tile: int = 16) -> dict:
    """One view of ``pc`` (renderer/gaussian_renderer/__init__.py:9-149).

    Returns the reference's dict: ``render``, ``viewspace_points``,
    ``visibility_filter``, ``radii``, ``rend_alpha``, ``rend_normal`` (world
    space), ``rend_dist``, ``surf_depth`` and ``surf_normal``.
    """
    xyz = pc.get_xyz
    screenspace_points = torch.zeros_like(xyz, dtype=xyz.dtype, requires_grad=True, device=xyz.device) + 0
    try:
        screenspace_points.retain_grad()
    except RuntimeError:
        pass

    settings = GaussianRasterizationSettings(
        image_height=int(viewpoint_camera.image_height),
        image_width=int(viewpoint_camera.image_width),
        tanfovx=math.tan(viewpoint_camera.FoVx * 0.5),
        tanfovy=math.tan(viewpoint_camera.FoVy * 0.5),
        bg=bg_color,
        scale_modifier=scaling_modifier,
        viewmatrix=viewpoint_camera.world_view_transform,
        projmatrix=viewpoint_camera.full_proj_transform,
        sh_degree=pc.active_sh_degree,
        campos=viewpoint_camera.camera_center,
        prefiltered=False,
        debug=False,
        tile=tile,
    )
    rasterizer = GaussianRasterizer(raster_settings=settings)

    scales = rotations = cov3D_precomp = None
    if getattr(pipe, "compute_cov3D_python", False):
        cov3D_precomp = _cov3d_precomp(viewpoint_camera, pc, scaling_modifier)
    else:
        scales = pc.get_scaling
        rotations = pc.get_rotation

    # The reference forces convert_SHs_python off (gaussian_renderer/__init__.py:73), so
    # SH evaluation always runs inside the rasteriser unless colours override it.
    shs = colors_precomp = None
    if override_color is None:
        shs = pc.get_features
    else:
        colors_precomp = override_color

    rendered_image, radii, allmap = rasterizer(means3D=xyz, means2D=screenspace_points, shs=shs,
                                               colors_precomp=colors_precomp, opacities=pc.get_opacity,
                                               scales=scales, rotations=rotations, cov3D_precomp=cov3D_precomp)
    rets = {"render": rendered_image, "viewspace_points": screenspace_points, "visibility_filter": radii > 0,
            "radii": radii}
    rets.update(postprocess_allmap(viewpoint_camera, allmap, pipe.depth_ratio))
    return rets


def render_part(viewpoint_camera, pc, pipe, bg_color: torch.Tensor, scaling_modifier: float = 1.0,
                override_color=None, tile: int = 16) -> dict:
    """One view of ``pc`` through the _part rasteriser
    (renderer/gaussian_renderer_2d/__init__.py:10-150).

    ``pc.get_semantic`` (P, S) with S <= 16 is blended into ``render_semantic``
    (S, H, W); the other keys are those of ``render``.  With
    ``compute_cov3D_python`` the reference hands ``pc.get_covariance`` to the
    rasteriser as the precomputed homography unchanged, and so does this.
    """
    xyz = pc.get_xyz
    screenspace_points = torch.zeros_like(xyz, dtype=xyz.dtype, requires_grad=True, device=xyz.device) + 0
    try:
        screenspace_points.retain_grad()
    except RuntimeError:
        pass
    settings = rasterizer_part.GaussianRasterizationSettings(
        image_height=int(viewpoint_camera.image_height),
        image_width=int(viewpoint_camera.image_width),
        tanfovx=math.tan(viewpoint_camera.FoVx * 0.5),
        tanfovy=math.tan(viewpoint_camera.FoVy * 0.5),
        bg=bg_color,
        scale_modifier=scaling_modifier,
        viewmatrix=viewpoint_camera.world_view_transform,
        projmatrix=viewpoint_camera.full_proj_transform,
        sh_degree=pc.active_sh_degree,
        campos=viewpoint_camera.camera_center,
        prefiltered=False,
        debug=False,
        tile=tile,
    )
    rasterizer = rasterizer_part.GaussianRasterizer(raster_settings=settings)
    scales = rotations = cov3D_precomp = None
    if getattr(pipe, "compute_cov3D_python", False):
        cov3D_precomp = pc.get_covariance(scaling_modifier)
    else:
        scales = pc.get_scaling
        rotations = pc.get_rotation
    shs = colors_precomp = None
    if override_color is None:
        shs = pc.get_features
    else:
        colors_precomp = override_color
    rendered_image, semantic, radii, allmap = rasterizer(
        means3D=xyz, means2D=screenspace_points, opacities=pc.get_opacity, semantics=pc.get_semantic, shs=shs,
        colors_precomp=colors_precomp, scales=scales, rotations=rotations, cov3D_precomp=cov3D_precomp)
    rets = {"render": rendered_image, "render_semantic": semantic, "viewspace_points": screenspace_points,
            "visibility_filter": radii > 0, "radii": radii}
    rets.update(postprocess_allmap(viewpoint_camera, allmap, pipe.depth_ratio))
    return rets


def postprocess_allmap(viewpoint_camera, allmap: torch.Tensor, depth_ratio: float) -> dict:
    """The regularisation maps of renderer/gaussian_renderer/__init__.py:109-147.

    allmap channels: 0 alpha-weighted depth, 1 alpha, 2-4 view-space normal,
    5 median depth, 6 distortion.
    """
    render_alpha = allmap[1:2]
    wvt = viewpoint_camera.world_view_transform.to(allmap.device)
    render_normal = (allmap[2:5].permute(1, 2, 0) @ (wvt[:3, :3].T)).permute(2, 0, 1)
    render_depth_median = torch.nan_to_num(allmap[5:6], 0, 0)
    render_depth_expected = torch.nan_to_num(allmap[0:1] / render_alpha, 0, 0)
    render_dist = allmap[6:7]
    surf_depth = render_depth_expected * (1 - depth_ratio) + depth_ratio * render_depth_median
    surf_normal = depth_to_normal(viewpoint_camera, surf_depth).permute(2, 0, 1)
    surf_normal = surf_normal * render_alpha.detach()
    return {"rend_alpha": render_alpha, "rend_normal": render_normal, "rend_dist": render_dist,
            "surf_depth": surf_depth, "surf_normal": surf_normal}
