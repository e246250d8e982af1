"""The render() post-processing (renderer/gaussian_renderer/__init__.py:104-149,
utils/point_utils.py:4-33), pinned by geometry: back-projected points land on
their pixels at their depth, planes give their own normal, the world-space
normal map is the camera rotation applied to the view-space one."""
import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from partgs_amd.camera import look_at, random_sphere_camera
from partgs_amd.renderer import depth_to_normal, depths_to_points, postprocess_allmap


def _c2w_rot(cam):
    return cam.world_view_transform.double()[:3, :3].numpy()  # = (W2C rot)^T = C2W rot


def test_depths_to_points_reproject_to_pixels():
    cam = random_sphere_camera(4, 3.0, 40, 30, fov_deg=50.0)
    rng = np.random.default_rng(0)
    depth = torch.tensor(rng.uniform(1.0, 5.0, size=(1, 30, 40)), dtype=torch.float32)
    pts = depths_to_points(cam, depth).double()
    hom = torch.cat([pts, torch.ones(pts.shape[0], 1, dtype=torch.float64)], 1)
    view = hom @ cam.world_view_transform.double()
    clip = hom @ cam.full_proj_transform.double()
    ndc = clip[:, :2] / clip[:, 3:4]
    W, H = 40, 30
    px = ndc[:, 0] * W / 2 + W / 2
    py = ndc[:, 1] * H / 2 + H / 2
    gy, gx = np.mgrid[0:H, 0:W]
    assert np.allclose(px.numpy(), gx.reshape(-1), atol=2e-3)
    assert np.allclose(py.numpy(), gy.reshape(-1), atol=2e-3)
    # the depth is the view-space z of the point
    assert np.allclose(view[:, 2].numpy(), depth.reshape(-1).double().numpy(), rtol=1e-5)


@pytest.mark.parametrize("normal_cam", [(0.0, 0.0, -1.0), (0.3, -0.2, -0.9), (-0.5, 0.4, -0.6)])
def test_depth_to_normal_of_a_plane(normal_cam):
    """Depth of the plane n.X = c seen from the camera: d(u,v) = c / (n . r),
    r the pixel ray with z = 1.  The pseudo normal is n (in world space),
    facing the camera, on every interior pixel."""
    W, H = 48, 36
    cam = look_at((0.5, -1.0, 3.0), width=W, height=H, fovx=math.radians(55), fovy=math.radians(45))
    n = np.asarray(normal_cam, dtype=np.float64)
    n /= np.linalg.norm(n)
    c = -2.0  # plane in front of the camera with the normal towards it
    # pixel rays r = K^-1 (x, y, 1), with the principal point of depths_to_points
    fx = W / (2 * math.tan(cam.FoVx / 2))
    fy = H / (2 * math.tan(cam.FoVy / 2))
    gy, gx = np.mgrid[0:H, 0:W].astype(np.float64)
    rx, ry = (gx - W / 2) / fx, (gy - H / 2) / fy
    d = c / (n[0] * rx + n[1] * ry + n[2])
    assert (d > 0).all()
    depth = torch.tensor(d[None], dtype=torch.float32)
    out = depth_to_normal(cam, depth).double().numpy()
    want = _c2w_rot(cam) @ n  # view -> world of a column vector
    inner = out[1:-1, 1:-1].reshape(-1, 3)
    assert np.allclose(inner, want[None], atol=2e-3), np.abs(inner - want).max()
    assert (out[0] == 0).all() and (out[-1] == 0).all() and (out[:, 0] == 0).all() and (out[:, -1] == 0).all()


def test_postprocess_allmap_channels():
    W, H = 24, 16
    cam = random_sphere_camera(9, 2.5, W, H)
    rng = np.random.default_rng(3)
    allmap = torch.tensor(rng.uniform(0.1, 2.0, size=(7, H, W)), dtype=torch.float32)
    allmap[1, :4] = 0.0  # alpha 0: expected depth 0/0 -> 0
    allmap[0, :4] = 0.0
    allmap[5, 0, 0] = float("nan")
    out = postprocess_allmap(cam, allmap, depth_ratio=0.25)
    R = _c2w_rot(cam)
    n_world = np.einsum("ij,jhw->ihw", R, allmap[2:5].double().numpy())
    assert np.allclose(out["rend_normal"].double().numpy(), n_world, atol=1e-5)
    exp_depth = allmap[0] / allmap[1]
    exp_depth[:4] = 0.0
    med = allmap[5].clone()
    med[0, 0] = 0.0
    surf = 0.75 * exp_depth + 0.25 * med
    assert torch.allclose(out["surf_depth"][0], surf, rtol=1e-6, atol=1e-6)
    assert torch.equal(out["rend_alpha"], allmap[1:2])
    assert torch.equal(out["rend_dist"], allmap[6:7])
    sn = depth_to_normal(cam, out["surf_depth"]).permute(2, 0, 1) * allmap[1:2]
    assert torch.allclose(out["surf_normal"], sn)


class _Model:
    """The GaussianModel properties render() reads, over fixed tensors."""

    def __init__(self, scene, dev):
        self.get_xyz = scene.means3D.to(dev).requires_grad_(True)
        self.get_opacity = scene.opacities.to(dev).requires_grad_(True)
        self.get_scaling = scene.scales.to(dev).requires_grad_(True)
        self.get_rotation = scene.rotations.to(dev).requires_grad_(True)
        self.get_features = scene.shs.to(dev).requires_grad_(True)
        self.active_sh_degree = scene.sh_degree


@pytest.mark.gpu
def test_render_matches_direct_rasteriser_call():
    from partgs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    from partgs_amd.renderer import render
    from partgs_amd.scenes import superquadric_surfels

    scene = superquadric_surfels(5, P=20000, n_sq=4, width=160, height=120, sh_degree=2)
    dev = torch.device("cuda")
    cam = scene.camera.to(dev)
    pc = _Model(scene, dev)
    pipe = SimpleNamespace(compute_cov3D_python=False, depth_ratio=0.0)
    bg = scene.bg.to(dev)
    out = render(cam, pc, pipe, bg)
    print(f"[render] P={scene.P} visible={int(out['visibility_filter'].sum())}", flush=True)
    rs = GaussianRasterizationSettings(image_height=120, image_width=160, tanfovx=math.tan(cam.FoVx * 0.5),
                                       tanfovy=math.tan(cam.FoVy * 0.5), bg=bg, scale_modifier=1.0,
                                       viewmatrix=cam.world_view_transform, projmatrix=cam.full_proj_transform,
                                       sh_degree=scene.sh_degree, campos=cam.camera_center, prefiltered=False,
                                       debug=False)
    m2 = torch.zeros_like(pc.get_xyz, requires_grad=True)
    color, radii, allmap = GaussianRasterizer(rs)(means3D=pc.get_xyz.detach(), means2D=m2,
                                                  shs=pc.get_features.detach(), opacities=pc.get_opacity.detach(),
                                                  scales=pc.get_scaling.detach(),
                                                  rotations=pc.get_rotation.detach())
    assert torch.equal(out["render"], color)
    assert torch.equal(out["radii"], radii)
    assert torch.equal(out["rend_alpha"], allmap[1:2])
    ref = postprocess_allmap(cam, allmap, 0.0)
    for k in ("rend_normal", "surf_depth", "surf_normal", "rend_dist"):
        assert torch.equal(out[k], ref[k]), k
    # gradients reach every model tensor and the screen-space points
    loss = out["render"].mean() + out["rend_dist"].mean() + (1 - (out["rend_normal"] * out["surf_normal"]).sum(0)).mean()
    loss.backward()
    for t in (pc.get_xyz, pc.get_opacity, pc.get_scaling, pc.get_rotation, pc.get_features):
        assert t.grad is not None and torch.isfinite(t.grad).all()
    assert out["viewspace_points"].grad is not None
