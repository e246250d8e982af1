"""GPU parity of the _part rasteriser (submodules/diff-surfel-rasterization_part)
against the oracle's _part mode on the same inputs.

Bars as for the plain variant: integer state (radii, tiles, point list,
ranges, contributor counts), the record and every image channel (colour,
semantics, 8 aux channels, the three final-T planes) bit-exact against the
FP32 oracle; gradients (semantics included) within 1e-5, element-relative
with a 1e-5 max|g| floor, of the oracle's FP64-accumulated sums.  Cases: 0, 5
and 16 semantic channels, tiles 8/16/32, SH degree 0 and 3, precomputed
colours, a precomputed camera-space homography, 100,000 superquadric
surfels at 400x300, the autograd API, and run-to-run determinism.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from partgs_amd import rasterizer_part as RP
from partgs_amd.scenes import superquadric_surfels, surfel_cloud
from tests.helpers import assert_close_scaled, read_state, rel_err

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-5
GRAD_NAMES = ("dL_dmeans2D", "dL_dcolors", "dL_dsemantics", "dL_dopacity", "dL_dmeans3D", "dL_dtransMat", "dL_dsh",
              "dL_dscales", "dL_drotations")


def _semantics(P, S, seed=0):
    g = torch.Generator().manual_seed(100 + seed)
    return torch.rand(P, S, generator=g) * 2 - 1


def _upstream(S, H, W, seed=0):
    g = torch.Generator().manual_seed(200 + seed)
    return (torch.randn(3, H, W, generator=g), torch.randn(S, H, W, generator=g), torch.randn(8, H, W, generator=g))


def _gpu_forward(scene, sem, *, tile=16, colors=None, transmat=None):
    dev = torch.device("cuda")
    cam = scene.camera
    empty = torch.empty(0, device=dev)
    args = dict(bg=scene.bg.to(dev), means3D=scene.means3D.to(dev),
                colors=colors.to(dev) if colors is not None else empty, opacity=scene.opacities.to(dev),
                semantics=sem.to(dev), scales=scene.scales.to(dev) if transmat is None else empty,
                rotations=scene.rotations.to(dev) if transmat is None else empty, scale_modifier=1.0,
                transMat_precomp=transmat.to(dev) if transmat is not None else empty,
                viewmatrix=cam.world_view_transform.to(dev), projmatrix=cam.full_proj_transform.to(dev),
                tan_fovx=math.tan(cam.FoVx * 0.5), tan_fovy=math.tan(cam.FoVy * 0.5),
                image_height=cam.image_height, image_width=cam.image_width,
                sh=scene.shs.to(dev) if colors is None else empty, degree=scene.sh_degree,
                campos=cam.camera_center.to(dev), prefiltered=False, debug=False, tile=tile)
    L, color, semantic, others, radii, geom, binb, img = RP.rasterize_gaussians_fwd(**args)
    torch.cuda.synchronize()
    return dict(L=L, color=color, semantic=semantic, others=others, radii=radii, geom=geom, bin=binb, img=img,
                args=args, tile=tile, capacity=L)


def _gpu_backward(res, gc, gs, go):
    a = res["args"]
    dev = torch.device("cuda")
    out = RP.rasterize_gaussians_bwd(
        a["bg"], a["means3D"], res["radii"], a["colors"], a["semantics"], a["scales"], a["rotations"],
        a["scale_modifier"], a["transMat_precomp"], a["viewmatrix"], a["projmatrix"], a["tan_fovx"], a["tan_fovy"],
        gc.to(dev), gs.to(dev), go.to(dev), a["sh"], a["degree"], a["campos"], res["geom"], res["L"], res["bin"],
        res["img"], False, res["tile"])
    torch.cuda.synchronize()
    return {n: t.cpu().numpy() for n, t in zip(GRAD_NAMES, out)}


def _oracle(scene, sem, *, tile=16, colors=None, transmat=None):
    cam = scene.camera
    kw = dict(means3D=scene.means3D, opacities=scene.opacities, viewmatrix=cam.world_view_transform,
              projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=scene.bg, W=cam.image_width,
              H=cam.image_height, tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5), tile=tile)
    if colors is None:
        kw.update(shs=scene.shs, sh_degree=scene.sh_degree)
    else:
        kw.update(colors_precomp=colors)
    if transmat is None:
        kw.update(scales=scene.scales, rotations=scene.rotations)
    else:
        kw.update(transMat_precomp=transmat)
    o = O.Oracle()
    assert o.forward_part(semantics=sem.numpy(), **kw) == 0
    return o


def _check(scene, sem, *, tile=16, colors=None, transmat=None, seed=0):
    W, H = scene.camera.image_width, scene.camera.image_height
    S = sem.shape[1]
    print(f"[part] P={scene.P} {W}x{H} tile={tile} S={S} colors={colors is not None} "
          f"transmat={transmat is not None}", flush=True)
    o = _oracle(scene, sem, tile=tile, colors=colors, transmat=transmat)
    res = _gpu_forward(scene, sem, tile=tile, colors=colors, transmat=transmat)
    assert res["L"] == o.num_rendered
    radii = res["radii"].cpu().numpy()
    np.testing.assert_array_equal(radii, o.get("radii"))
    st = read_state(res, scene)
    vis = radii > 0
    np.testing.assert_array_equal(st["tiles_touched"], o.get("tiles_touched"))
    np.testing.assert_array_equal(st["depths"][vis].view(np.uint32), o.get("depths")[vis].view(np.uint32))
    for k in ("means2D", "transMat", "normal_opacity", "rgb"):
        np.testing.assert_array_equal(st[k][vis], o.get(k)[vis], err_msg=k)
    np.testing.assert_array_equal(st["ranges"], o.get("ranges"))
    np.testing.assert_array_equal(st["point_list"], o.get("point_list"))
    np.testing.assert_array_equal(st["n_contrib"], o.get("n_contrib"))
    np.testing.assert_array_equal(st["final_T"], o.get("final_T"))
    np.testing.assert_array_equal(res["color"].cpu().numpy(), o.get("out_color"))
    np.testing.assert_array_equal(res["others"].cpu().numpy(), o.get("out_others"))
    np.testing.assert_array_equal(res["semantic"].cpu().numpy(), o.get("out_semantic"))
    assert res["others"][1].max().item() > 0.1  # something is drawn
    gc, gs, go = _upstream(S, H, W, seed)
    o.backward_part(gc.numpy(), gs.numpy(), go.numpy())
    g = _gpu_backward(res, gc, gs, go)
    report = {}
    for name in GRAD_NAMES:
        ref = o.get(name)
        if ref.size == 0:
            continue
        report[name] = rel_err(g[name], ref)
        assert_close_scaled(g[name], ref, GRAD_TOL, name)
    print("[part] grad max rel err:", {k: f"{v:.2e}" for k, v in report.items()}, flush=True)
    return res, g, o


@pytest.mark.parametrize("S", [0, 5, 16])
def test_part_cloud_semantic_channels(S):
    scene = surfel_cloud(3, P=6000, width=160, height=120)
    _check(scene, _semantics(scene.P, S))


@pytest.mark.parametrize("tile", [8, 32])
def test_part_tiles(tile):
    scene = superquadric_surfels(1, P=20_000, width=200, height=152, sh_degree=3)
    _check(scene, _semantics(scene.P, 4, 1), tile=tile, seed=1)


def test_part_precomputed_colours_and_homography():
    scene = surfel_cloud(5, P=3000, width=128, height=96)
    sem = _semantics(scene.P, 3, 2)
    g = torch.Generator().manual_seed(7)
    colors = torch.rand(scene.P, 3, generator=g)
    # the camera-space homography the _part preprocess builds, handed back as
    # the precomputed one (then scales and rotations get no gradient)
    o = _oracle(scene, sem)
    tm = torch.from_numpy(np.ascontiguousarray(o.get("transMat"))).float()
    _check(scene, sem, colors=colors, transmat=tm, seed=2)


def test_part_superquadric_100k():
    scene = superquadric_surfels(0, P=100_000, width=400, height=300, sh_degree=3)
    _check(scene, _semantics(scene.P, 16, 3), seed=3)


def test_part_backward_is_deterministic():
    scene = superquadric_surfels(2, P=30_000, width=200, height=150, sh_degree=3)
    sem = _semantics(scene.P, 8, 4)
    res = _gpu_forward(scene, sem)
    gc, gs, go = _upstream(8, 150, 200, 4)
    g1 = _gpu_backward(res, gc, gs, go)
    g2 = _gpu_backward(_gpu_forward(scene, sem), gc, gs, go)
    for name in GRAD_NAMES:
        np.testing.assert_array_equal(g1[name], g2[name], err_msg=name)


def test_part_empty_and_culled():
    scene = surfel_cloud(0, P=500, width=64, height=48)
    dev = torch.device("cuda")
    empty_scene = type(scene)(scene.means3D[:0], scene.scales[:0], scene.rotations[:0], scene.opacities[:0],
                              scene.shs[:0], 0, scene.camera, scene.bg)
    res = _gpu_forward(empty_scene, torch.empty(0, 3))
    assert res["L"] == 0
    assert res["semantic"].shape == (3, 48, 64) and not res["semantic"].any()
    assert torch.equal(res["others"], torch.zeros(8, 48, 64, device=dev))
    # every surfel behind the camera
    behind = type(scene)(scene.means3D * 0.01 + scene.camera.camera_center * 1.5, scene.scales, scene.rotations,
                         scene.opacities, scene.shs, 0, scene.camera, scene.bg)
    res, g, _ = _check_empty(behind)
    assert res["L"] == 0 and not res["radii"].any()
    assert all(not np.any(v) for v in g.values())


def _check_empty(scene):
    sem = _semantics(scene.P, 2)
    o = _oracle(scene, sem)
    res = _gpu_forward(scene, sem)
    assert res["L"] == o.num_rendered == 0
    np.testing.assert_array_equal(res["color"].cpu().numpy(), o.get("out_color"))
    gc, gs, go = _upstream(2, 48, 64)
    return res, _gpu_backward(res, gc, gs, go), o


def test_part_autograd_api_matches_low_level():
    from diff_surfel_rasterization_part import GaussianRasterizationSettings, GaussianRasterizer

    scene = surfel_cloud(4, P=4000, width=128, height=96)
    dev = torch.device("cuda")
    sem = _semantics(scene.P, 6, 5)
    cam = scene.camera
    settings = GaussianRasterizationSettings(
        image_height=96, image_width=128, tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5),
        bg=scene.bg.to(dev), scale_modifier=1.0, viewmatrix=cam.world_view_transform.to(dev),
        projmatrix=cam.full_proj_transform.to(dev), sh_degree=0, campos=cam.camera_center.to(dev),
        prefiltered=False, debug=False)
    leaves = {k: v.to(dev).requires_grad_(True) for k, v in dict(
        means3D=scene.means3D, opacities=scene.opacities, semantics=sem, shs=scene.shs, scales=scene.scales,
        rotations=scene.rotations).items()}
    means2D = torch.zeros_like(leaves["means3D"], requires_grad=True)
    rast = GaussianRasterizer(settings)
    color, semantic, radii, allmap = rast(means3D=leaves["means3D"], means2D=means2D,
                                          opacities=leaves["opacities"], semantics=leaves["semantics"],
                                          shs=leaves["shs"], scales=leaves["scales"], rotations=leaves["rotations"])
    assert allmap.shape == (8, 96, 128) and semantic.shape == (6, 96, 128)
    gc, gs, go = _upstream(6, 96, 128, 5)
    loss = (color * gc.to(dev)).sum() + (semantic * gs.to(dev)).sum() + (allmap * go.to(dev)).sum()
    loss.backward()
    res = _gpu_forward(scene, sem)
    assert torch.equal(color, res["color"]) and torch.equal(semantic, res["semantic"])
    assert torch.equal(allmap, res["others"]) and torch.equal(radii, res["radii"])
    g = _gpu_backward(res, gc, gs, go)
    for name, t in (("dL_dmeans3D", leaves["means3D"]), ("dL_dopacity", leaves["opacities"]),
                    ("dL_dsemantics", leaves["semantics"]), ("dL_dsh", leaves["shs"]),
                    ("dL_dscales", leaves["scales"]), ("dL_drotations", leaves["rotations"]),
                    ("dL_dmeans2D", means2D)):
        np.testing.assert_array_equal(t.grad.cpu().numpy(), g[name], err_msg=name)
    vis = rast.markVisible(leaves["means3D"].detach())
    assert vis.dtype == torch.bool and vis.shape == (scene.P,)
