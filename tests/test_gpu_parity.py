"""GPU parity: the HIP library against the CPU oracle on the same inputs.

Bars: integer outputs (radii, tiles_touched, num_rendered, point_list,
ranges, n_contrib) and the depth bits bit-exact; images within 1e-5 of the
tensor scale; gradients within 1e-5 (element-relative with a 1e-5 * max|g|
floor per tensor) of the oracle's FP64-accumulated sums.
"""
import math

import numpy as np
import pytest
import torch

from partgs_amd import _lib
from partgs_amd.scenes import superquadric_surfels, surfel_cloud, upstream_grads
from tests.helpers import (assert_close_scaled, gpu_backward, gpu_forward, oracle_run, read_state, rel_err)

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-5
IMG_TOL = 1e-5


def _check_forward(scene, res, o, *, exact_images=True):
    print(f"[parity] P={scene.P} {scene.camera.image_width}x{scene.camera.image_height} tile={res['tile']} "
          f"L={res['L']}", flush=True)
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
    # the reference's sort keys, rebuilt from tile ids and depth bits
    ranges = st["ranges"]
    keys = np.zeros(res["L"], dtype=np.uint64)
    for t in np.nonzero(ranges[:, 1] > ranges[:, 0])[0]:
        keys[ranges[t, 0]:ranges[t, 1]] = np.uint64(t) << np.uint64(32)
    keys |= st["depths"].view(np.uint32)[st["point_list"]].astype(np.uint64)
    np.testing.assert_array_equal(keys, o.get("keys"))
    np.testing.assert_array_equal(st["n_contrib"], o.get("n_contrib"))
    color = res["color"].cpu().numpy()
    others = res["others"].cpu().numpy()
    if exact_images:
        np.testing.assert_array_equal(color, o.get("out_color"))
        np.testing.assert_array_equal(others, o.get("out_others"))
        np.testing.assert_array_equal(st["final_T"], o.get("final_T"))
    assert rel_err(color, o.get("out_color")) <= IMG_TOL
    assert rel_err(others, o.get("out_others")) <= IMG_TOL


def _check_backward(scene, res, o, seed=0, tol=GRAD_TOL):
    H, W = scene.camera.image_height, scene.camera.image_width
    gc, go = upstream_grads(seed, H, W)
    o.backward(gc.numpy(), go.numpy())
    g = gpu_backward(res, gc, go)
    report = {}
    for name in ("dL_dtransMat", "dL_dopacity", "dL_dcolors", "dL_dmeans3D", "dL_dscales", "dL_drotations",
                 "dL_dsh", "dL_dmeans2D"):
        ref = o.get(name)
        if name == "dL_dsh" and ref.size == 0:
            continue
        report[name] = rel_err(g[name], ref)
        assert_close_scaled(g[name], ref, tol, name)
    print("[parity] grad max rel err:", {k: f"{v:.2e}" for k, v in report.items()}, flush=True)


def test_forward_backward_cloud_c1():
    scene = surfel_cloud(0)
    o, rc = oracle_run(scene)
    assert rc == 0
    res = gpu_forward(scene)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o)


@pytest.mark.parametrize("pair_store", [0, 1])
def test_backward_is_deterministic(pair_store):
    """Every gradient sum is formed in a fixed order without atomics: two
    backward passes over the same forward state give the same bits, and a
    second forward + backward does too.  pair_store = 1 sends most quadrants
    through the path that recomputes their pairs."""
    scene = superquadric_surfels(3, P=20000, width=160, height=120)
    H, W = scene.camera.image_height, scene.camera.image_width
    gc, go = upstream_grads(3, H, W)
    res = gpu_forward(scene, pair_store=pair_store)
    print(f"[parity] determinism L={res['L']} pair_store={pair_store}", flush=True)
    g1 = gpu_backward(res, gc, go)
    g2 = gpu_backward(res, gc, go)
    g3 = gpu_backward(gpu_forward(scene, pair_store=pair_store), gc, go)
    for k in g1:
        np.testing.assert_array_equal(g1[k], g2[k], err_msg=k)
        np.testing.assert_array_equal(g1[k], g3[k], err_msg=k)


def test_recompute_path_gives_stored_path_bits():
    """A quadrant whose pairs missed the accepted-pair store runs the backward
    by recomputing them; it sums each entry's partial gradients over the
    pixels in the same order as the stored path, so the gradients do not
    depend on which quadrants fit (which the store's first-come reservation
    decides)."""
    scene = superquadric_surfels(3, P=20000, width=160, height=120)
    H, W = scene.camera.image_height, scene.camera.image_width
    gc, go = upstream_grads(3, H, W)
    full = gpu_backward(gpu_forward(scene), gc, go)
    print("[parity] stored vs recomputed pairs", flush=True)
    for cap in (1, 20000, 200000):
        res = gpu_forward(scene, capacity=40000, pair_capacity=cap)
        g = gpu_backward(res, gc, go)
        for k in g:
            np.testing.assert_array_equal(g[k], full[k], err_msg=f"{k} pair_capacity={cap}")


def test_compact_slots_give_the_same_bits():
    """The partial-gradient slots sized from the measured count (and from
    the exact count when no capacity is given) hold the same sums in the
    same order as one slot per (instance, quadrant): bitwise the same
    gradients.  A slot capacity below the count is reported as a capacity
    error with the instances within the capacity."""
    scene = superquadric_surfels(4, P=20000, width=160, height=120)
    H, W = scene.camera.image_height, scene.camera.image_width
    gc, go = upstream_grads(4, H, W)
    exact = gpu_forward(scene)
    g_exact = gpu_backward(exact, gc, go)
    L = exact["L"]
    counts = _lib.slot_count()
    assert counts is not None and counts[1] == L
    S = counts[0]
    assert L <= S <= 4 * L
    print(f"[parity] slots {S} for {L} instances", flush=True)
    full_f = gpu_forward(scene, capacity=L + 100)
    full = gpu_backward(full_f, gc, go)
    tight = gpu_forward(scene, capacity=L + 100, slot_capacity=S)
    assert tight["bin"].numel() < full_f["bin"].numel()
    g_tight = gpu_backward(tight, gc, go)
    for k in g_exact:
        np.testing.assert_array_equal(g_exact[k], full[k], err_msg=k)
        np.testing.assert_array_equal(g_exact[k], g_tight[k], err_msg=k)
    with pytest.raises(_lib.CapacityError) as e:
        gpu_forward(scene, capacity=L + 100, slot_capacity=S - 1)
    assert e.value.slots and e.value.needed == L


def test_autograd_recovers_from_a_slot_overflow(monkeypatch):
    """The autograd wrapper sizes the slots from the last forward's ratio;
    when a view needs more it renders again with one slot per (instance,
    quadrant), with the same results."""
    from partgs_amd import rasterizer as R

    scene = superquadric_surfels(5, P=20000, width=160, height=120)
    cam = scene.camera
    s = scene.to("cuda")
    rs = R.GaussianRasterizationSettings(image_height=cam.image_height, image_width=cam.image_width,
                                         tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5),
                                         bg=s.bg, scale_modifier=1.0, viewmatrix=s.camera.world_view_transform,
                                         projmatrix=s.camera.full_proj_transform, sh_degree=s.sh_degree,
                                         campos=s.camera.camera_center, prefiltered=False, debug=False)
    rast = R.GaussianRasterizer(rs)
    outs = []
    for step in range(3):
        if step == 2:  # report a slot count far below what this view needs
            monkeypatch.setattr(_lib, "slot_count", lambda: (1, 1000))
        leaves = [t.clone().requires_grad_(True) for t in (s.means3D, s.shs, s.opacities, s.scales, s.rotations)]
        m2 = torch.zeros_like(s.means3D, requires_grad=True)
        color, radii, allmap = rast(means3D=leaves[0], means2D=m2, shs=leaves[1], opacities=leaves[2],
                                    scales=leaves[3], rotations=leaves[4])
        (color.sum() + allmap.sum()).backward()
        torch.cuda.synchronize()
        outs.append((color.detach().cpu(), leaves[0].grad.cpu()))
    for c, g in outs[1:]:
        assert torch.equal(c, outs[0][0])
        assert torch.equal(g, outs[0][1])


@pytest.mark.parametrize("tile", [8, 32])
def test_tile_sizes(tile):
    scene = surfel_cloud(1, P=4096, width=200, height=150)
    o, _ = oracle_run(scene, tile=tile)
    res = gpu_forward(scene, tile=tile)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=1)


@pytest.mark.parametrize("pair_store", [0, 1])
def test_superquadric_sh3_small(pair_store):
    """pair_store = 1 overflows the accepted-pair store, so most chunks take
    the recompute path."""
    scene = superquadric_surfels(0, P=20000, width=160, height=120)
    o, _ = oracle_run(scene)
    res = gpu_forward(scene, pair_store=pair_store)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=2)


def test_precomputed_colors_and_scale_modifier():
    scene = surfel_cloud(3, P=3000, width=96, height=80)
    g = torch.Generator().manual_seed(3)
    colors = torch.rand(scene.P, 3, generator=g)
    o, _ = oracle_run(scene, colors=colors, scale_modifier=1.7)
    res = gpu_forward(scene, colors=colors, scale_modifier=1.7)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=3)


def test_precomputed_transmat():
    scene = surfel_cloud(4, P=3000, width=96, height=80)
    # transMat from a first oracle pass (a valid homography per surfel)
    o0, _ = oracle_run(scene)
    tm = torch.from_numpy(o0.get("transMat").copy())
    o, _ = oracle_run(scene, transmat=tm)
    res = gpu_forward(scene, transmat=tm)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=4)


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("P", [12000, 40000])
def test_large_tile_sorts(P, tile):
    """Tiles above 2048 instances are sorted as 2048-key chunks merged by rank;
    above 16384, by one workgroup (LDS chunks, merge path in global memory).
    Tile 8 has one quadrant per tile, tile 32 sixteen (four packed counters)."""
    scene = surfel_cloud(5, P=P, width=32, height=32)
    cam = scene.camera
    # pull every centre onto the optical axis region so all land in few tiles
    c2w = torch.linalg.inv(cam.world_view_transform.T)
    fwd = c2w[:3, 2]
    centre = cam.camera_center + 3.0 * fwd
    g = torch.Generator().manual_seed(5)
    scene.means3D = (centre + 0.05 * (torch.rand(scene.P, 3, generator=g) - 0.5)).float()
    o, _ = oracle_run(scene, tile=tile)
    st_counts = np.diff(o.get("ranges"), axis=1).max()
    assert st_counts > (16384 if P > 20000 else 2048), st_counts
    res = gpu_forward(scene, tile=tile)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=5)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_early_reject_stress(tile):
    """Opacities around 1/255, grazing and edge-on surfels, huge and sub-pixel
    scales: the quadrant culling and the early rejection must not drop a
    pair the reference blends."""
    from partgs_amd.scenes import random_rotations, quat_from_matrix

    scene = surfel_cloud(11, P=6000, width=136, height=104)
    rng = np.random.default_rng(11)
    P = scene.P
    opa = np.concatenate([rng.uniform(0.9 / 255, 1.6 / 255, P // 3), rng.uniform(0.0, 1.0, P - P // 3)])
    scene.opacities = torch.from_numpy(opa.reshape(-1, 1)).float()
    sc = np.exp(rng.uniform(np.log(1e-4), np.log(0.6), (P, 2)))
    scene.scales = torch.from_numpy(sc).float()
    # a third of the surfels almost edge-on to the camera
    cam = scene.camera
    c2w = torch.linalg.inv(cam.world_view_transform.T).numpy().astype(np.float64)
    R = random_rotations(rng, P)
    k = P // 3
    view_dir = c2w[:3, 2]
    for i in range(k):
        n = np.cross(view_dir, rng.normal(size=3))
        n /= np.linalg.norm(n)
        n = n + rng.normal(scale=1e-3, size=3)
        n /= np.linalg.norm(n)
        t = np.cross(n, rng.normal(size=3))
        t /= np.linalg.norm(t)
        R[i] = np.stack([t, np.cross(n, t), n], axis=1)
    scene.rotations = torch.from_numpy(quat_from_matrix(R)).float()
    o, _ = oracle_run(scene, tile=tile)
    res = gpu_forward(scene, tile=tile)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=11)


def test_empty_and_culled():
    scene = surfel_cloud(6, P=1000, width=64, height=48)
    # P = 0: zero images, like the reference glue
    empty = surfel_cloud(6, P=1000, width=64, height=48)
    for k in ("means3D", "scales", "rotations", "opacities", "shs"):
        setattr(empty, k, getattr(empty, k)[:0])
    res = gpu_forward(empty)
    assert res["L"] == 0
    assert float(res["color"].abs().max()) == 0.0
    # everything behind the near plane: background only, no instances
    cam = scene.camera
    c2w = torch.linalg.inv(cam.world_view_transform.T)
    scene.means3D = (cam.camera_center - 2.0 * c2w[:3, 2] + 0.01 * scene.means3D).float()
    o, _ = oracle_run(scene)
    res = gpu_forward(scene)
    assert res["L"] == 0 == o.num_rendered
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=6)
    # the same with the binning enqueued before the (zero) count is known
    res = gpu_forward(scene, capacity=64)
    assert res["L"] == 0
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=6)


def test_capacity_mode_matches_exact():
    """A binning capacity given up front (no synchronisation mid-forward)
    gives the same instance count, state, images and gradients as sizing the
    buffer from the count; too small a capacity raises CapacityError carrying
    the count, and the exact count is enough."""
    from partgs_amd._lib import CapacityError

    scene = superquadric_surfels(4, P=20000, width=160, height=120)
    H, W = scene.camera.image_height, scene.camera.image_width
    gc, go = upstream_grads(4, H, W)
    ref = gpu_forward(scene)
    L = ref["L"]
    print(f"[parity] capacity mode L={L}", flush=True)
    g_ref = gpu_backward(ref, gc, go)
    st_ref = read_state(ref, scene)
    with pytest.raises(CapacityError) as ei:
        gpu_forward(scene, capacity=L // 2)
    assert ei.value.needed == L
    for cap in (L, L + 1000):
        res = gpu_forward(scene, capacity=cap)
        assert res["L"] == L
        np.testing.assert_array_equal(res["color"].cpu().numpy(), ref["color"].cpu().numpy())
        np.testing.assert_array_equal(res["others"].cpu().numpy(), ref["others"].cpu().numpy())
        st = read_state(res, scene)
        for k in ("point_list", "ranges", "n_contrib", "final_T"):
            np.testing.assert_array_equal(st[k], st_ref[k], err_msg=k)
        g = gpu_backward(res, gc, go)
        for k in g:
            np.testing.assert_array_equal(g[k], g_ref[k], err_msg=k)


def test_prefiltered_violation_raises():
    scene = surfel_cloud(7, P=500, width=64, height=48)
    cam = scene.camera
    c2w = torch.linalg.inv(cam.world_view_transform.T)
    scene.means3D[:10] = (cam.camera_center - 1.0 * c2w[:3, 2]).float()  # behind the camera
    o, rc = oracle_run(scene, prefiltered=True)
    assert rc == 1
    with pytest.raises(RuntimeError, match="prefiltered"):
        gpu_forward(scene, prefiltered=True)


def test_autograd_api_matches_low_level():
    from partgs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer

    scene = surfel_cloud(8, P=5000, width=128, height=96).to("cuda")
    cam = scene.camera
    rs = GaussianRasterizationSettings(image_height=cam.image_height, image_width=cam.image_width,
                                       tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5), bg=scene.bg,
                                       scale_modifier=1.0, viewmatrix=cam.world_view_transform,
                                       projmatrix=cam.full_proj_transform, sh_degree=0, campos=cam.camera_center,
                                       prefiltered=False, debug=False)
    means3D = scene.means3D.clone().requires_grad_(True)
    means2D = torch.zeros_like(means3D, requires_grad=True)
    shs = scene.shs.clone().requires_grad_(True)
    opac = scene.opacities.clone().requires_grad_(True)
    scales = scene.scales.clone().requires_grad_(True)
    rots = scene.rotations.clone().requires_grad_(True)
    color, radii, allmap = GaussianRasterizer(rs)(means3D=means3D, means2D=means2D, shs=shs, opacities=opac,
                                                  scales=scales, rotations=rots)
    gc, go = upstream_grads(8, cam.image_height, cam.image_width)
    ((color * gc.cuda()).sum() + (allmap * go.cuda()).sum()).backward()
    res = gpu_forward(scene.to("cpu"))
    g = gpu_backward(res, gc, go)
    np.testing.assert_array_equal(color.detach().cpu().numpy(), res["color"].cpu().numpy())
    for t, name in ((means3D, "dL_dmeans3D"), (opac, "dL_dopacity"), (scales, "dL_dscales"),
                    (rots, "dL_drotations"), (means2D, "dL_dmeans2D")):
        assert_close_scaled(t.grad.cpu().numpy(), g[name], 1e-5, name)
    # later forwards reuse the binning capacity of earlier ones (no
    # synchronisation mid-forward), growing it when outgrown: same results
    from partgs_amd import rasterizer as R
    key = (means3D.device, 16, cam.image_width, cam.image_height)
    for hint in (None, 1):
        if hint is not None:
            R._capacity_hint[key] = hint  # far too small: the forward retries
        leaves = (means3D, means2D, shs, opac, scales, rots)
        for t in leaves:
            t.grad = None
        color2, _, allmap2 = GaussianRasterizer(rs)(means3D=means3D, means2D=means2D, shs=shs, opacities=opac,
                                                    scales=scales, rotations=rots)
        ((color2 * gc.cuda()).sum() + (allmap2 * go.cuda()).sum()).backward()
        assert R._capacity_hint[key] >= res["L"]
        np.testing.assert_array_equal(color2.detach().cpu().numpy(), res["color"].cpu().numpy())
        np.testing.assert_array_equal(allmap2.detach().cpu().numpy(), res["others"].cpu().numpy())
        for t, name in ((means3D, "dL_dmeans3D"), (opac, "dL_dopacity"), (scales, "dL_dscales"),
                        (rots, "dL_drotations"), (means2D, "dL_dmeans2D")):
            assert_close_scaled(t.grad.cpu().numpy(), g[name], 1e-5, name)


def test_mark_visible_matches_oracle_across_near_plane():
    """markVisible (dsr/cuda_rasterizer/rasterizer_impl.cu:54-66,141-153):
    present = view z > 0.2.  Surfels placed at view depths in [0, 0.4], a
    thousand of them within float rounding of 0.2, the rest on both sides."""
    from oracle import oracle as O
    from partgs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer

    scene = surfel_cloud(9, P=20000, width=64, height=48)
    cam = scene.camera
    c2w = torch.linalg.inv(cam.world_view_transform.T).double()
    rng = np.random.default_rng(9)
    z = rng.uniform(0.0, 0.4, scene.P)
    z[:1000] = 0.2 + rng.integers(-4, 5, 1000) * np.spacing(np.float32(0.2))
    xy = rng.uniform(-0.3, 0.3, (scene.P, 2))
    local = torch.tensor(np.stack([xy[:, 0], xy[:, 1], z, np.ones(scene.P)], 1))
    means = (local @ c2w.T)[:, :3].float()
    rs = GaussianRasterizationSettings(image_height=48, image_width=64, tanfovx=0.5, tanfovy=0.5,
                                       bg=torch.ones(3, device="cuda"), scale_modifier=1.0,
                                       viewmatrix=cam.world_view_transform.cuda(),
                                       projmatrix=cam.full_proj_transform.cuda(), sh_degree=0,
                                       campos=cam.camera_center.cuda(), prefiltered=False, debug=False)
    vis = GaussianRasterizer(rs).markVisible(means.cuda()).cpu().numpy()
    want = O.mark_visible(means, cam.world_view_transform)
    print(f"[parity] mark_visible: {int(want.sum())} of {scene.P} in front", flush=True)
    assert 0 < want.sum() < scene.P and 0 < want[:1000].sum() < 1000
    np.testing.assert_array_equal(vis, want)
