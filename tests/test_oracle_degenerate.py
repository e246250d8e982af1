"""The FP64 preprocess of near-degenerate surfels, on the oracle.

The reference evaluates every surfel's homography, camera-facing test and
screen AABB in FP32 (dsr/cuda_rasterizer/forward.cu:75-145, 183-247).  Four
of its decisions can go the other way than exact arithmetic would: the
normal flip and the cos == 0 cull for grazing surfels (forward.cu:209-214),
the d == 0 exit of compute_aabb (:125-126), the radius ceil(sqrt(max(1e-4,
h0))) where h0 = p^2 - f T T cancels (:141, 230), and the tile rectangle
(getRect, auxiliary.h:67-77).  Surfels whose FP32 values lie within a
rounding bound of one of these decisions, or whose footprint is under
1/128 pixel, take that part of the preprocess in FP64 (the HIP kernel and
the oracle select them identically; flip = +-2 marks them).

Checks against the oracle's FP64 truth mode on scenes built to be
degenerate: with the FP64 branch the culls, radii and normal signs equal
the truth; pure FP32 (the reference's arithmetic) differs on thousands of
surfels.  And both sides of the grazing, radius and tiny thresholds."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from partgs_amd.camera import look_at
from partgs_amd.scenes import quat_from_matrix, random_rotations, surfel_cloud


def _run(scene, f64=False, colors=False):
    cam = scene.camera
    o = O.Oracle()
    kw = dict(means3D=scene.means3D, opacities=scene.opacities, viewmatrix=cam.world_view_transform,
              projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=scene.bg, W=cam.image_width,
              H=cam.image_height, tanfovx=math.tan(cam.FoVx / 2), tanfovy=math.tan(cam.FoVy / 2),
              scales=scene.scales, rotations=scene.rotations, f64=f64)
    if colors:
        kw["colors_precomp"] = torch.rand(scene.P, 3, generator=torch.Generator().manual_seed(0))
    else:
        kw.update(shs=scene.shs, sh_degree=0)
    assert o.forward(**kw) == 0
    return o


def degenerate_scene(seed=5, P=60000, W=800, H=600):
    """Half the surfels almost edge-on (|cos| from 1e-9 to 1e-2 of |p|),
    scales log-uniform from 1e-7 to 0.3 (tiny to large)."""
    sc = surfel_cloud(seed, P=P, width=W, height=H)
    rng = np.random.default_rng(seed)
    cam = sc.camera
    eye = torch.linalg.inv(cam.world_view_transform.T).numpy().astype(np.float64)[:3, 3]
    means = sc.means3D.numpy().astype(np.float64)
    sc.scales = torch.from_numpy(np.exp(rng.uniform(np.log(1e-7), np.log(0.3), (P, 2)))).float()
    R = random_rotations(rng, P)
    for i in range(P // 2):
        vd = means[i] - eye
        vd /= np.linalg.norm(vd)
        n = np.cross(vd, rng.normal(size=3))
        n /= np.linalg.norm(n)
        n = n + vd * rng.normal() * 10 ** rng.uniform(-9, -2)
        n /= np.linalg.norm(n)
        t = np.cross(n, rng.normal(size=3))
        t /= np.linalg.norm(t)
        R[i] = np.stack([t, np.cross(n, t), n], 1)
    sc.rotations = torch.from_numpy(quat_from_matrix(R)).float()
    return sc


def _divergence(a, truth):
    ra, rt = a.get("radii"), truth.get("radii")
    vis = (ra > 0) & (rt > 0)
    na, nt = a.get("normal_opacity")[:, :3], truth.get("normal_opacity")[:, :3]
    return dict(culls=int(((ra > 0) != (rt > 0)).sum()), radii=int((ra != rt).sum()),
                flips=int((((na * nt).sum(1) < 0) & vis).sum()),
                tiles=int((a.get("tiles_touched") != truth.get("tiles_touched")).sum()),
                centre=float(np.abs(a.get("means2D")[vis] - truth.get("means2D")[vis]).max()))


def test_fp64_branch_takes_the_exact_decisions():
    sc = degenerate_scene()
    truth = _run(sc, f64=True)
    O.set_fp64_degenerate(False)
    try:
        pure = _divergence(_run(sc), truth)
    finally:
        O.set_fp64_degenerate(True)
    o = _run(sc)
    got = _divergence(o, truth)
    n_deg = int((np.abs(o.get("flip")) == 2).sum())
    print(f"[degenerate] {n_deg} of {sc.P} surfels in FP64; pure FP32 vs truth {pure}; with FP64 {got}")
    assert n_deg > 1000
    # the reference's FP32 takes thousands of these decisions differently
    assert pure["culls"] > 100 and pure["radii"] > 100 and pure["flips"] > 100
    assert got["culls"] == 0 and got["radii"] == 0 and got["flips"] == 0
    # the rectangle comes from the centre stored in FP32, as the binning
    # recomputes it: a centre within half an FP32 ulp of a tile edge can
    # round across it
    assert got["tiles"] <= 8
    assert got["centre"] < 1e-3


def test_grazing_threshold_both_sides():
    """Surfels at the image centre whose normal makes cos = eps |p| with the
    view ray, eps from 1e-9 to 1e-3: below 2^-16 (minus the FP32 rounding of
    the test itself) they take the FP64 path, above it the FP32 one; on both
    sides the normal sign equals the truth."""
    W, H = 64, 48
    cam = look_at([0.0, 0.0, -3.0], width=W, height=H)
    eps = np.geomspace(1e-9, 1e-3, 400)
    P = eps.size
    rng = np.random.default_rng(2)
    means = rng.uniform(-0.05, 0.05, (P, 3))
    eye = np.array([0.0, 0.0, -3.0])
    R = np.zeros((P, 3, 3))
    for i in range(P):
        vd = means[i] - eye
        vd /= np.linalg.norm(vd)
        n = np.cross(vd, rng.normal(size=3))
        n /= np.linalg.norm(n)
        n = np.sqrt(1 - eps[i] ** 2) * n - eps[i] * vd * (1 if i % 2 else -1)
        t = np.cross(n, rng.normal(size=3))
        t /= np.linalg.norm(t)
        R[i] = np.stack([t, np.cross(n, t), n], 1)
    sc = surfel_cloud(0, P=P, width=W, height=H)
    sc.camera = cam
    sc.means3D = torch.from_numpy(means).float()
    sc.rotations = torch.from_numpy(quat_from_matrix(R)).float()
    sc.scales = torch.full((P, 2), 0.05)
    o, truth = _run(sc), _run(sc, f64=True)
    fl = o.get("flip")
    inside = eps < 2.0 ** -16 * 0.25
    outside = eps > 2.0 ** -16 * 4
    assert (np.abs(fl[inside]) == 2).all()
    # outside the grazing band the FP32 path is kept unless another test fires
    assert (np.abs(fl[outside]) == 1).mean() > 0.9
    d = _divergence(o, truth)
    assert d["flips"] == 0 and d["culls"] == 0, d


def test_radius_threshold_both_sides():
    """A fronto-parallel surfel far off the image centre (h0 cancels against
    p^2) whose extent sweeps across h0 = 16 (radius 4 -> 5): near the step
    it takes the FP64 path and its radius equals the exact one; away from
    the step the FP32 path is kept and agrees too."""
    W, H = 800, 600
    cam = look_at([0.0, 0.0, -3.0], width=W, height=H)
    tx = math.tan(cam.FoVx / 2)
    focal = W / (2 * tx)
    P = 600
    # centre near the right edge, x pixel about 780
    x_view = (780 - (W - 1) / 2) * 3.0 / focal
    h = np.linspace(3.9, 4.1, P)  # 3-sigma half extent in pixels
    sigma_world = h / 3.0 * 3.0 / focal
    sc = surfel_cloud(0, P=P, width=W, height=H)
    sc.camera = cam
    sc.means3D = torch.tensor(np.tile([x_view, 0.0, 0.0], (P, 1))).float()
    sc.rotations = torch.tensor(np.tile([1.0, 0.0, 0.0, 0.0], (P, 1))).float()
    sc.scales = torch.from_numpy(np.stack([sigma_world, sigma_world], 1)).float()
    o, truth = _run(sc), _run(sc, f64=True)
    fl, r, rt = o.get("flip"), o.get("radii"), truth.get("radii")
    assert set(np.unique(rt)) == {4, 5}
    np.testing.assert_array_equal(r, rt)
    step = np.argmax(rt == 5)
    assert np.abs(fl[step - 2:step + 2]).min() == 2  # the surfels at the step took FP64
    assert np.abs(fl[:P // 10]).max() == 1 and np.abs(fl[-P // 10:]).max() == 1  # far from it: FP32


def test_tiny_threshold_both_sides():
    """Footprints from 1/1000 to 1/10 pixel: below 1/128 pixel FP64, above
    FP32 (unless another test fires)."""
    W, H = 64, 48
    cam = look_at([0.0, 0.0, -3.0], width=W, height=H)
    focal = W / (2 * math.tan(cam.FoVx / 2))
    px_size = np.geomspace(1e-3, 1e-1, 200)
    P = px_size.size
    sc = surfel_cloud(0, P=P, width=W, height=H)
    sc.camera = cam
    rng = np.random.default_rng(1)
    sc.means3D = torch.from_numpy(rng.uniform(-0.3, 0.3, (P, 3)) * [1, 1, 0.1]).float()
    sc.rotations = torch.tensor(np.tile([1.0, 0.0, 0.0, 0.0], (P, 1))).float()
    s = px_size * 3.0 / focal
    sc.scales = torch.from_numpy(np.stack([s, s], 1)).float()
    o, truth = _run(sc), _run(sc, f64=True)
    fl = o.get("flip")
    assert (np.abs(fl[px_size < 1 / 128 * 0.8]) == 2).all()
    assert (np.abs(fl[px_size > 1 / 128 * 1.25]) == 1).mean() > 0.9
    assert _divergence(o, truth)["radii"] == 0


def test_backward_uses_forward_flip():
    """The backward takes the camera-facing sign from the forward's record
    (the reference re-derives it from p_view . normal, backward.cu:550-555).
    With the FP64 branch the FP32 backward's rotation gradients (through
    the normal channels) agree with the FP64 truth's; pure FP32 re-derives
    the sign its forward got wrong and is off on those surfels."""
    sc = degenerate_scene(7, P=6000, W=160, H=120)
    cam = sc.camera
    rng = np.random.default_rng(8)
    go = np.zeros((7, 120, 160))
    go[2:5] = rng.normal(size=(3, 120, 160))  # the normal channels only
    gc = np.zeros((3, 120, 160))

    def grads(f64):
        o = _run(sc, f64=f64)
        o.backward(gc.astype(np.float64 if f64 else np.float32), go.astype(np.float64 if f64 else np.float32))
        return o.get("dL_drotations").astype(np.float64), o.get("radii")

    g64, r64 = grads(True)
    g32, r32 = grads(False)
    O.set_fp64_degenerate(False)
    try:
        gp, rp = grads(False)
    finally:
        O.set_fp64_degenerate(True)
    vis = (r64 > 0) & (r32 > 0) & (rp > 0)
    # a wrong sign turns the gradient around: (g . g64) < 0.  (Where the
    # signs agree the FP32 gradients of grazing surfels still carry the
    # ill-conditioning of the grazing homography, up to a few percent.)
    bad = lambda g: ((g * g64).sum(1) < 0)[vis]  # noqa: E731
    print(f"[degenerate] rotation gradients turned against the FP64 truth: FP64 branch {int(bad(g32).sum())}, "
          f"pure FP32 {int(bad(gp).sum())} of {int(vis.sum())}")
    assert bad(g32).sum() == 0
    assert bad(gp).sum() > 0
