"""Pins the oracle's backward: FP64 central differences of the FP64 forward.

loss = <dL_dcolor, color> + <dL_dothers, others> over all ten output
channels, so every adjoint path (colour, depth, alpha, normal, median depth,
distortion, the screen-space low-pass branch and the AABB-centre adjoint) is
exercised.  The FP32 backward is then checked against the FP64 one.

Reference quirks kept on purpose and accounted for here:
* the rotation gradient is that of the normalised quaternion without the
  normalisation Jacobian (dsr/cuda_rasterizer/auxiliary.h:238-282); central
  differences see only its component tangent to the unit sphere;
* dL_dmeans2D is overwritten with a densification statistic
  (dsr/cuda_rasterizer/backward.cu:626-629) and is not compared.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from partgs_amd.camera import look_at
from partgs_amd.scenes import quat_from_matrix, random_rotations


def make_scene(seed=0, P=8, W=40, H=32, deg=3, scale_range=(0.08, 0.25)):
    rng = np.random.default_rng(seed)
    cam = look_at([0.6, -0.5, -3.0], width=W, height=H)
    means = rng.uniform(-0.45, 0.45, (P, 3))
    rots = quat_from_matrix(random_rotations(rng, P))
    scales = rng.uniform(scale_range[0], scale_range[1], (P, 2))
    opac = rng.uniform(0.3, 0.85, (P, 1))
    M = (deg + 1) ** 2
    shs = rng.normal(0.0, 0.15, (P, M, 3))
    shs[:, 0, :] = rng.uniform(0.0, 1.2, (P, 3))
    colors = rng.uniform(0.0, 1.0, (P, 3))
    gc = rng.normal(size=(3, H, W))
    go = rng.normal(size=(7, H, W))
    return dict(cam=cam, means3D=means, rotations=rots, scales=scales, opacities=opac, shs=shs, deg=deg,
                colors=colors, gc=gc, go=go)


def forward(sc, f64=True, use_colors=False, tm_base=None, **over):
    cam = sc["cam"]
    p = {k: over.get(k, sc[k]) for k in ("means3D", "rotations", "scales", "opacities", "shs", "colors")}
    o = O.Oracle()
    kw = dict(means3D=p["means3D"], opacities=p["opacities"], viewmatrix=cam.world_view_transform,
              projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=np.array([0.3, 0.5, 0.7]),
              W=cam.image_width, H=cam.image_height, tanfovx=math.tan(cam.FoVx / 2),
              tanfovy=math.tan(cam.FoVy / 2), f64=f64)
    if use_colors:
        kw["colors_precomp"] = p["colors"]
    else:
        kw.update(shs=p["shs"], sh_degree=sc["deg"])
    if tm_base is not None:
        kw["transMat_precomp"] = over.get("transmat", tm_base)
    else:
        kw.update(scales=p["scales"], rotations=p["rotations"])
    assert o.forward(**kw) == 0
    return o


def loss(o, sc):
    return float((o.get("out_color") * sc["gc"]).sum() + (o.get("out_others") * sc["go"]).sum())


def fd_grad(sc, name, shape_src, h=1e-6, **kw):
    base = np.array(shape_src, dtype=np.float64)
    g = np.zeros_like(base)
    flat = base.reshape(-1)
    for i in range(flat.size):
        step = h * max(1.0, abs(flat[i]))
        xp = flat.copy()
        xp[i] += step
        xm = flat.copy()
        xm[i] -= step
        lp = loss(forward(sc, **kw, **{name: xp.reshape(base.shape)}), sc)
        lm = loss(forward(sc, **kw, **{name: xm.reshape(base.shape)}), sc)
        g.reshape(-1)[i] = (lp - lm) / (2 * step)
    return g


def analytic(sc, **kw):
    o = forward(sc, **kw)
    o.backward(sc["gc"], sc["go"])
    return o


def assert_grad(fd, an, name, tol=1e-6):
    fd = np.asarray(fd, np.float64)
    an = np.asarray(an, np.float64)
    scale = np.abs(fd).max()
    assert scale > 0, name
    err = np.abs(fd - an)
    bad = err > tol * scale + 1e-5 * np.abs(fd)
    assert not bad.any(), f"{name}: max err {err.max():.3e} (scale {scale:.3e}), {bad.sum()} bad"


@pytest.mark.parametrize("seed", [0, 1])
def test_fd_means_scales_opacity_sh(seed):
    sc = make_scene(seed)
    o = analytic(sc)
    assert o.num_rendered > 0
    assert_grad(fd_grad(sc, "means3D", sc["means3D"]), o.get("dL_dmeans3D"), "means3D")
    assert_grad(fd_grad(sc, "scales", sc["scales"]), o.get("dL_dscales"), "scales")
    assert_grad(fd_grad(sc, "opacities", sc["opacities"]), o.get("dL_dopacity"), "opacities")
    assert_grad(fd_grad(sc, "shs", sc["shs"]), o.get("dL_dsh"), "shs")


def test_fd_screen_space_filter_branch():
    """Sub-pixel surfels blend through the low-pass branch min(rho3d, rho2d) =
    rho2d, whose gradient reaches the surfel through the AABB-centre adjoint."""
    sc = make_scene(6, P=10, scale_range=(5e-4, 2e-3))
    o = analytic(sc)
    assert np.abs(o.get("acc_dL_dmean2D")).max() > 0  # the low-pass branch was taken
    assert_grad(fd_grad(sc, "means3D", sc["means3D"]), o.get("dL_dmeans3D"), "means3D")
    assert_grad(fd_grad(sc, "opacities", sc["opacities"]), o.get("dL_dopacity"), "opacities")
    fd = fd_grad(sc, "rotations", sc["rotations"])
    g = o.get("dL_drotations")
    q = sc["rotations"]
    assert_grad(fd, g - (g * q).sum(1, keepdims=True) * q, "rotations")


def test_fd_rotations_tangent():
    sc = make_scene(2)
    o = analytic(sc)
    g = o.get("dL_drotations")
    q = sc["rotations"]
    fd = fd_grad(sc, "rotations", q)
    # project the analytic gradient onto the tangent space of the unit quaternion
    g_t = g - (g * q).sum(1, keepdims=True) * q
    assert_grad(fd, g_t, "rotations")
    # and the finite differences see nothing along q itself
    assert np.abs((fd * q).sum(1)).max() < 1e-6 * np.abs(fd).max()


def test_fd_precomputed_colors():
    sc = make_scene(3)
    o = analytic(sc, use_colors=True)
    assert_grad(fd_grad(sc, "colors", sc["colors"], use_colors=True), o.get("dL_dcolors"), "colors")
    assert_grad(fd_grad(sc, "means3D", sc["means3D"], use_colors=True), o.get("dL_dmeans3D"), "means3D")


def test_fd_precomputed_transmat():
    sc = make_scene(4, P=6)
    tm = forward(sc).get("transMat")
    o = analytic(sc, use_colors=True, tm_base=tm)
    fd = fd_grad(sc, "transmat", tm, use_colors=True, tm_base=tm)
    assert_grad(fd, o.get("dL_dtransMat"), "transMat")


def test_fp32_backward_close_to_fp64():
    sc = make_scene(5, P=12, W=48, H=40)
    o64 = analytic(sc)
    o32 = analytic(sc, f64=False)
    for n in ("dL_dmeans3D", "dL_dscales", "dL_drotations", "dL_dopacity", "dL_dsh", "dL_dtransMat"):
        a, b = o32.get(n).astype(np.float64), o64.get(n)
        assert np.abs(a - b).max() <= 1e-3 * np.abs(b).max(), n
