"""Pins the oracle's frozen-decision replay and, through it, the backward on
dense scenes where plain finite differences cross thresholds.

A recorded forward keeps every discrete choice (culls, normal flips, radii
and tile lists, colour clamps, and per pixel which entries blend with which
branch and whether they set the median).  A replay on perturbed inputs holds
them, so its outputs are smooth in the inputs and their derivative is what
the backward computes (dsr/cuda_rasterizer/backward.cu:143-630 differentiates
with the same choices held).  Checks:
* a replay on the recorded inputs reproduces the recorded outputs bit for
  bit, in FP32 and FP64;
* on a dense scene (600 surfels, SH degree 3 with clamped channels) central
  differences of the FP64 replay equal the FP64 backward for means, scales,
  opacities, rotations (tangent part) and SH.  This pins the replay as the
  function the backward differentiates, so it can stand in for the forward
  in finite differences at sizes where a plain forward crosses thresholds
  within the step (the block-layer-through-rasteriser check at 65,536
  surfels);
* the densification statistic the reference writes into dL_dmeans2D
  (backward.cu:626-629) equals 0.5 W' dL/d(delta) for a fronto-parallel
  precomputed T shifted by delta pixels (Tu <- Tu + delta Tw), and likewise
  for y with Tv and H'.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from partgs_amd.camera import look_at
from partgs_amd.scenes import quat_from_matrix, random_rotations


def dense_scene(seed=0, P=600, W=64, H=48):
    rng = np.random.default_rng(seed)
    cam = look_at([0.4, -0.3, -3.0], width=W, height=H)
    means = rng.uniform(-0.6, 0.6, (P, 3))
    rots = quat_from_matrix(random_rotations(rng, P))
    scales = np.exp(rng.uniform(np.log(5e-3), np.log(0.2), (P, 2)))
    opac = rng.uniform(0.05, 0.9, (P, 1))
    shs = rng.normal(0.0, 0.15, (P, 16, 3))
    shs[:, 0, :] = rng.uniform(-2.5, 1.5, (P, 3))  # about a sixth of the channels clamp at 0
    gc = rng.normal(size=(3, H, W))
    go = rng.normal(size=(7, H, W))
    return dict(cam=cam, means3D=means, rotations=rots, scales=scales, opacities=opac, shs=shs, gc=gc, go=go)


def _kw(sc, f64=True, **over):
    cam = sc["cam"]
    p = {k: over.get(k, sc[k]) for k in ("means3D", "rotations", "scales", "opacities", "shs")}
    return dict(means3D=p["means3D"], opacities=p["opacities"], viewmatrix=cam.world_view_transform,
                projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=np.array([0.3, 0.5, 0.7]),
                W=cam.image_width, H=cam.image_height, tanfovx=math.tan(cam.FoVx / 2), tanfovy=math.tan(cam.FoVy / 2),
                shs=p["shs"], sh_degree=3, scales=p["scales"], rotations=p["rotations"], f64=f64)


def _loss(o, sc):
    return float((o.get("out_color") * sc["gc"]).sum() + (o.get("out_others") * sc["go"]).sum())


@pytest.mark.parametrize("f64", [False, True])
def test_replay_on_recorded_inputs_is_bit_identical(f64):
    sc = dense_scene(1)
    o = O.Oracle()
    o.forward(**_kw(sc, f64), record=True)
    ref = {k: o.get(k).copy() for k in ("out_color", "out_others", "final_T", "means2D", "transMat", "rgb")}
    o.replay(**_kw(sc, f64))
    for k, v in ref.items():
        np.testing.assert_array_equal(o.get(k), v, err_msg=k)


def test_replay_fd_equals_backward_on_dense_scene():
    sc = dense_scene(2)
    o = O.Oracle()
    o.forward(**_kw(sc), record=True)
    assert o.num_rendered > 1000 and o.alpha_capped == 0
    o.backward(sc["gc"], sc["go"])
    an = {k: o.get(k).copy() for k in ("dL_dmeans3D", "dL_dscales", "dL_dopacity", "dL_drotations", "dL_dsh")}
    rep = O.Oracle()
    rep.forward(**_kw(sc), record=True)
    rng = np.random.default_rng(3)
    vis = np.nonzero(o.get("radii") > 0)[0]
    picks = rng.choice(vis, size=24, replace=False)
    h = 1e-6
    for name, key in (("means3D", "dL_dmeans3D"), ("scales", "dL_dscales"), ("opacities", "dL_dopacity"),
                      ("rotations", "dL_drotations"), ("shs", "dL_dsh")):
        base = np.array(sc[name], dtype=np.float64)
        g_an = an[key]
        if name == "rotations":  # the backward omits the normalisation Jacobian: compare tangent parts
            q = base / np.linalg.norm(base, axis=1, keepdims=True)
            g_an = g_an - (g_an * q).sum(1, keepdims=True) * q
        fd_vals, an_vals = [], []
        for i in picks:
            for j in range(min(base[i].size, 6)):
                idx = np.unravel_index(j, base[i].shape)
                step = h * max(1.0, abs(base[i][idx]))
                vals = []
                for sgn in (1, -1):
                    x = base.copy()
                    x[i][idx] += sgn * step
                    rep.replay(**_kw(sc, **{name: x}))
                    vals.append(_loss(rep, sc))
                fd_vals.append((vals[0] - vals[1]) / (2 * step))
                an_vals.append(g_an[i][idx])
        fd_vals, an_vals = np.array(fd_vals), np.array(an_vals)
        scale = np.abs(fd_vals).max()
        assert scale > 0, name
        err = np.abs(fd_vals - an_vals)
        assert (err <= 1e-6 * scale + 1e-5 * np.abs(fd_vals)).all(), (name, err.max(), scale)


def _fronto_scene(seed=0, P=6, W=48, H=40):
    """Precomputed homographies of fronto-parallel surfels: Tw = (0, 0, z)."""
    rng = np.random.default_rng(seed)
    cam = look_at([0.0, 0.0, -3.0], width=W, height=H)
    tx, ty = math.tan(cam.FoVx / 2), math.tan(cam.FoVy / 2)
    fx, fy = W / (2 * tx), H / (2 * ty)
    T = np.zeros((P, 9))
    for i in range(P):
        z = rng.uniform(2.5, 3.5)
        cx, cy = rng.uniform(8, W - 8), rng.uniform(8, H - 8)
        su, sv = rng.uniform(0.1, 0.3, 2)
        th = rng.uniform(0, np.pi)
        # homogeneous pixel coordinates of the surfel's (u, v, 1) frame
        u_ax = np.array([math.cos(th), math.sin(th)]) * su
        v_ax = np.array([-math.sin(th), math.cos(th)]) * sv
        Tu = np.array([fx * u_ax[0], fx * v_ax[0], cx * z])
        Tv = np.array([fy * u_ax[1], fy * v_ax[1], cy * z])
        Tw = np.array([0.0, 0.0, z])
        T[i] = np.concatenate([Tu, Tv, Tw])
    return dict(cam=cam, T=T, P=P, colors=rng.uniform(0, 1, (P, 3)), opac=rng.uniform(0.3, 0.8, (P, 1)),
                means=rng.uniform(-0.2, 0.2, (P, 3)), gc=rng.normal(size=(3, H, W)), go=rng.normal(size=(7, H, W)))


def test_densification_statistic_is_screen_shift_gradient():
    sc = _fronto_scene(0)
    cam = sc["cam"]
    W, H = cam.image_width, cam.image_height
    tx, ty = math.tan(cam.FoVx / 2), math.tan(cam.FoVy / 2)

    def kw(T):
        return dict(means3D=sc["means"], opacities=sc["opac"], viewmatrix=cam.world_view_transform,
                    projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=np.array([1.0, 1.0, 1.0]),
                    W=W, H=H, tanfovx=tx, tanfovy=ty, colors_precomp=sc["colors"], transMat_precomp=T, f64=True)

    o = O.Oracle()
    o.forward(**kw(sc["T"]), record=True)
    assert o.alpha_capped == 0 and (o.get("radii") > 0).all()
    o.backward(sc["gc"], sc["go"])
    stat = o.get("dL_dmeans2D")
    # the backward's image size: int(focal * tanfov * 2) (backward.cu:607-608)
    Wp = int((W / (2 * tx)) * tx * 2)
    Hp = int((H / (2 * ty)) * ty * 2)
    rep = O.Oracle()
    rep.forward(**kw(sc["T"]), record=True)
    h = 1e-5
    for i in range(sc["P"]):
        for axis, (row, size) in enumerate(((0, Wp), (1, Hp))):
            vals = []
            for sgn in (1, -1):
                T = sc["T"].copy()
                T[i, 3 * row:3 * row + 3] += sgn * h * T[i, 6:9]  # screen shift by delta = +-h pixels
                rep.replay(**kw(T))
                vals.append(_loss(rep, sc))
            want = 0.5 * size * (vals[0] - vals[1]) / (2 * h)
            assert abs(stat[i, axis] - want) <= 1e-6 * (abs(want) + 1.0), (i, axis, stat[i, axis], want)
    assert np.abs(stat[:, :2]).max() > 1.0
