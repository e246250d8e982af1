"""Pins the oracle's _part mode (submodules/diff-surfel-rasterization_part,
the rasteriser renderer/gaussian_renderer_2d/__init__.py:86-96 calls).

* the camera-space homography's closed forms: a fronto-parallel surfel at
  view position (X, Y, Z) has centre (fx X/Z + W/2, fy Y/Z + H/2) and
  radius ceil(3 max(f sigma / Z, 0.7071)) (dsp/cuda_rasterizer/forward.cu:
  75-163, 235-239);
* against the plain rasteriser (pinned on its own): the same scene renders
  to the same images up to rounding, the pixel centres of the _part sitting
  half a pixel further (its u = plain u + 0.5);
* the semantic blend (forward.cu:440, 460): semantics equal to the colours
  with a black background give the colour image bit for bit; all-ones
  semantics give the alpha channel;
* the median-weight channel (:402-404, 471) of a single surfel is its alpha;
* the backward (dsp/cuda_rasterizer/backward.cu): FP64 central differences
  of the FP64 forward, loss over colour, semantic and all 8 aux channels
  (the semantic term for the semantics' own gradient only, as the reference
  computes it).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from partgs_amd.camera import look_at
from partgs_amd.scenes import quat_from_matrix, random_rotations


def _kw(cam, means, scales, rots, opac, colors=None, shs=None, deg=0, f64=True, bg=(0.3, 0.5, 0.7), tile=16,
        transmat=None):
    kw = dict(means3D=means, opacities=opac, viewmatrix=cam.world_view_transform,
              projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=np.array(bg),
              W=cam.image_width, H=cam.image_height, tanfovx=math.tan(cam.FoVx / 2),
              tanfovy=math.tan(cam.FoVy / 2), f64=f64, tile=tile)
    if colors is not None:
        kw["colors_precomp"] = colors
    else:
        kw.update(shs=shs, sh_degree=deg)
    if transmat is not None:
        kw["transMat_precomp"] = transmat
    else:
        kw.update(scales=scales, rotations=rots)
    return kw


def scene(seed=0, P=10, W=40, H=32, S=5, scale_range=(0.08, 0.25)):
    rng = np.random.default_rng(seed)
    cam = look_at([0.6, -0.5, -3.0], width=W, height=H)
    return dict(cam=cam, means3D=rng.uniform(-0.45, 0.45, (P, 3)),
                rotations=quat_from_matrix(random_rotations(rng, P)),
                scales=rng.uniform(*scale_range, (P, 2)), opacities=rng.uniform(0.3, 0.85, (P, 1)),
                shs=np.concatenate([rng.uniform(0.0, 1.2, (P, 1, 3)), rng.normal(0.0, 0.15, (P, 15, 3))], 1),
                colors=rng.uniform(0.0, 1.0, (P, 3)), sem=rng.uniform(-1.0, 1.0, (P, S)),
                gc=rng.normal(size=(3, H, W)), gs=rng.normal(size=(S, H, W)), go=rng.normal(size=(8, H, W)))


def run_part(sc, f64=True, use_colors=False, transmat=None, **over):
    p = {k: over.get(k, sc[k]) for k in ("means3D", "rotations", "scales", "opacities", "shs", "colors", "sem")}
    o = O.Oracle()
    kw = _kw(sc["cam"], p["means3D"], p["scales"], p["rotations"], p["opacities"],
             colors=p["colors"] if use_colors else None, shs=None if use_colors else p["shs"], deg=3, f64=f64,
             transmat=over.get("transmat", transmat))
    assert o.forward_part(semantics=p["sem"], **kw) == 0
    return o


def test_fronto_parallel_centre_and_radius():
    W, H = 64, 48
    cam = look_at([0.0, 0.0, -3.0], width=W, height=H)
    focal_x = W / (2 * math.tan(cam.FoVx / 2))
    focal_y = H / (2 * math.tan(cam.FoVy / 2))
    rng = np.random.default_rng(0)
    P = 6
    view = np.stack([rng.uniform(-0.4, 0.4, P), rng.uniform(-0.3, 0.3, P), rng.uniform(2.5, 3.5, P)], 1)
    c2w = np.linalg.inv(cam.world_view_transform.numpy().T.astype(np.float64))
    means = (np.concatenate([view, np.ones((P, 1))], 1) @ c2w.T)[:, :3]
    # rotation whose z axis is the view axis: camera-to-world rotation itself
    R = np.repeat(c2w[None, :3, :3], P, 0)
    sig = rng.uniform(0.02, 0.1, (P, 2))
    o = O.Oracle()
    o.forward_part(semantics=np.zeros((P, 0)), **_kw(cam, means, sig, quat_from_matrix(R), np.full((P, 1), 0.5),
                                                     colors=np.ones((P, 3))))
    m2 = o.get("means2D")
    np.testing.assert_allclose(m2[:, 0], focal_x * view[:, 0] / view[:, 2] + W / 2, atol=1e-9)
    np.testing.assert_allclose(m2[:, 1], focal_y * view[:, 1] / view[:, 2] + H / 2, atol=1e-9)
    ext = np.maximum(focal_x * sig[:, 0], focal_y * sig[:, 1]) / view[:, 2]
    np.testing.assert_array_equal(o.get("radii"), np.ceil(3 * np.maximum(ext, 0.7071067811865476)).astype(int))
    # the plain rasteriser puts the same surfel half a pixel lower
    plain = O.Oracle()
    plain.forward(**_kw(cam, means, sig, quat_from_matrix(R), np.full((P, 1), 0.5), colors=np.ones((P, 3))))
    # (the plain path reads the FP32 projection matrix, this one the focal length)
    np.testing.assert_allclose(m2 - plain.get("means2D"), 0.5, atol=1e-5)


def test_images_agree_with_plain_rasteriser():
    """The same scene through both variants renders the same images up to
    rounding: the camera-space homography with pixel centres at +0.5 is the
    plain one's pixel-space homography shifted by half a pixel.  Surfels
    face the camera (then both low-pass centres are the projected centre)
    with opacity below 0.3 (then nothing outside a tile rectangle reaches
    1/255, so the half-pixel shift of the rectangles changes nothing)."""
    sc = scene(1, P=40, W=64, H=48, scale_range=(0.1, 0.3))
    c2w = np.linalg.inv(sc["cam"].world_view_transform.numpy().T.astype(np.float64))
    sc["rotations"] = quat_from_matrix(np.repeat(c2w[None, :3, :3], 40, 0))
    sc["opacities"] = sc["opacities"] * 0.35
    part = run_part(sc)
    plain = O.Oracle()
    plain.forward(**_kw(sc["cam"], sc["means3D"], sc["scales"], sc["rotations"], sc["opacities"], shs=sc["shs"],
                        deg=3))
    np.testing.assert_allclose(part.get("out_color"), plain.get("out_color"), atol=1e-6)
    np.testing.assert_allclose(part.get("out_others")[:7], plain.get("out_others"), atol=1e-5)
    assert part.get("out_others")[1].max() > 0.2


def test_semantic_blend_equals_colour_blend():
    sc = scene(2, P=30, W=48, H=40)
    sem = sc["colors"].copy()
    o = O.Oracle()
    o.forward_part(semantics=sem, **_kw(sc["cam"], sc["means3D"], sc["scales"], sc["rotations"], sc["opacities"],
                                        colors=sc["colors"], f64=False, bg=(0.0, 0.0, 0.0)))
    np.testing.assert_array_equal(o.get("out_semantic"), o.get("out_color"))
    ones = O.Oracle()
    ones.forward_part(semantics=np.ones((30, 2)), **_kw(sc["cam"], sc["means3D"], sc["scales"], sc["rotations"],
                                                         sc["opacities"], colors=sc["colors"], f64=True))
    alpha = ones.get("out_others")[1]
    np.testing.assert_allclose(ones.get("out_semantic"), np.stack([alpha, alpha]), atol=1e-12)
    assert alpha.max() > 0.5


def test_median_weight_of_single_surfel_is_alpha():
    sc = scene(3, P=1, W=40, H=32, scale_range=(0.2, 0.3))
    o = run_part(sc, use_colors=True)
    others = o.get("out_others")
    np.testing.assert_allclose(others[7], others[1], atol=1e-12)
    assert others[1].max() > 0.1


def _loss(o, sc, S=True):
    v = float((o.get("out_color") * sc["gc"]).sum() + (o.get("out_others") * sc["go"]).sum())
    if S:
        v += float((o.get("out_semantic") * sc["gs"]).sum())
    return v


def _fd(sc, name, base, h=1e-6, S=False, **kw):
    """Central differences of the loss.  The semantic term enters only for
    the semantics themselves: the reference does not carry the semantic
    images' gradient into the geometry (its alpha term is commented out,
    dsp/cuda_rasterizer/backward.cu:346), so the geometry gradients are
    those of the colour and aux terms."""
    g = np.zeros_like(base, dtype=np.float64)
    flat = base.reshape(-1)
    for i in range(flat.size):
        step = h * max(1.0, abs(flat[i]))
        vals = []
        for sgn in (1, -1):
            x = flat.copy()
            x[i] += sgn * step
            vals.append(_loss(run_part(sc, **kw, **{name: x.reshape(base.shape)}), sc, S=S))
        g.reshape(-1)[i] = (vals[0] - vals[1]) / (2 * step)
    return g


def _check(fd, an, name, tol=1e-6):
    scale = np.abs(fd).max()
    assert scale > 0, name
    err = np.abs(fd - an)
    assert (err <= tol * scale + 1e-5 * np.abs(fd)).all(), f"{name}: max err {err.max():.3e} of {scale:.3e}"


@pytest.mark.parametrize("seed", [0, 4])
def test_backward_matches_fd64(seed):
    sc = scene(seed)
    o = run_part(sc)
    o.backward_part(sc["gc"], sc["gs"], sc["go"])
    _check(_fd(sc, "means3D", sc["means3D"]), o.get("dL_dmeans3D"), "means3D")
    _check(_fd(sc, "scales", sc["scales"]), o.get("dL_dscales"), "scales")
    _check(_fd(sc, "opacities", sc["opacities"]), o.get("dL_dopacity"), "opacities")
    _check(_fd(sc, "shs", sc["shs"]), o.get("dL_dsh"), "shs")
    _check(_fd(sc, "sem", sc["sem"], S=True), o.get("dL_dsemantics"), "semantics")
    q = sc["rotations"]
    g = o.get("dL_drotations")
    _check(_fd(sc, "rotations", q), g - (g * q).sum(1, keepdims=True) * q, "rotations")


def test_backward_precomputed_colours_and_transmat():
    sc = scene(5, P=6)
    o = run_part(sc, use_colors=True)
    o.backward_part(sc["gc"], sc["gs"], sc["go"])
    _check(_fd(sc, "colors", sc["colors"], use_colors=True), o.get("dL_dcolors"), "colors")
    tm = run_part(sc).get("transMat")
    o = run_part(sc, use_colors=True, transmat=tm)
    o.backward_part(sc["gc"], sc["gs"], sc["go"])
    _check(_fd(sc, "transmat", tm, use_colors=True), o.get("dL_dtransMat"), "transMat")


def test_low_pass_branch_gradients():
    """Sub-pixel surfels take the screen-space filter (rho2d < rho3d); its
    gradient reaches the surfel through the AABB-centre adjoint."""
    sc = scene(6, P=8, scale_range=(5e-4, 2e-3))
    o = run_part(sc)
    o.backward_part(sc["gc"], sc["gs"], sc["go"])
    assert np.abs(o.get("acc_dL_dmean2D")).max() > 0
    _check(_fd(sc, "means3D", sc["means3D"]), o.get("dL_dmeans3D"), "means3D")
    _check(_fd(sc, "opacities", sc["opacities"]), o.get("dL_dopacity"), "opacities")


def test_fp32_close_to_fp64():
    sc = scene(7, P=12, W=48, H=40)
    o64 = run_part(sc)
    o32 = run_part(sc, f64=False)
    for k in ("out_color", "out_semantic", "out_others"):
        np.testing.assert_allclose(o32.get(k), o64.get(k), atol=2e-5)
    o64.backward_part(sc["gc"], sc["gs"], sc["go"])
    o32.backward_part(sc["gc"], sc["gs"], sc["go"])
    for n in ("dL_dmeans3D", "dL_dscales", "dL_drotations", "dL_dopacity", "dL_dsh", "dL_dsemantics"):
        a, b = o32.get(n).astype(np.float64), o64.get(n)
        assert np.abs(a - b).max() <= 1e-3 * np.abs(b).max(), n
