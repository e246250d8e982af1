"""Pins the CPU oracle's forward against mathematics and golden vectors.

Nothing here retypes the oracle's formulas: each check derives the expected
value independently (geometry of the ray/plane intersection, the projection
of a fronto-parallel disk, orthonormality of the SH basis, brute-force
sorting, closed-form compositing) or reads vectors written by the
reference's own Python functions (tests/golden).
"""
import math
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from partgs_amd.camera import camera_from_Rt, look_at
from partgs_amd.scenes import random_rotations, surfel_cloud

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def quat_of(R):
    from partgs_amd.scenes import quat_from_matrix

    return quat_from_matrix(R[None])[0]


def run(cam, means, scales, rots, opac, colors=None, shs=None, deg=0, f64=True, tile=16, bg=(1.0, 1.0, 1.0),
        mod=1.0):
    o = O.Oracle()
    kw = dict(means3D=np.asarray(means, np.float64), opacities=np.asarray(opac, np.float64).reshape(-1, 1),
              viewmatrix=cam.world_view_transform, projmatrix=cam.full_proj_transform, campos=cam.camera_center,
              bg=np.asarray(bg, np.float64), W=cam.image_width, H=cam.image_height,
              tanfovx=math.tan(cam.FoVx / 2), tanfovy=math.tan(cam.FoVy / 2), scales=np.asarray(scales),
              rotations=np.asarray(rots), f64=f64, tile=tile, scale_modifier=mod)
    if colors is not None:
        kw["colors_precomp"] = np.asarray(colors, np.float64)
    else:
        kw["shs"] = shs
        kw["sh_degree"] = deg
    assert o.forward(**kw) == 0
    return o


def axis_camera(W=64, H=48, fov=math.radians(60)):
    return camera_from_Rt(np.eye(3), np.zeros(3), fov, fov, W, H)


def test_expf_within_one_ulp():
    x = np.linspace(-87.0, 0.0, 1_000_001).astype(np.float32)
    x = np.concatenate([x, -np.logspace(-8, 1.9, 200_000).astype(np.float32), np.float32([0.0, -0.0])])
    y = O.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert (np.abs(y - ref) / ulp).max() <= 1.0
    assert O.expf([0.0])[0] == 1.0
    assert O.expf([-90.0])[0] == 0.0


def test_fronto_parallel_centre_and_extent():
    W, H = 96, 64
    cam = axis_camera(W, H)
    rng = np.random.default_rng(0)
    P = 40
    z = rng.uniform(1.0, 5.0, P)
    means = np.stack([rng.uniform(-0.3, 0.3, P) * z, rng.uniform(-0.2, 0.2, P) * z, z], 1)
    scales = np.exp(rng.uniform(math.log(0.01), math.log(0.1), (P, 2)))
    rots = np.tile([1.0, 0.0, 0.0, 0.0], (P, 1))
    o = run(cam, means, scales, rots, np.full(P, 0.5), colors=np.full((P, 3), 0.5))
    # projection through the camera's own (float32) matrices, evaluated in FP64
    Pm = cam.full_proj_transform.numpy().astype(np.float64)
    clip = np.concatenate([means, np.ones((P, 1))], 1) @ Pm
    ndc = clip[:, :2] / clip[:, 3:4]
    cx = ((ndc[:, 0] + 1) * W - 1) / 2
    cy = ((ndc[:, 1] + 1) * H - 1) / 2
    fx = W / 2 * Pm[0, 0]
    fy = H / 2 * Pm[1, 1]
    m2 = o.get("means2D")
    np.testing.assert_allclose(m2[:, 0], cx, rtol=0, atol=1e-9)
    np.testing.assert_allclose(m2[:, 1], cy, rtol=0, atol=1e-9)
    ext = np.maximum(3 * scales[:, 0] * fx / z, 3 * scales[:, 1] * fy / z)
    ext = np.maximum(ext, 3.0 * 0.707106)
    np.testing.assert_array_equal(o.get("radii"), np.ceil(ext - 1e-9).astype(np.int32))
    np.testing.assert_allclose(o.get("depths"), z, rtol=1e-12)
    # the normal faces the camera: (0, 0, -1) in view space
    np.testing.assert_allclose(o.get("normal_opacity")[:, :3], np.tile([0.0, 0.0, -1.0], (P, 1)), atol=1e-12)


def _pixel_ray(cam, px, py):
    """World ray through integer pixel (px, py), inverting the camera's own matrices."""
    W, H = cam.image_width, cam.image_height
    xn = (2 * px + 1) / W - 1
    yn = (2 * py + 1) / H - 1
    Pm = cam.projection_matrix.numpy().astype(np.float64)  # row-vector convention
    d_cam = np.array([xn / Pm[0, 0], yn / Pm[1, 1], 1.0])
    w2c = cam.world_view_transform.numpy().astype(np.float64).T
    c2w = np.linalg.inv(w2c)
    return c2w[:3, 3], c2w[:3, :3] @ d_cam, w2c


def test_homography_matches_ray_plane_intersection():
    rng = np.random.default_rng(1)
    cam = look_at([0.5, -2.5, 1.5], width=80, height=60)
    P = 25
    means = rng.uniform(-0.4, 0.4, (P, 3))
    Rs = random_rotations(rng, P)
    rots = np.stack([quat_of(R) for R in Rs])
    scales = rng.uniform(0.05, 0.2, (P, 2))
    mod = 1.3
    o = run(cam, means, scales, rots, np.full(P, 0.5), colors=np.full((P, 3), 0.5), mod=mod)
    Tm = o.get("transMat")
    radii = o.get("radii")
    checked = 0
    for i in range(P):
        if radii[i] == 0:
            continue
        Tu, Tv, Tw = Tm[i, 0:3], Tm[i, 3:6], Tm[i, 6:9]
        centre = o.get("means2D")[i]
        for px, py in centre + rng.uniform(-1.0, 1.0, (6, 2)) * radii[i]:
            k = px * Tw - Tu
            l = py * Tw - Tv
            p = np.cross(k, l)
            s = p[:2] / p[2]
            orig, d, w2c = _pixel_ray(cam, px, py)
            n = Rs[i][:, 2]
            t = np.dot(means[i] - orig, n) / np.dot(d, n)
            X = orig + t * d
            u = np.dot(X - means[i], Rs[i][:, 0]) / (mod * scales[i, 0])
            v = np.dot(X - means[i], Rs[i][:, 1]) / (mod * scales[i, 1])
            if max(abs(u), abs(v)) > 10.0:  # grazing rays amplify the float32 matrix rounding
                continue
            np.testing.assert_allclose(s, [u, v], rtol=1e-6, atol=1e-6)
            z_view = (w2c @ np.append(X, 1.0))[2]
            np.testing.assert_allclose(s[0] * Tw[0] + s[1] * Tw[1] + Tw[2], z_view, rtol=1e-6)
            checked += 1
        # camera-facing view-space normal
        nv = w2c[:3, :3] @ Rs[i][:, 2]
        pv = (w2c @ np.append(means[i], 1.0))[:3]
        if np.dot(pv, nv) > 0:
            nv = -nv
        np.testing.assert_allclose(o.get("normal_opacity")[i, :3], nv, atol=1e-12)
    assert checked > 50


def test_rect_counts_match_tile_overlap():
    sc = surfel_cloud(2, P=3000, width=150, height=100)
    for tile in (8, 16, 32):
        o = O.run_scene(sc, tile=tile)
        r = o.get("radii").astype(np.float32)
        m = o.get("means2D")
        gx, gy = -(-150 // tile), -(-100 // tile)
        B = np.float32(tile)

        def span(c, g):
            lo = np.clip(np.floor((c - r) / B), 0, g)
            hi = np.clip(np.floor(((c + r) + B - np.float32(1)) / B), 0, g)
            return hi - lo

        expect = (span(m[:, 0], gx) * span(m[:, 1], gy)).astype(np.int64)
        expect[r == 0] = 0
        np.testing.assert_array_equal(o.get("tiles_touched"), expect)
        assert ((o.get("radii") > 0) == (o.get("tiles_touched") > 0)).all()


def test_binning_brute_force():
    sc = surfel_cloud(3, P=4000, width=120, height=90)
    for tile in (8, 16):
        o = O.run_scene(sc, tile=tile)
        radii, m, d = o.get("radii"), o.get("means2D"), o.get("depths")
        gx, gy = -(-120 // tile), -(-90 // tile)
        B = np.float32(tile)
        tiles, dbits, ids = [], [], []
        for i in np.nonzero(radii > 0)[0]:
            r = np.float32(radii[i])
            x0 = int(min(gx, max(0, np.floor((m[i, 0] - r) / B))))
            x1 = int(min(gx, max(0, np.floor(((m[i, 0] + r) + B - np.float32(1)) / B))))
            y0 = int(min(gy, max(0, np.floor((m[i, 1] - r) / B))))
            y1 = int(min(gy, max(0, np.floor(((m[i, 1] + r) + B - np.float32(1)) / B))))
            for y in range(y0, y1):
                for x in range(x0, x1):
                    tiles.append(y * gx + x)
                    dbits.append(np.float32(d[i]).view(np.uint32))
                    ids.append(i)
        tiles, dbits, ids = np.array(tiles), np.array(dbits, np.uint64), np.array(ids)
        order = np.lexsort((np.arange(len(ids)), dbits, tiles))
        keys = (tiles[order].astype(np.uint64) << np.uint64(32)) | dbits[order]
        assert o.num_rendered == len(ids)
        np.testing.assert_array_equal(o.get("keys"), keys)
        np.testing.assert_array_equal(o.get("point_list"), ids[order])
        np.testing.assert_array_equal(o.get("point_offsets"), np.cumsum(o.get("tiles_touched")))
        ranges = o.get("ranges")
        st = tiles[order]
        for t in range(gx * gy):
            a, b = np.searchsorted(st, t, "left"), np.searchsorted(st, t, "right")
            exp = (a, b) if b > a else (0, 0)
            assert tuple(ranges[t]) == exp


def test_higher_msb_covers_tile_ids():
    for n in list(range(1, 70)) + [255, 256, 257, 1900, 4800, 65535, 65536, 1 << 20]:
        b = O.higher_msb(n)
        assert (n - 1) >> b == 0, (n, b)


def _geom_alpha(cam, mean, R, s, opa, px, py, centre2d=None):
    """Alpha and depth at pixel (px, py) from the ray/plane intersection; with
    centre2d, the screen-space low-pass filter min(rho3d, 2 |pix - c|^2) of the
    reference (depth then falls back to the centre's view depth)."""
    orig, d, w2c = _pixel_ray(cam, px, py)
    n = R[:, 2]
    t = np.dot(mean - orig, n) / np.dot(d, n)
    X = orig + t * d
    u = np.dot(X - mean, R[:, 0]) / s[0]
    v = np.dot(X - mean, R[:, 1]) / s[1]
    rho = u * u + v * v
    z = (w2c @ np.append(X, 1.0))[2]
    if centre2d is not None:
        rho2d = 2.0 * ((centre2d[0] - px) ** 2 + (centre2d[1] - py) ** 2)
        if rho2d < rho:
            rho = rho2d
            z = (w2c @ np.append(mean, 1.0))[2]
    return min(0.99, opa * math.exp(-0.5 * rho)), z, rho


def test_screen_aabb_of_projected_disk():
    """means2D / radius are the centre and half-size of the screen bounding box
    of the projected 3-sigma ellipse (sampled boundary, any orientation)."""
    rng = np.random.default_rng(7)
    cam = look_at([0.4, 0.3, -3.0], width=120, height=90)
    P = 30
    means = rng.uniform(-0.5, 0.5, (P, 3))
    Rs = random_rotations(rng, P)
    rots = np.stack([quat_of(R) for R in Rs])
    scales = rng.uniform(0.03, 0.15, (P, 2))
    o = run(cam, means, scales, rots, np.full(P, 0.5), colors=np.full((P, 3), 0.5))
    Pm = cam.full_proj_transform.numpy().astype(np.float64)
    W, H = cam.image_width, cam.image_height
    th = np.linspace(0, 2 * np.pi, 20001)
    checked = 0
    for i in range(P):
        R = Rs[i]
        w2c = cam.world_view_transform.numpy().astype(np.float64).T
        nv = w2c[:3, :3] @ R[:, 2]
        pv = (w2c @ np.append(means[i], 1.0))[:3]
        if abs(np.dot(nv, pv / np.linalg.norm(pv))) < 0.3 or o.get("radii")[i] == 0:
            continue  # grazing views: bounding box of a strongly foreshortened conic
        pts = means[i] + 3 * (np.outer(np.cos(th), R[:, 0] * scales[i, 0]) + np.outer(np.sin(th), R[:, 1] * scales[i, 1]))
        clip = np.concatenate([pts, np.ones((len(pts), 1))], 1) @ Pm
        x = ((clip[:, 0] / clip[:, 3] + 1) * W - 1) / 2
        y = ((clip[:, 1] / clip[:, 3] + 1) * H - 1) / 2
        cx, cy = (x.max() + x.min()) / 2, (y.max() + y.min()) / 2
        hx, hy = (x.max() - x.min()) / 2, (y.max() - y.min()) / 2
        np.testing.assert_allclose(o.get("means2D")[i], [cx, cy], atol=2e-3)
        np.testing.assert_allclose(max(hx, hy), np.ceil(max(hx, hy)) - (np.ceil(max(hx, hy)) - max(hx, hy)), atol=0)
        assert o.get("radii")[i] == max(3, math.ceil(max(hx, hy) - 1e-6)), (i, hx, hy)
        checked += 1
    assert checked >= 10


def test_single_large_surfel_closed_form():
    cam = look_at([0.3, -0.4, -3.0], width=48, height=40)
    rng = np.random.default_rng(4)
    R = random_rotations(rng, 1)[0]
    if np.dot(R[:, 2], [0.3, -0.4, -3.0]) < 0:
        R[:, 1:] *= -1  # keep a right-handed frame whose normal faces the camera
    q = quat_of(R)
    s = np.array([0.35, 0.25])
    c = np.array([0.2, 0.6, 0.9])
    opa = 0.8
    o = run(cam, [[0.0, 0.0, 0.0]], [s], [q], [opa], colors=[c], bg=(0.1, 0.2, 0.3))
    img, oth = o.get("out_color"), o.get("out_others")
    nv = o.get("normal_opacity")[0, :3]
    c2, r = o.get("means2D")[0], int(o.get("radii")[0])
    # a surfel is blended only inside the tiles its screen box [c - r, c + r] overlaps
    tx0, tx1 = int((c2[0] - r) // 16), int((c2[0] + r + 15) // 16)
    ty0, ty1 = int((c2[1] - r) // 16), int((c2[1] + r + 15) // 16)
    hits = 0
    for py in range(40):
        for px in range(48):
            a, z, rho = _geom_alpha(cam, np.zeros(3), R, s, opa, px, py, centre2d=c2)
            if not (tx0 <= px // 16 < tx1 and ty0 <= py // 16 < ty1):
                a = 0.0
            if abs(a - 1 / 255) < 1e-5:
                continue  # threshold decided by float32-level matrix rounding
            if a < 1 / 255:
                np.testing.assert_allclose(img[:, py, px], [0.1, 0.2, 0.3], atol=1e-12)
                continue
            hits += 1
            np.testing.assert_allclose(img[:, py, px], a * c + (1 - a) * np.array([0.1, 0.2, 0.3]), rtol=1e-6, atol=1e-9)
            np.testing.assert_allclose(oth[0, py, px], a * z, rtol=1e-6, atol=1e-9)
            np.testing.assert_allclose(oth[1, py, px], a, rtol=1e-6, atol=1e-9)
            np.testing.assert_allclose(oth[2:5, py, px], a * nv, rtol=1e-6, atol=1e-9)
            np.testing.assert_allclose(oth[5, py, px], z, rtol=1e-6, atol=1e-9)
            assert oth[6, py, px] == 0.0
    assert hits > 150


def test_small_surfel_uses_screen_space_filter():
    """A sub-pixel fronto-parallel surfel blends with exp(-|pix - centre|^2)."""
    W, H = 32, 32
    cam = axis_camera(W, H)
    mean = np.array([0.013, -0.021, 2.0])
    o = run(cam, [mean], [[1e-4, 1e-4]], [[1, 0, 0, 0]], [0.9], colors=[[1.0, 0.0, 0.0]], bg=(0, 0, 0))
    Pm = cam.full_proj_transform.numpy().astype(np.float64)
    clip = np.append(mean, 1.0) @ Pm
    cx = ((clip[0] / clip[3] + 1) * W - 1) / 2
    cy = ((clip[1] / clip[3] + 1) * H - 1) / 2
    img = o.get("out_color")
    for py in range(H):
        for px in range(W):
            a = min(0.99, 0.9 * math.exp(-((px - cx) ** 2 + (py - cy) ** 2)))
            exp = a if a >= 1 / 255 else 0.0
            np.testing.assert_allclose(img[0, py, px], exp, rtol=1e-9, atol=1e-12)


def test_two_surfel_distortion_and_median():
    W, H = 32, 32
    cam = axis_camera(W, H)
    z1, z2 = 1.5, 3.0
    means = [[0.0, 0.0, z1], [0.0, 0.0, z2]]
    sc = [[0.3, 0.3], [0.6, 0.6]]
    for o1, o2 in ((0.3, 0.7), (0.8, 0.6)):
        o = run(cam, means, sc, [[1, 0, 0, 0]] * 2, [o1, o2], colors=[[1, 0, 0], [0, 1, 0]], bg=(0, 0, 0))
        oth = o.get("out_others")
        px = py = 15
        a1, d1, _ = _geom_alpha(cam, np.array(means[0]), np.eye(3), np.array(sc[0]), o1, px, py)
        a2, d2, _ = _geom_alpha(cam, np.array(means[1]), np.eye(3), np.array(sc[1]), o2, px, py)
        w1, w2 = a1, a2 * (1 - a1)
        m = lambda z: 100 / (100 - 0.2) * (1 - 0.2 / z)  # noqa: E731
        np.testing.assert_allclose(oth[6, py, px], w1 * w2 * (m(d1) - m(d2)) ** 2, rtol=1e-6)
        np.testing.assert_allclose(oth[0, py, px], w1 * d1 + w2 * d2, rtol=1e-9)
        np.testing.assert_allclose(oth[5, py, px], d1 if a1 >= 0.5 else d2, rtol=1e-9)
        np.testing.assert_allclose(oth[1, py, px], 1 - (1 - a1) * (1 - a2), rtol=1e-9)


def test_background_linearity_gives_transmittance():
    sc = surfel_cloud(5, P=2000, width=64, height=64)
    o0 = O.run_scene(sc)
    sc.bg = torch.zeros(3)
    o1 = O.run_scene(sc)
    diff = o0.get("out_color") - o1.get("out_color")
    alpha = o0.get("out_others")[1]
    np.testing.assert_allclose(diff, np.broadcast_to(1 - alpha, diff.shape), atol=2e-7)


def test_sh_basis_orthonormal():
    """Gauss-Legendre x uniform-phi quadrature of Y_i Y_j over the sphere."""
    nodes, weights = np.polynomial.legendre.leggauss(8)
    nphi = 16
    dirs, wts = [], []
    for ct, w in zip(nodes, weights):
        st = math.sqrt(1 - ct * ct)
        for k in range(nphi):
            ph = 2 * math.pi * k / nphi
            dirs.append([st * math.cos(ph), st * math.sin(ph), ct])
            wts.append(w * 2 * math.pi / nphi)
    dirs, wts = np.array(dirs), np.array(wts)
    n = len(dirs)
    Y = np.zeros((16, n))
    for b in range(16):
        sh = np.zeros((n, 16, 3))
        sh[:, b, :] = 0.1
        rgb, cl = O.sh_rgb(3, sh, dirs, np.zeros(3), f64=True)
        assert not cl.any()
        Y[b] = (rgb[:, 0] - 0.5) / 0.1
    G = (Y * wts) @ Y.T
    np.testing.assert_allclose(G, np.eye(16), atol=2e-7)


def test_sh_matches_reference_eval_sh_golden():
    g = np.load(os.path.join(GOLDEN, "eval_sh.npz"))
    dirs, coeffs = g["dirs"], g["coeffs"]  # coeffs (n, 3, 16): channel-major as eval_sh takes them
    sh = np.transpose(coeffs, (0, 2, 1)).copy()  # (n, 16, 3) as the rasteriser stores them
    for deg in range(4):
        rgb, cl = O.sh_rgb(deg, sh, dirs * 2.0, np.zeros(3), f64=True)
        ref = g[f"deg{deg}"] + 0.5
        np.testing.assert_allclose(rgb, np.maximum(ref, 0.0), rtol=1e-6, atol=1e-7)  # float constants
        np.testing.assert_array_equal(cl, ref < 0)


def test_mark_visible_near_plane():
    cam = axis_camera()
    pts = np.array([[0, 0, 0.1], [0, 0, 0.2], [0, 0, 0.2001], [0, 0, 5.0], [0, 0, -1.0]], np.float32)
    np.testing.assert_array_equal(O.mark_visible(pts, cam.world_view_transform), [False, False, True, True, False])


@pytest.mark.parametrize("f64", [False, True])
def test_oracle_deterministic(f64):
    sc = surfel_cloud(6, P=1500, width=48, height=40)
    a = O.run_scene(sc, f64=f64)
    b = O.run_scene(sc, f64=f64)
    for k in ("out_color", "out_others", "point_list", "n_contrib"):
        np.testing.assert_array_equal(a.get(k), b.get(k))
