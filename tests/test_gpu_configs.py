"""GPU parity at the benchmark configurations' own sizes.

* 300,000 superquadric surfels (SH degree 3, all aux channels) at 400x300
  and 800x600 -- the benchmarked workload itself -- through the exact-size
  path and through the drop-in autograd API the benchmark times (binning
  capacity and pair store sized ahead, no synchronisation mid-forward), and
  with a pair store small enough that most quadrants recompute their pairs.
* 1,000,000 surfels over the benchmark's 32 ragged views, batched, at tiles
  8, 16 and 32: one view of each resolution against the oracle run on that
  view alone.
* 16 superquadric blocks x 64 x 64 surfels rasterised at 400x300: the block
  parameter gradient through the rendered image against FP64 central
  differences (h = 1e-4) of the oracle's frozen-decision replay.
* The proposal network's fused point MLP + max-pool at 20,000 points.

Bars as in test_gpu_parity: integer state and images bit-exact, gradients
within 1e-5 (element-relative with a 1e-5 max|g| floor) of the oracle's
FP64-accumulated sums.
"""
import math
import os
from dataclasses import replace

import numpy as np
import pytest
import torch

from oracle import oracle as O
from partgs_amd.scenes import batched_views, superquadric_surfels, upstream_grads
from tests.helpers import assert_close_scaled, gpu_backward, gpu_forward, oracle_run, rel_err
from tests.test_gpu_parity import GRAD_TOL, _check_backward, _check_forward

pytestmark = pytest.mark.gpu

GRAD_NAMES = ("dL_dopacity", "dL_dcolors", "dL_dmeans3D", "dL_dscales", "dL_drotations", "dL_dsh", "dL_dmeans2D")


@pytest.fixture(scope="module", autouse=True)
def _threads():
    O.set_threads(min(16, os.cpu_count() or 1))


@pytest.mark.parametrize("W,H", [(400, 300), (800, 600)])
def test_superquadric_300k(W, H):
    from partgs_amd import rasterizer as R
    from partgs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer

    scene = superquadric_surfels(0, P=300_000, width=W, height=H, sh_degree=3)
    o, rc = oracle_run(scene)
    assert rc == 0
    print(f"[configs] 300k {W}x{H}: oracle L={o.num_rendered}", flush=True)
    res = gpu_forward(scene)
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=0)
    gc, go = upstream_grads(0, H, W)
    # the pair store at 1/16 of a pair per instance: most quadrants take the
    # recompute path, with the same results
    print(f"[configs] 300k {W}x{H}: small pair store", flush=True)
    small = gpu_forward(scene, capacity=res["L"], pair_capacity=max(res["L"] // 16, 1))
    for k in ("color", "others", "radii"):
        assert torch.equal(small[k], res[k]), k
    g_small = gpu_backward(small, gc, go)
    for name in GRAD_NAMES:
        assert_close_scaled(g_small[name], o.get(name), GRAD_TOL, name)
    # the benchmarked path: GaussianRasterizer forward + backward, three
    # steps (exact sizing, then capacity and pair store sized ahead)
    cam = scene.camera
    s = scene.to("cuda")
    rs = GaussianRasterizationSettings(image_height=H, image_width=W, tanfovx=math.tan(cam.FoVx * 0.5),
                                       tanfovy=math.tan(cam.FoVy * 0.5), bg=s.bg, scale_modifier=1.0,
                                       viewmatrix=s.camera.world_view_transform,
                                       projmatrix=s.camera.full_proj_transform, sh_degree=3,
                                       campos=s.camera.camera_center, prefiltered=False, debug=False)
    leaves = [t.clone().requires_grad_(True) for t in (s.means3D, s.shs, s.opacities, s.scales, s.rotations)]
    means2D = torch.zeros_like(s.means3D, requires_grad=True)
    ref_color, ref_others = o.get("out_color"), o.get("out_others")
    for step in range(3):
        print(f"[configs] 300k {W}x{H}: autograd step {step}", flush=True)
        for t in leaves + [means2D]:
            t.grad = None
        color, radii, allmap = GaussianRasterizer(rs)(means3D=leaves[0], means2D=means2D, shs=leaves[1],
                                                      opacities=leaves[2], scales=leaves[3], rotations=leaves[4])
        torch.autograd.backward((color, allmap), (gc.cuda(), go.cuda()))
        np.testing.assert_array_equal(color.detach().cpu().numpy(), ref_color)
        np.testing.assert_array_equal(allmap.detach().cpu().numpy(), ref_others)
        np.testing.assert_array_equal(radii.cpu().numpy(), o.get("radii"))
        for t, name in zip(leaves + [means2D], ("dL_dmeans3D", "dL_dsh", "dL_dopacity", "dL_dscales",
                                                "dL_drotations", "dL_dmeans2D")):
            assert_close_scaled(t.grad.cpu().numpy(), o.get(name), GRAD_TOL, name)
    key = (torch.device("cuda", torch.cuda.current_device()), 16, W, H)
    assert R._capacity_hint.get(key, 0) >= res["L"]


@pytest.fixture(scope="module")
def million():
    return superquadric_surfels(0, P=1_000_000, width=800, height=600, sh_degree=3)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_batched_million_surfels_per_view_against_oracle(million, tile):
    from partgs_amd.batch import BatchConfig, BatchRasterizer

    cams = batched_views(0, 32)
    picks = []
    for res in ((400, 300), (640, 480), (800, 600)):
        picks.append(next(v for v, c in enumerate(cams) if (c.image_width, c.image_height) == res))
    dev = torch.device("cuda")
    s = million.to(dev)
    br = BatchRasterizer([c.to(dev) for c in cams], s.bg, 3, BatchConfig(tile=tile))
    print(f"[configs] batch 1M x 32 views tile {tile}: forward", flush=True)
    color, others, radii = br.forward_checked(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations)
    torch.cuda.synchronize()
    need, flags = br.read_status()
    assert flags == 0
    print(f"[configs] tile {tile}: L={need} pairs={br.read_pairs()} workspace {br.workspace.numel() / 1e9:.1f} GB",
          flush=True)
    # upstream gradients on the picked views only: the summed per-surfel
    # gradients are then the sum of those views' gradients
    gcs, gos = [torch.zeros(3 * n) for n in br.pixels], [torch.zeros(7 * n) for n in br.pixels]
    for v in picks:
        gc, go = upstream_grads(v, br.heights[v], br.widths[v])
        gcs[v], gos[v] = gc.reshape(-1), go.reshape(-1)
    gc_flat = torch.cat(gcs).to(dev)
    go_flat = torch.cat(gos).to(dev)
    grads = br.new_grads(s.P, s.shs.shape[1])
    print(f"[configs] tile {tile}: backward", flush=True)
    br.backward_raw(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations, radii, gc_flat, go_flat, grads)
    torch.cuda.synchronize()
    col_v, oth_v = br.split(color, 3), br.split(others, 7)
    sums = {}
    for v in picks:
        cam = cams[v]
        print(f"[configs] tile {tile}: oracle view {v} {cam.image_width}x{cam.image_height}", flush=True)
        o, rc = oracle_run(replace(million, camera=cam), tile=tile)
        assert rc == 0
        np.testing.assert_array_equal(radii[v].cpu().numpy(), o.get("radii"))
        np.testing.assert_array_equal(col_v[v].cpu().numpy(), o.get("out_color"))
        np.testing.assert_array_equal(oth_v[v].cpu().numpy(), o.get("out_others"))
        o.backward(gcs[v].reshape(3, cam.image_height, cam.image_width).numpy(),
                   gos[v].reshape(7, cam.image_height, cam.image_width).numpy())
        assert_close_scaled(grads["means2D"][v].cpu().numpy(), o.get("dL_dmeans2D"), GRAD_TOL, f"means2D[{v}]")
        for name in ("dL_dopacity", "dL_dcolors", "dL_dmeans3D", "dL_dscales", "dL_drotations", "dL_dsh"):
            sums[name] = sums.get(name, 0.0) + o.get(name).astype(np.float64)
    for key, name in (("opacities", "dL_dopacity"), ("means3D", "dL_dmeans3D"), ("scales", "dL_dscales"),
                      ("rotations", "dL_drotations"), ("shs", "dL_dsh")):
        got = grads[key].cpu().numpy().reshape(sums[name].shape)
        print(f"[configs] tile {tile}: {name} max rel {rel_err(got, sums[name]):.2e}", flush=True)
        assert_close_scaled(got, sums[name], GRAD_TOL, name)
    del br, grads, color, others, radii
    torch.cuda.empty_cache()


def test_blocks_through_rasteriser_fd64():
    """16 blocks x 64 x 64 surfels rendered at 400x300.  The GPU gradient of
    L = <gc, color> + <go, allmap> with respect to 64 random block parameters
    (block kernel -> rasteriser -> their adjoints, FP32) against FP64 central
    differences with h = 1e-4 of the oracle chain: the numpy block layer in
    FP64, then the rasteriser oracle's FP64 replay of its recorded forward
    (every discrete decision held, so the image is smooth in the parameters
    and its derivative is what the backward computes).  The block layer's own
    kinks remain (see below); at most 8 of the 64 entries may be excluded for
    one."""
    from oracle import block_oracle as BO
    from partgs_amd.blocks import generate_surfels, make_grid
    from partgs_amd.camera import random_sphere_camera
    from partgs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer

    rng = np.random.default_rng(31)
    params, occ = BO.random_blocks(rng, 16, scale_range=(0.1, 0.5), eps_range=(0.2, 1.8))
    params = params.astype(np.float32).astype(np.float64)
    occ = np.clip(occ, -3.0, 3.0).astype(np.float32).astype(np.float64)  # opacity below the 0.99 alpha cap
    grid = make_grid(64, 64)
    W, H = 400, 300
    cam = random_sphere_camera(31, 3.0, W, H)
    P = 16 * 64 * 64
    colors = rng.uniform(0.0, 1.0, size=(P, 3))
    gc = rng.normal(size=(3, H, W))
    go = rng.normal(size=(7, H, W))
    bg = np.ones(3)
    # GPU: block kernel -> rasteriser, autograd
    dev = torch.device("cuda")
    p = torch.tensor(params, dtype=torch.float32, device=dev, requires_grad=True)
    o_ = torch.tensor(occ, dtype=torch.float32, device=dev, requires_grad=True)
    print("[configs] blocks 16 x 64x64 -> 400x300: GPU forward/backward", flush=True)
    means, scales, rots, opac = generate_surfels(p, o_, grid)
    camd = cam.to(dev)
    rs = GaussianRasterizationSettings(image_height=H, image_width=W, tanfovx=math.tan(cam.FoVx * 0.5),
                                       tanfovy=math.tan(cam.FoVy * 0.5), bg=torch.ones(3, device=dev),
                                       scale_modifier=1.0, viewmatrix=camd.world_view_transform,
                                       projmatrix=camd.full_proj_transform, sh_degree=0, campos=camd.camera_center,
                                       prefiltered=False, debug=False)
    means2D = torch.zeros_like(means, requires_grad=True)
    color, radii, allmap = GaussianRasterizer(rs)(means3D=means, means2D=means2D, opacities=opac,
                                                  colors_precomp=torch.tensor(colors, dtype=torch.float32,
                                                                              device=dev),
                                                  scales=scales, rotations=rots)
    assert int((radii > 0).sum()) > P // 4
    torch.autograd.backward([color, allmap], [torch.tensor(gc, dtype=torch.float32, device=dev),
                                              torch.tensor(go, dtype=torch.float32, device=dev)])
    g_params = p.grad.cpu().numpy().astype(np.float64)
    g_occ = o_.grad.cpu().numpy().astype(np.float64)
    # oracle chain in FP64
    eta = grid.eta.cpu().numpy().astype(np.float64)
    om = grid.omega.cpu().numpy().astype(np.float64)

    def kw(pp, oo):
        m, sc, q, a = BO.generate(pp, oo, eta, om)
        return dict(means3D=m, opacities=a, viewmatrix=cam.world_view_transform, projmatrix=cam.full_proj_transform,
                    campos=cam.camera_center, bg=bg, W=W, H=H, tanfovx=math.tan(cam.FoVx * 0.5),
                    tanfovy=math.tan(cam.FoVy * 0.5), colors_precomp=colors, scales=sc, rotations=q, f64=True)

    def loss(orc):
        return float((orc.get("out_color") * gc).sum() + (orc.get("out_others") * go).sum())

    base = O.Oracle()
    base.forward(**kw(params, occ), record=True)
    assert base.alpha_capped == 0
    print(f"[configs] blocks: oracle L={base.num_rendered}", flush=True)

    def fd(b, j, h):
        """Central difference of the replayed loss in block parameter (b, j)."""
        vals = []
        for sgn in (1.0, -1.0):
            pp, oo = params.copy(), occ.copy()
            if j < 12:
                pp[b, j] += sgn * h
            else:
                oo[b] += sgn * h
            base.replay(**kw(pp, oo))
            vals.append(loss(base))
        return (vals[0] - vals[1]) / (2 * h)

    # central differences at h = 1e-4, h / 2 and h / 4; two Richardson
    # estimates (4 D(h/2) - D(h)) / 3 and (4 D(h/4) - D(h/2)) / 3 cancel the
    # h^2 term.  Nearly edge-on surfels make the AABB-centre map steep (its
    # denominator 9 (Tw.x^2 + Tw.y^2) - Tw.z^2 nears 0), so a translation's
    # h^2 term can be large; where the two estimates still disagree the
    # function has a kink within h (the block layer's 1e-8 guards make
    # rot_to_quat switch branches on slightly non-orthonormal frames) and
    # the entry is excluded.
    h = 1e-4
    picks = np.random.default_rng(7).choice(16 * 13, size=64, replace=False)
    est, spread, an = [], [], []
    for k in picks:
        b, j = divmod(int(k), 13)
        d1, d2, d4 = fd(b, j, h), fd(b, j, h / 2), fd(b, j, h / 4)
        r1, r2 = (4 * d2 - d1) / 3, (4 * d4 - d2) / 3
        est.append(r1)
        spread.append(abs(r1 - r2))
        an.append(g_params[b, j] if j < 12 else g_occ[b])
    est, spread, an = np.array(est), np.array(spread), np.array(an)
    scale = np.abs(est).max()
    smooth = spread <= 1e-3 * (np.abs(est) + 1e-2 * scale)
    err = np.abs(est - an)[smooth]
    fds = est[smooth]
    print(f"[configs] blocks: {smooth.sum()} smooth entries, {(~smooth).sum()} with a kink within h; "
          f"|fd| max {scale:.3e}, max abs err {err.max():.3e} (rel to max {err.max() / scale:.2e}), "
          f"max elementwise rel {np.max(err / (np.abs(fds) + 1e-3 * scale)):.2e}", flush=True)
    assert (~smooth).sum() <= 8
    assert (err <= 1e-4 * scale + 1e-3 * np.abs(fds)).all(), (err.max(), scale)


def test_proposal_pool_at_20000_points():
    from oracle import proposal_oracle as PO
    from partgs_amd import proposal as SP

    print("[configs] proposal pool 3 x 20000 points", flush=True)
    net = SP.ProposalNet(seed=5)
    prm = {k: v.cpu().numpy() for k, v in net.p.items()}
    rng = np.random.default_rng(6)
    pts = rng.normal(scale=0.4, size=(3, 20000, 3)).astype(np.float32)
    val, idx = net.pool(torch.tensor(pts, device="cuda"))
    o = PO.Oracle(prm, bf16=True)
    for b in range(3):
        _, _, h3 = o.shared(pts[b])
        want = h3.max(0)
        got = val[b].cpu().numpy()
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 2e-2 * scale, np.abs(got - want).max() / scale
        srt = np.sort(h3, axis=0)
        clear = (srt[-1] - srt[-2]) > 5e-2 * scale
        assert clear.sum() > 0
        assert np.array_equal(idx[b].cpu().numpy()[clear], h3.argmax(0)[clear])
