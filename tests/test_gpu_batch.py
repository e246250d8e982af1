"""Batched views against the single-view path, view by view.

The forward of a batch must reproduce, bit for bit, what the single-view
entry point gives for each view (which the parity tests pin to the oracle);
per-surfel gradients must equal the sum over views of the single-view
gradients (FP32 atomics reorder the sums: 1e-5 scaled tolerance).  A hipGraph
replay must equal the eager passes on every replay."""
import math

import numpy as np
import pytest
import torch

from helpers import assert_close_scaled
from partgs_amd import _lib
from partgs_amd.batch import BatchConfig, BatchRasterizer, ViewGroups, autotune, capture_grouped_step, capture_step
from partgs_amd.camera import random_sphere_camera
from partgs_amd.rasterizer import rasterize_gaussians_bwd, rasterize_gaussians_fwd
from partgs_amd.scenes import superquadric_surfels, surfel_cloud

pytestmark = pytest.mark.gpu

RES = [(160, 120), (96, 72), (200, 150)]


def _setup(seed=0, P=20000, res=RES, sh_degree=3):
    scene = superquadric_surfels(seed, P=P, n_sq=4, width=res[0][0], height=res[0][1], sh_degree=sh_degree)
    cams = [random_sphere_camera(seed * 10 + k, 3.0, w, h) for k, (w, h) in enumerate(res)]
    dev = torch.device("cuda")
    s = scene.to(dev)
    return s, [c.to(dev) for c in cams]


def _single(s, cam, tile, gc=None, go=None):
    e = torch.empty(0, device="cuda")
    tx, ty = math.tan(cam.FoVx * 0.5), math.tan(cam.FoVy * 0.5)
    H, W = cam.image_height, cam.image_width
    L, color, others, radii, geom, binb, img = rasterize_gaussians_fwd(
        s.bg, s.means3D, e, s.opacities, s.scales, s.rotations, 1.0, e, cam.world_view_transform,
        cam.full_proj_transform, tx, ty, H, W, s.shs, s.sh_degree, cam.camera_center, False, False, tile)
    out = {"color": color, "others": others, "radii": radii}
    if gc is not None:
        g = rasterize_gaussians_bwd(s.bg, s.means3D, radii, e, s.scales, s.rotations, 1.0, e,
                                    cam.world_view_transform, cam.full_proj_transform, tx, ty, gc, go, s.shs,
                                    s.sh_degree, cam.camera_center, geom, L, binb, img, False, tile)
        out.update(means2D=g[0], colors=g[1], opacities=g[2], means3D=g[3], shs=g[5], scales=g[6], rotations=g[7])
    return out


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_batched_forward_equals_single_views(tile):
    s, cams = _setup()
    print(f"[batch] fwd tile {tile} views {RES}", flush=True)
    br = BatchRasterizer(cams, s.bg, s.sh_degree, BatchConfig(tile=tile))
    color, others, radii = br.forward_checked(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations)
    torch.cuda.synchronize()
    need, flags = br.read_status()
    assert flags == 0
    total = 0
    for v, cam in enumerate(cams):
        ref = _single(s, cam, tile)
        assert torch.equal(br.split(color, 3)[v], ref["color"]), v
        assert torch.equal(br.split(others, 7)[v], ref["others"]), v
        assert torch.equal(radii[v], ref["radii"]), v
        total += int(ref["radii"].gt(0).sum())
    assert total > 0 and need > 0


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_batched_dense_tiles_equal_single_views(tile):
    """Tiles above 2048 instances: a batch sorts each with one workgroup, a
    single view in 2048-key chunks merged by rank; both must give the same
    images, bit for bit."""
    scene = surfel_cloud(5, P=12000, width=32, height=32)
    cam = scene.camera
    c2w = torch.linalg.inv(cam.world_view_transform.T)
    centre = cam.camera_center + 3.0 * c2w[:3, 2]
    g = torch.Generator().manual_seed(5)
    scene.means3D = (centre + 0.05 * (torch.rand(scene.P, 3, generator=g) - 0.5)).float()
    s = scene.to("cuda")
    cams = [cam.to("cuda"), cam.to("cuda")]
    print(f"[batch] dense tiles, tile {tile}", flush=True)
    br = BatchRasterizer(cams, s.bg, s.sh_degree, BatchConfig(tile=tile))
    color, others, radii = br.forward_checked(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations)
    torch.cuda.synchronize()
    for v, c in enumerate(cams):
        ref = _single(s, c, tile)
        assert torch.equal(br.split(color, 3)[v], ref["color"]), v
        assert torch.equal(br.split(others, 7)[v], ref["others"]), v
        assert torch.equal(radii[v], ref["radii"]), v


def test_batched_backward_sums_single_views():
    s, cams = _setup(1)
    print("[batch] bwd", flush=True)
    br = BatchRasterizer(cams, s.bg, s.sh_degree)
    color, others, radii = br.forward_checked(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations)
    g = torch.Generator(device="cpu").manual_seed(3)
    gcs = [torch.randn(3, h, w, generator=g).cuda() for w, h in RES]
    gos = [torch.randn(7, h, w, generator=g).cuda() for w, h in RES]
    grads = br.new_grads(s.P, s.shs.shape[1])
    br.backward_raw(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations, radii, br.pack(gcs, 3),
                    br.pack(gos, 7), grads)
    torch.cuda.synchronize()
    acc = {}
    for v, cam in enumerate(cams):
        ref = _single(s, cam, 16, gcs[v], gos[v])
        assert_close_scaled(grads["means2D"][v].cpu().numpy(), ref["means2D"].cpu().numpy(), 1e-5,
                            f"means2D view {v}")
        for k in ("colors", "opacities", "means3D", "shs", "scales", "rotations"):
            acc[k] = acc.get(k, 0) + ref[k].double()
    for k, want in acc.items():
        assert_close_scaled(grads[k].cpu().numpy(), want.cpu().numpy(), 1e-5, k)


def test_capacity_overflow_is_reported_and_recovered():
    s, cams = _setup(2)
    print("[batch] capacity", flush=True)
    br = BatchRasterizer(cams, s.bg, s.sh_degree, capacity=1000)
    br.ensure_workspace(s.P)
    out = br.new_outputs(s.P)
    br.forward_raw(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations, *out)
    need, flags = br.read_status()
    assert flags & _lib.SR_STATUS_CAPACITY and need > 1000
    color, others, radii = br.forward_checked(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations)
    assert br.capacity >= need
    ref = _single(s, cams[0], 16)
    assert torch.equal(br.split(color, 3)[0], ref["color"])


def test_slot_overflow_is_reported_and_recovered():
    """A batch whose partial-gradient slots outgrow the slot capacity sets
    the capacity flag with the instances within the capacity and reports the
    slots needed; forward_checked then re-sizes the slots and renders the
    same images as a batch sized from the start."""
    s, cams = _setup(2)
    print("[batch] slot overflow", flush=True)
    br = BatchRasterizer(cams, s.bg, s.sh_degree)
    ref_color, _, _ = br.forward_checked(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations)
    need, _ = br.read_status()
    slots = br.read_slots()
    assert need <= slots <= 4 * need
    br.ensure_workspace(s.P, br.capacity, br.pair_capacity, slots // 2)
    out = br.new_outputs(s.P)
    br.forward_raw(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations, *out)
    n2, flags = br.read_status()
    assert flags & _lib.SR_STATUS_CAPACITY and n2 == need and br.read_slots() == slots
    color, _, _ = br.forward_checked(s.means3D, s.shs, None, s.opacities, s.scales, s.rotations)
    assert br.slot_capacity >= slots
    assert torch.equal(color, ref_color)


def test_graph_replay_equals_eager():
    """Forward + backward captured once, replayed three times with eager work
    and input changes between replays; every replay is checked against the
    eager passes on the same inputs."""
    s, cams = _setup(3)
    print("[batch] graph", flush=True)
    br = BatchRasterizer(cams, s.bg, s.sh_degree)
    means = s.means3D.clone()
    gen = torch.Generator(device="cpu").manual_seed(5)
    gc = torch.randn(3 * br.N, generator=gen).cuda()
    go = torch.randn(7 * br.N, generator=gen).cuda()
    step = capture_step(br, means, s.shs, None, s.opacities, s.scales, s.rotations, gc, go)
    eager = BatchRasterizer(cams, s.bg, s.sh_degree, capacity=br.capacity)
    for k in range(3):
        means.copy_(s.means3D + 0.002 * k)
        step.replay()
        torch.cuda.synchronize()
        assert step.ok()
        eager.ensure_workspace(s.P)
        out = eager.new_outputs(s.P)
        eager.forward_raw(means, s.shs, None, s.opacities, s.scales, s.rotations, *out)
        grads = eager.new_grads(s.P, s.shs.shape[1])
        eager.backward_raw(means, s.shs, None, s.opacities, s.scales, s.rotations, out[2], gc, go, grads)
        torch.cuda.synchronize()
        assert torch.equal(step.outputs[0], out[0]), k
        assert torch.equal(step.outputs[1], out[1]), k
        assert torch.equal(step.outputs[2], out[2]), k
        for name in ("means3D", "scales", "rotations", "opacities", "shs", "means2D"):
            assert_close_scaled(step.grads[name].cpu().numpy(), grads[name].cpu().numpy(), 1e-5, f"{name} {k}")
        _ = torch.randn(1000, device="cuda").sum().item()  # unrelated eager work between replays


def test_view_groups_equal_one_batch():
    """Views in two groups on one shared workspace: images, radii and the
    per-view screen-space statistic equal the one-batch replay bit for bit;
    per-surfel gradients are the same sums regrouped.  Replayed twice with
    an input change in between."""
    s, cams = _setup(6)
    print("[batch] view groups", flush=True)
    means = s.means3D.clone()
    br = BatchRasterizer(cams, s.bg, s.sh_degree)
    gen = torch.Generator(device="cpu").manual_seed(7)
    gc = torch.randn(3 * br.N, generator=gen).cuda()
    go = torch.randn(7 * br.N, generator=gen).cuda()
    one = capture_step(br, means, s.shs, None, s.opacities, s.scales, s.rotations, gc, go)
    vg = ViewGroups(cams, s.bg, s.sh_degree, groups=2)
    assert [p.V for p in vg.parts] == [1, 2] and vg.N == br.N
    two = capture_grouped_step(vg, means, s.shs, None, s.opacities, s.scales, s.rotations, gc, go)
    assert vg.workspace.numel() < br.workspace.numel()
    for k in range(2):
        means.copy_(s.means3D + 0.003 * k)
        one.replay()
        two.replay()
        torch.cuda.synchronize()
        assert one.ok() and two.ok()
        for i in range(3):
            assert torch.equal(one.outputs[i], two.outputs[i]), (k, i)
        assert torch.equal(one.grads["means2D"], two.grads["means2D"]), k
        for name in ("means3D", "scales", "rotations", "opacities", "shs", "colors"):
            assert_close_scaled(two.grads[name].cpu().numpy(), one.grads[name].cpu().numpy(), 1e-5, f"{name} {k}")


def test_autograd_through_batch():
    s, cams = _setup(4, P=8000)
    print("[batch] autograd", flush=True)
    br = BatchRasterizer(cams, s.bg, s.sh_degree)
    leaves = [t.clone().requires_grad_(True) for t in (s.means3D, s.shs, s.opacities, s.scales, s.rotations)]
    colors, allmaps, radii = br(means3D=leaves[0], shs=leaves[1], opacities=leaves[2], scales=leaves[3],
                                rotations=leaves[4])
    loss = sum(c.sum() for c in colors) + sum(a[6].sum() for a in allmaps)
    loss.backward()
    for t in leaves:
        assert t.grad is not None and torch.isfinite(t.grad).all()
    assert radii.shape == (len(cams), s.P)


def test_autotune_picks_a_candidate(tmp_path):
    s, cams = _setup(5, P=8000, res=[(96, 72), (64, 48)])
    print("[batch] autotune", flush=True)
    path = str(tmp_path / "tune.json")
    cfg = autotune(cams, s.bg, s.sh_degree, s.means3D, s.shs, None, s.opacities, s.scales, s.rotations,
                   tiles=(8, 16), pair_stores=(0, 8), reps=1, cache_path=path, log=print)
    assert cfg.tile in (8, 16) and cfg.pair_store in (0, 8)
    cfg2 = autotune(cams, s.bg, s.sh_degree, s.means3D, s.shs, None, s.opacities, s.scales, s.rotations,
                    cache_path=path)
    assert cfg2 == cfg
