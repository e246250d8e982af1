"""Block-stage losses (games/block_mesh_splatting/scene/block_gaussian_model.py:
510-648, utils/superquadric.py:25-46, scene/cameras.py:79-96).

CPU: the oracle's superquadric SDF pinned by closed forms (eps = 1 sphere,
the block layer's surface at eps1 != eps2, the clamped y axis, the flat
region beyond the +-5 clamp), the torch reductions against the oracle's on
random fields and against hand-made fields with known answers, and cast_rays
against the camera's own projection.  GPU: the fused SDF kernel against the
FP64 oracle (1e-5 relative), its adjoint against FP64 central differences of
the oracle, run to run, and summarize_losses term by term against the oracle
on the same random draws.
"""
import math

import numpy as np
import pytest
import torch

from oracle import block_oracle as BO
from partgs_amd import block_losses as BL
from partgs_amd.camera import random_sphere_camera


def _blocks(seed, B):
    rng = np.random.default_rng(seed)
    params, occ = BO.random_blocks(rng, B)
    return params, occ, rng


def _pinv_independent(params, pts, ratio=0.25, scale_min=0.2):
    S, R, T, _, _ = BO.block_frame(params, scale_min)
    # world = ((v ratio S) @ R) + T  =>  v = ((world - T) @ inv(R)) / (ratio S)
    return np.stack([((pts - T[b]) @ np.linalg.inv(R[b])) / (ratio * S[b]) for b in range(len(S))])


def test_sdf_sphere_at_eps_one():
    params, _, rng = _blocks(0, 4)
    params[:, 10:12] = math.log((1.0 - 0.1) / 1.8 / (1 - (1.0 - 0.1) / 1.8))  # eps = 1
    S, R, T, _, _ = BO.block_frame(params)
    for b in range(4):
        pts = (rng.uniform(-3, 3, (500, 3)) * 0.25 * S[b]) @ R[b] + T[b]
        pinv = _pinv_independent(params[b:b + 1], pts)[0]
        ok = np.all(pinv ** 2 >= 1e-6, axis=-1)
        sdf = BO.superquadric_sdf(params[b:b + 1], pts)[0]
        np.testing.assert_allclose(sdf[ok], np.linalg.norm(pinv, axis=-1)[ok] - 1.0, rtol=1e-12, atol=1e-12)
        assert ok.mean() > 0.9


def test_sdf_zero_on_block_surface_unequal_eps():
    params, _, rng = _blocks(1, 5)
    e = np.array([[0.3, 1.6], [1.7, 0.25], [0.9, 1.4], [1.2, 0.5], [0.6, 0.6]])
    params[:, 10:12] = np.log((e - 0.1) / 1.8 / (1 - (e - 0.1) / 1.8))
    S, R, T, e1, e2 = BO.block_frame(params)
    eta = rng.uniform(-np.pi / 2 + 0.05, np.pi / 2 - 0.05, 400)
    om = rng.uniform(-np.pi, np.pi, 400)
    om = np.where(np.min(np.abs(om[:, None] - np.array([-np.pi, -np.pi / 2, 0, np.pi / 2, np.pi])), 1) < 0.05,
                  om + 0.1, om)
    eta[np.abs(eta) < 0.05] += 0.1
    for b in range(5):
        u = BO.parametric_sq(eta, om, e1[b], e2[b])
        # keep samples away from the safe-power clamps: every squared
        # coordinate and x + z = |cos eta|^(2 e1 / e2) above 1e-5
        keep = np.all(u ** 2 > 1e-5, axis=1) & (np.abs(np.cos(eta)) ** (2 * e1[b] / e2[b]) > 1e-5)
        assert keep.sum() > 100
        w = (u[keep] * 0.25 * S[b]) @ R[b] + T[b]
        sdf = BO.superquadric_sdf(params[b:b + 1], w)[0]
        assert np.abs(sdf).max() < 1e-9, (b, np.abs(sdf).max())
        if abs(e1[b] - e2[b]) > 0.1:  # swapped exponents put it off the surface
            sw = BO.implicit_sq_sdf(_pinv_independent(params[b:b + 1], w)[0], e2[b], e1[b])
            assert np.abs(sw).max() > 1e-3


def test_sdf_clamped_axis_and_flat_region():
    params, _, _ = _blocks(2, 3)
    S, R, T, e1, e2 = BO.block_frame(params)
    for b in range(3):
        for y in (0.3, 0.8, 2.0):
            w = (np.array([0.0, y, 0.0]) * 0.25 * S[b]) @ R[b] + T[b]
            got = BO.superquadric_sdf(params[b:b + 1], w[None])[0, 0]
            want = (2 ** (e2[b] / e1[b]) * 1e-6 ** (1 / e1[b]) + y ** (2 / e1[b])) ** (e1[b] / 2) - 1
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (b, y, got, want)
        far = (np.array([[6.0, 7.0, -8.0], [9.0, 12.0, -6.0]]) * 0.25 * S[b]) @ R[b] + T[b]
        v = BO.superquadric_sdf(params[b:b + 1], far)[0]
        assert v[0] == v[1]


def test_reductions_against_oracle_and_hand_made_fields():
    rng = np.random.default_rng(3)
    B, Rn, Sm = 5, 64, 16
    sdf = rng.normal(0, 0.3, (B, Rn, Sm))
    mask = (rng.uniform(size=Rn) > 0.5).astype(np.float64)
    occ = rng.normal(size=(B, 1))
    t = lambda a: torch.tensor(a, dtype=torch.float64)  # noqa: E731
    ri, ro, bce = BO.ray_terms(sdf, mask, occ)
    assert abs(BL.ray_coverage_loss(t(sdf), t(mask))["ray_in_mask"].item() - ri) < 1e-12
    assert abs(BL.ray_uncoverage_loss(t(sdf), t(mask))["ray_out_mask"].item() - ro) < 1e-12
    assert abs(BL.occupancy_opacity_loss(t(sdf), t(mask), t(occ))["primitive_opacity"].item() - bce) < 1e-12
    # hand-made: every ray inside every block (sdf = -0.5): coverage 0, and
    # the uncovered rays each contribute 0.5 - 1e-3
    inside = np.full((B, Rn, Sm), -0.5)
    ri, ro, _ = BO.ray_terms(inside, mask, occ)
    assert ri == 0.0 and abs(ro - (0.5 - 1e-3)) < 1e-15
    # overlap: two blocks both deep inside at every sample, alpha = 1 each:
    # sum 2 -> excess 0.05
    assert abs(BO.overlap_value(np.full((2, 10), -1.0), [1.0, 1.0]) - 0.05) < 1e-12
    par = BL.parsimony_loss(t(np.array([[0.0], [2.0]])))["parsimony"].item()
    assert abs(par - 0.5 * (math.sqrt(0.5) + math.sqrt(1 / (1 + math.exp(-2))))) < 1e-12


def test_cast_rays_hit_their_pixels():
    cam = random_sphere_camera(3, 3.0, 64, 48)
    g = torch.Generator().manual_seed(0)
    pts, idx = BL.cast_rays(cam, 2.0, 5.0, num_samples=8, num_rays=100, device="cpu", generator=g)
    assert pts.shape == (100, 8, 3) and idx.unique().numel() == 100
    hom = torch.cat([pts.double(), torch.ones(100, 8, 1, dtype=torch.float64)], -1)
    view = hom @ cam.world_view_transform.double()
    fx = 64 / (2 * math.tan(cam.FoVx / 2))
    fy = 48 / (2 * math.tan(cam.FoVy / 2))
    u = view[..., 0] / view[..., 2] * fx + 32
    v = view[..., 1] / view[..., 2] * fy + 24
    assert torch.allclose(u, (idx % 64).double()[:, None] + 0.5, atol=1e-3)
    assert torch.allclose(v, torch.div(idx, 64, rounding_mode="floor").double()[:, None] + 0.5, atol=1e-3)
    # stratified depths: sample k lies in [2 + 3k/8, 2 + 3(k+1)/8) along the z = 1 ray
    z = view[..., 2]
    k = torch.arange(8, dtype=torch.float64)
    assert bool(((z >= 2 + 3 * k / 8 - 1e-4) & (z < 2 + 3 * (k + 1) / 8 + 1e-4)).all())


def _gpu_points(rng, N):
    return rng.uniform(-1.2, 1.2, (N, 3)).astype(np.float32)


@pytest.mark.gpu
def test_gpu_sdf_matches_oracle():
    params, _, rng = _blocks(4, 16)
    pts = _gpu_points(rng, 32768)
    print("[sdf] 16 blocks x 32768 points", flush=True)
    p32 = params.astype(np.float32)
    got = BL.superquadric_sdf(torch.from_numpy(p32).cuda(), torch.from_numpy(pts).cuda()).cpu().numpy()
    want = BO.superquadric_sdf(p32.astype(np.float64), pts.astype(np.float64))
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want) + 1.0)
    print(f"[sdf] max scaled err {err.max():.2e}", flush=True)
    assert err.max() <= 1e-5


@pytest.mark.gpu
def test_gpu_sdf_adjoint_against_fd64():
    params, _, rng = _blocks(5, 6)
    p32 = params.astype(np.float32)
    pts = _gpu_points(rng, 8192)
    g = rng.normal(size=(6, 8192)).astype(np.float32)
    print("[sdf] adjoint, 6 blocks x 8192 points", flush=True)
    P = torch.from_numpy(p32).cuda().requires_grad_(True)
    sdf = BL.superquadric_sdf(P, torch.from_numpy(pts).cuda())
    (sdf * torch.from_numpy(g).cuda()).sum().backward()
    got = P.grad.cpu().numpy()
    base = p32.astype(np.float64)
    fd = np.zeros_like(base)
    h = 1e-6
    for b in range(6):
        for k in range(12):
            pp, pm = base.copy(), base.copy()
            pp[b, k] += h
            pm[b, k] -= h
            fp = (BO.superquadric_sdf(pp[b:b + 1], pts)[0] * g[b]).sum()
            fm = (BO.superquadric_sdf(pm[b:b + 1], pts)[0] * g[b]).sum()
            fd[b, k] = (fp - fm) / (2 * h)
    scale = np.abs(fd).max(axis=1, keepdims=True)
    rel = np.abs(got - fd) / (np.abs(fd) + 1e-3 * scale)
    print(f"[sdf] adjoint max rel err {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-3
    P2 = torch.from_numpy(p32).cuda().requires_grad_(True)
    (BL.superquadric_sdf(P2, torch.from_numpy(pts).cuda()) * torch.from_numpy(g).cuda()).sum().backward()
    assert torch.equal(P.grad, P2.grad)


@pytest.mark.gpu
def test_gpu_summarize_losses_terms():
    params, occ, rng = _blocks(6, 8)
    dev = torch.device("cuda")
    P = torch.from_numpy(params.astype(np.float32)).to(dev)
    O = torch.from_numpy(occ.astype(np.float32)).to(dev).reshape(-1, 1)
    cam = random_sphere_camera(1, 3.0, 80, 60)
    mask = torch.from_numpy((rng.uniform(size=(60, 80)) > 0.6).astype(np.float32)).to(dev)
    img = torch.rand(3, 60, 80, device=dev)
    gt = torch.rand(3, 60, 80, device=dev)
    opt = BL.BlockLossWeights()
    print("[losses] summarize_losses, 8 blocks", flush=True)
    gen = torch.Generator(device=dev).manual_seed(11)
    Ll1, loss, d = BL.summarize_losses(img, gt, cam, opt, P, O, mask=mask, generator=gen, near_far=(1.0, 5.0))
    # the same draws again, reduced by the oracle
    gen = torch.Generator(device=dev).manual_seed(11)
    pts, ray_idx = BL.cast_rays(cam, 1.0, 5.0, num_samples=128, device=dev, generator=gen)
    S, R, T, _, _ = BL.blocks_srt(P)
    ov = torch.rand(8, BL.OVERLAP_N_POINTS, 3, device=dev, generator=gen) * 2 - 1
    ov = ((ov * 0.25 * S[:, None]) @ R + T[:, None]).reshape(-1, 3)
    noise = torch.randn(O.shape, device=dev, generator=gen)
    lm = mask.flatten()[ray_idx].cpu().numpy()
    sdf = BO.superquadric_sdf(params.astype(np.float32).astype(np.float64),
                              pts.reshape(-1, 3).cpu().numpy().astype(np.float64)).reshape(8, len(lm), 128)
    ri, ro, bce = BO.ray_terms(sdf, lm, occ.astype(np.float32))
    alpha = 1 / (1 + np.exp(-(occ.astype(np.float32).astype(np.float64) + noise.cpu().numpy().reshape(-1))))
    ovv = BO.overlap_value(BO.superquadric_sdf(params.astype(np.float32).astype(np.float64),
                                               ov.cpu().numpy().astype(np.float64)), alpha)
    print({k: float(v) for k, v in d.items()}, ri, ro, bce, ovv, flush=True)
    assert abs(d["ray_in_mask"].item() - ri) <= 1e-4 * max(1.0, ri)
    assert abs(d["ray_out_mask"].item() - ro) <= 1e-4 * max(1.0, ro)
    # a sample within FP32 rounding of the surface may fall on either side of
    # the sdf <= 1e-8 test, moving one ray's max opacity: about 1e-3 per ray
    assert abs(d["primitive_opacity"].item() - bce) <= 3e-3
    assert abs(d["overlap"].item() - ovv) <= 1e-3 * max(1.0, ovv)
    loss.backward() if loss.requires_grad else None
    Pg = P.clone().requires_grad_(True)
    gen = torch.Generator(device=dev).manual_seed(11)
    _, loss2, _ = BL.summarize_losses(img, gt, cam, opt, Pg, O, mask=mask, generator=gen, near_far=(1.0, 5.0))
    loss2.backward()
    assert torch.isfinite(Pg.grad).all() and Pg.grad.abs().sum() > 0
