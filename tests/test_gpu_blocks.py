"""GPU parity of the block layer: the HIP generation kernel against the FP64
numpy oracle, and its adjoint against FP64 central differences (h = 1e-4)."""
import math

import numpy as np
import pytest
import torch

from oracle import block_oracle as BO
from oracle import oracle as O
from partgs_amd.blocks import generate_surfels, make_grid
from partgs_amd.camera import random_sphere_camera
from partgs_amd.rasterizer import rasterize_gaussians_bwd, rasterize_gaussians_fwd

pytestmark = pytest.mark.gpu


def _blocks(seed, B):
    rng = np.random.default_rng(seed)
    params, occ = BO.random_blocks(rng, B, scale_range=(0.1, 0.5), eps_range=(0.2, 1.8))
    return params, occ


def _run_gpu(params, occ, n_eta, n_omega):
    grid = make_grid(n_eta, n_omega)
    p = torch.tensor(params, dtype=torch.float32, device="cuda", requires_grad=True)
    o = torch.tensor(occ, dtype=torch.float32, device="cuda", requires_grad=True)
    outs = generate_surfels(p, o, grid)
    return grid, p, o, outs


def _rotmat(q):
    """Rotation matrices of (w, x, y, z) quaternions (numpy or torch), flattened."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    rows = [w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y),
            2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x),
            2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z]
    return torch.stack(rows, 1) if isinstance(q, torch.Tensor) else np.stack(rows, 1)


def test_block_forward_matches_oracle():
    params, occ = _blocks(0, 16)
    grid, p, o, outs = _run_gpu(params, occ, 64, 64)
    print(f"[blocks] B=16 grid 64x64 P={outs[0].shape[0]}", flush=True)
    ref = BO.generate(p.detach().cpu().numpy().astype(np.float64), o.detach().cpu().numpy().astype(np.float64),
                      grid.eta.cpu().numpy().astype(np.float64), grid.omega.cpu().numpy().astype(np.float64))
    outs = [outs[0], outs[1], _rotmat(outs[2]), outs[3]]
    ref = [ref[0], ref[1], _rotmat(ref[2]), ref[3]]
    for got, want, name in zip(outs, ref, ("means3D", "scales", "rotation matrices", "opacities")):
        got = got.detach().cpu().numpy().astype(np.float64)
        err = np.abs(got - want)
        tol = 2e-5 * np.abs(want).max() + 1e-4 * np.abs(want)
        assert (err <= tol).all(), (name, err.max(), np.argmax(err))


def test_block_backward_matches_fd64():
    """64 random block parameters, FP64 central differences with h = 1e-4.
    The loss reads the rotation through R(q): the reference's w >= 0
    standardisation flips q -> -q (the same rotation) discontinuously.
    The map has kinks (the relu floor on a face's second scale, frame flips
    of nearly collinear edges); an entry whose FD at h and at h / 2 disagree
    has one within h and is excluded, and at most 8 of the 64 may be."""
    params, occ = _blocks(1, 16)
    params = params.astype(np.float32).astype(np.float64)
    occ = occ.astype(np.float32).astype(np.float64)
    grid, p, o, outs = _run_gpu(params, occ, 64, 64)
    rng = np.random.default_rng(7)
    feats = [outs[0], outs[1], _rotmat(outs[2]), outs[3]]
    gs = [rng.normal(size=t.shape) for t in feats]
    loss = sum((t * torch.tensor(g, dtype=torch.float32, device="cuda")).sum() for t, g in zip(feats, gs))
    loss.backward()
    gp = p.grad.cpu().numpy().astype(np.float64)
    go = o.grad.cpu().numpy().astype(np.float64)
    eta = grid.eta.cpu().numpy().astype(np.float64)
    om = grid.omega.cpu().numpy().astype(np.float64)

    def f(pp, oo):
        m, sc, r, a = BO.generate(pp, oo, eta, om)
        return sum(float((x * g).sum()) for x, g in zip((m, sc, _rotmat(r), a), gs))

    def fd(b, j, h):
        pp, oo = params.copy(), occ.copy()
        tgt = (pp, (b, j)) if j < 12 else (oo, (b,))
        tgt[0][tgt[1]] += h
        lp = f(pp, oo)
        tgt[0][tgt[1]] -= 2 * h
        lm = f(pp, oo)
        return (lp - lm) / (2 * h)

    h = 1e-4
    picks = rng.choice(16 * 13, size=64, replace=False)
    fds, an, kinks = [], [], 0
    for k in picks:
        b, j = divmod(int(k), 13)
        d1, d2 = fd(b, j, h), fd(b, j, h / 2)
        if abs(d1 - d2) > 1e-4 * (abs(d1) + 1.0):
            kinks += 1
            continue
        fds.append(d1)
        an.append(gp[b, j] if j < 12 else go[b])
    fds, an = np.array(fds), np.array(an)
    scale = np.abs(fds).max()
    err = np.abs(fds - an)
    print(f"[blocks] {len(fds)} smooth entries, {kinks} with a kink within h; fd max {scale:.3e} "
          f"max abs err {err.max():.3e} rel {err.max() / scale:.2e}", flush=True)
    assert kinks <= 8
    assert (err <= 1e-3 * scale + 2e-3 * np.abs(fds)).all(), (err.max(), scale)


def test_blocks_through_rasteriser_autograd():
    """Block kernel -> GaussianRasterizer through torch autograd gives the same
    block gradients as chaining the two adjoints by hand (each adjoint is
    pinned by its own finite-difference test)."""
    from partgs_amd.blocks import _desc  # noqa: F401
    from partgs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer

    params, occ = _blocks(2, 4)
    grid, p, o, (means, scales, rots, opac) = _run_gpu(params, occ, 16, 24)
    cam = random_sphere_camera(3, 3.0, 96, 72).to("cuda")
    P = means.shape[0]
    rng = np.random.default_rng(9)
    colors = torch.tensor(rng.uniform(0, 1, size=(P, 3)), dtype=torch.float32, device="cuda")
    gc = torch.tensor(rng.normal(size=(3, 72, 96)), dtype=torch.float32, device="cuda")
    go = torch.tensor(rng.normal(size=(7, 72, 96)), dtype=torch.float32, device="cuda")
    tx = ty = math.tan(cam.FoVx * 0.5)
    rs = GaussianRasterizationSettings(image_height=72, image_width=96, tanfovx=tx, tanfovy=ty,
                                       bg=torch.ones(3, device="cuda"), scale_modifier=1.0,
                                       viewmatrix=cam.world_view_transform, projmatrix=cam.full_proj_transform,
                                       sh_degree=0, campos=cam.camera_center, prefiltered=False, debug=False)
    means2D = torch.zeros_like(means, requires_grad=True)
    color, radii, allmap = GaussianRasterizer(rs)(means3D=means, means2D=means2D, opacities=opac,
                                                  colors_precomp=colors, scales=scales, rotations=rots)
    assert int((radii > 0).sum()) > 0
    torch.autograd.backward([color, allmap], [gc, go])
    # by hand: rasteriser adjoint, then the block adjoint
    e = torch.empty(0, device="cuda")
    args = (rs.bg, means.detach(), colors, opac.detach(), scales.detach(), rots.detach(), 1.0, e,
            rs.viewmatrix, rs.projmatrix, tx, ty, 72, 96, e, 0, rs.campos, False, False)
    L, _, _, radii2, geom, binb, img = rasterize_gaussians_fwd(*args)
    g = rasterize_gaussians_bwd(rs.bg, means.detach(), radii2, colors, scales.detach(), rots.detach(), 1.0, e,
                                rs.viewmatrix, rs.projmatrix, tx, ty, gc, go, e, 0, rs.campos, geom, L, binb, img,
                                False)
    p2 = p.detach().clone().requires_grad_(True)
    o2 = o.detach().clone().requires_grad_(True)
    outs2 = generate_surfels(p2, o2, grid)
    torch.autograd.backward(list(outs2), [g[3], g[6], g[7], g[2]])
    for a, b_, name in ((p.grad, p2.grad, "params"), (o.grad, o2.grad, "occupancy")):
        a, b_ = a.cpu().numpy(), b_.cpu().numpy()
        assert np.abs(a - b_).max() <= 1e-4 * np.abs(b_).max() + 1e-7, name


def test_block_backward_is_deterministic():
    """The block gradients are summed in a fixed order (per-surfel
    contributions, then one workgroup per block and parameter): repeated
    backward passes give the same bits."""
    params, occ = _blocks(4, 16)
    grid, p, o, outs = _run_gpu(params, occ, 64, 64)
    print("[blocks] determinism B=16 64x64", flush=True)
    rng = np.random.default_rng(11)
    gs = [torch.tensor(rng.normal(size=t.shape), dtype=torch.float32, device="cuda") for t in outs]
    res = []
    for _ in range(3):
        p.grad = None
        o.grad = None
        outs = generate_surfels(p, o, grid)
        torch.autograd.backward(list(outs), gs)
        res.append((p.grad.clone(), o.grad.clone()))
    for a, b in res[1:]:
        assert torch.equal(a, res[0][0]) and torch.equal(b, res[0][1])
