"""The proposal network's MFMA kernels against the CPU oracle.

GEMM: bf16 products are exact in FP32, so only the FP32 accumulation order
differs: 1e-5 relative.  Fused pool and the full network: the oracle rounds
at the same points, but an FP32 sum that lands on the other side of a bf16
rounding boundary moves that activation by one bf16 step, so values compare
at 2e-2 of the tensor's scale and the argmax only where the maximum is
clear."""
import numpy as np
import pytest
import torch

from oracle import proposal_oracle as PO
from partgs_amd import proposal as SP

pytestmark = pytest.mark.gpu


def _bf(x):
    return torch.tensor(PO.bf16_bits(np.asarray(x, np.float32)).view(np.int16), device="cuda")


@pytest.mark.parametrize("M,N,K", [(37, 70, 45), (64, 64, 16), (130, 5, 300), (1, 192, 128)])
def test_gemm_matches_numpy(M, N, K):
    print(f"[proposal] gemm {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M * 1000 + N)
    A = PO.bf16_round(rng.normal(size=(M, K)).astype(np.float32))
    Bt = PO.bf16_round(rng.normal(size=(K, N)).astype(np.float32))  # stored K x N: B(n, k) = Bt[k, n]
    bias = rng.normal(size=N).astype(np.float32)
    want = A.astype(np.float64) @ Bt.astype(np.float64) * 0.5 + bias
    C = SP.gemm(_bf(A), _bf(Bt), M=M, N=N, K=K, a_strides=(K, 1), b_strides=(1, N), bias=torch.tensor(
        bias, device="cuda"), alpha=0.5)
    got = C.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-5 * (np.abs(want).max() + 1), np.abs(got - want).max()
    # ReLU + accumulate + bf16 copy, A read transposed (stored K x M)
    At = np.ascontiguousarray(A.T)
    C2 = torch.ones((M, N), device="cuda")
    out, Cb = SP.gemm(_bf(At), _bf(Bt), M=M, N=N, K=K, a_strides=(1, M), b_strides=(1, N), relu=True, out=C2,
                      accumulate=True, want_bf16=True)
    want2 = np.maximum(A.astype(np.float64) @ Bt, 0) + 1.0
    assert np.abs(out.cpu().numpy() - want2).max() <= 1e-5 * (np.abs(want2).max() + 1)
    assert np.array_equal(Cb.cpu().numpy().view(np.uint16), PO.bf16_bits(out.cpu().numpy()))


def _net(seed=0):
    net = SP.ProposalNet(seed=seed)
    return net, {k: v.cpu().numpy() for k, v in net.p.items()}


def test_pool_matches_oracle():
    print("[proposal] pool", flush=True)
    net, prm = _net(1)
    rng = np.random.default_rng(2)
    pts = rng.normal(size=(3, 1000, 3)).astype(np.float32)
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
        assert np.array_equal(idx[b].cpu().numpy()[clear], h3.argmax(0)[clear])


def test_forward_backward_match_oracle():
    print("[proposal] fwd/bwd", flush=True)
    net, prm = _net(3)
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(4, 700, 3)).astype(np.float32)
    out, cache = net.forward(torch.tensor(pts, device="cuda"))
    o = PO.Oracle(prm, bf16=True)
    want, oc = o.forward(pts)
    got = out.reshape(4, -1).cpu().numpy()
    assert np.abs(got - want).max() <= 2e-2 * np.abs(want).max()
    dout = rng.normal(size=(4, 192)).astype(np.float32)
    G = net.backward(torch.tensor(dout, device="cuda"), cache)
    # backward through the GPU's own argmax choice, so routing differences do not count
    oc["idx"] = cache["idx"].cpu().numpy().astype(np.int64)
    oc["g"] = cache["g"].cpu().numpy()
    oc["z1"] = cache["z1"].cpu().numpy()
    GW = o.backward(dout, oc)
    for k, want_g in GW.items():
        g = G[k].cpu().numpy()
        scale = np.abs(want_g).max() + 1e-12
        assert np.abs(g - want_g).max() <= 2e-2 * scale, (k, np.abs(g - want_g).max() / scale)


def test_adam_kernel_matches_formula():
    print("[proposal] adam", flush=True)
    net, prm = _net(5)
    rng = np.random.default_rng(6)
    grads = {k: torch.tensor(rng.normal(size=v.shape).astype(np.float32), device="cuda") for k, v in prm.items()}
    net.adam_step(grads, lr=1e-2)
    net.adam_step(grads, lr=1e-2)
    for k, p0 in prm.items():
        g = grads[k].cpu().numpy()
        p, m, v = PO.adam(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-2)
        p, m, v = PO.adam(p, g, m, v, 2, 1e-2)
        assert np.allclose(net.p[k].cpu().numpy(), p, rtol=1e-5, atol=1e-6), k
    for k in SP.BF16_PARAMS:
        assert np.array_equal(net.pb[k].cpu().numpy().view(np.uint16), PO.bf16_bits(net.p[k].cpu().numpy()))


def test_training_fits_fixed_targets():
    """Gradient descent through the MFMA forward/backward drives a regression
    onto fixed targets down by an order of magnitude."""
    print("[proposal] train", flush=True)
    net, _ = _net(7)
    rng = np.random.default_rng(8)
    pts = torch.tensor(rng.normal(size=(8, 512, 3)).astype(np.float32), device="cuda")
    target = torch.tensor(rng.normal(size=(8, 16, 12)).astype(np.float32), device="cuda") * 0.5
    losses = []
    for it in range(300):
        out, cache = net.forward(pts)
        diff = out - target
        losses.append(float((diff ** 2).mean()))
        grads = net.backward(2 * diff / diff.numel(), cache)
        net.adam_step(grads, lr=3e-3)
    print(f"[proposal] loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < 0.1 * losses[0]


def test_block_fit_reduces_image_loss():
    from partgs_amd.block_fit import fit_blocks, random_block_scene
    from partgs_amd.init import plain_block_init

    print("[proposal] block fit", flush=True)
    scene = random_block_scene(123, n_points=4000, views=3, width=64, height=48)
    p0, o0 = plain_block_init(scene.points, n_blocks=8, seed=0)
    fit = fit_blocks(scene, p0, o0, iters=40)
    print(f"[proposal] fit loss {fit.losses[0]:.4f} -> {fit.losses[-1]:.4f}", flush=True)
    assert all(np.isfinite(fit.losses))
    assert fit.losses[-1] < fit.losses[0]


def test_protocol_runs_end_to_end(tmp_path):
    from partgs_amd import proposal_train

    print("[proposal] protocol smoke", flush=True)
    proposal_train.main(["--train", "2", "--eval", "1", "--fit-iters", "10", "--epochs", "3", "--batch", "2",
                         "--out", str(tmp_path)])
    import json

    rep = json.load(open(tmp_path / "proposal_report.json"))
    assert rep["eval"] and (tmp_path / "proposal_weights.npz").exists()


def test_scene_points_match_full_scene():
    from partgs_amd.block_fit import block_scene_points, random_block_scene

    print("[proposal] scene points", flush=True)
    for sd in (5, 90001):
        assert torch.equal(block_scene_points(sd), random_block_scene(sd).points)


def test_block_fit_with_block_losses():
    from partgs_amd.block_fit import fit_blocks, random_block_scene
    from partgs_amd.block_losses import BlockLossWeights
    from partgs_amd.init import plain_block_init

    print("[proposal] block fit with the block-stage objective", flush=True)
    scene = random_block_scene(321, n_points=4000, views=3, width=64, height=48)
    assert scene.masks is not None and len(scene.masks) == 3 and scene.masks[0].shape == (48, 64)
    p0, o0 = plain_block_init(scene.points, n_blocks=8, seed=0)
    fit = fit_blocks(scene, p0, o0, iters=40, block_losses=BlockLossWeights(), near_far=(1.0, 5.0), seed=1)
    print(f"[proposal] fit loss {fit.losses[0]:.4f} -> {fit.losses[-1]:.4f}", flush=True)
    assert all(np.isfinite(fit.losses)) and fit.losses[-1] < fit.losses[0]
