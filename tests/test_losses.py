"""Block-stage losses (utils/loss_utils.py:6-55) and the proposal's matched
set loss, pinned by their defining identities."""
import math

import numpy as np
import torch

from partgs_amd.losses import l1_loss, photometric, ssim
from partgs_amd.proposal_train import canonical, matched_loss


def test_ssim_identities():
    g = torch.Generator().manual_seed(0)
    x = torch.rand(3, 40, 50, generator=g)
    assert abs(float(ssim(x, x)) - 1.0) < 1e-6
    y = (x + 0.1 * torch.randn(3, 40, 50, generator=g)).clamp(0, 1)
    s = float(ssim(x, y))
    assert 0.0 < s < 1.0
    assert abs(s - float(ssim(y, x))) < 1e-6
    assert float(photometric(x, x)) < 1e-6
    assert abs(float(l1_loss(x, x + 0.25)) - 0.25) < 1e-6


def test_ssim_of_constant_images_closed_form():
    """Away from the zero-padded border, two constant images a, b have
    sigma = 0 and SSIM = (2ab + C1) / (a^2 + b^2 + C1): an 11-pixel-wide
    image of height 1 has no interior, so use a large image and compare the
    map mean against the closed form with the border's share removed."""
    a, b = 0.3, 0.7
    H = W = 301
    x = torch.full((1, H, W), a)
    y = torch.full((1, H, W), b)
    s = float(ssim(x, y))
    C1 = 0.01 ** 2
    inner = (2 * a * b + C1) / (a * a + b * b + C1)
    frac = ((H - 10) * (W - 10)) / (H * W)
    # the border band (5 pixels) lies between inner and 1 in value: bound the mean
    assert inner * frac + 0.0 * (1 - frac) - 1e-6 <= s <= inner * frac + 1.0 * (1 - frac) + 1e-6


def test_matched_loss_is_permutation_and_sign_invariant():
    g = torch.Generator().manual_seed(1)
    label = canonical(torch.randn(2, 16, 12, generator=g))
    perm = torch.randperm(16, generator=g)
    pred = label[:, perm].clone()
    pred[:, :, 3:7] *= -1  # the same rotations
    assert float(matched_loss(pred, label)) < 1e-10
    noisy = pred + 0.1 * torch.randn(pred.shape, generator=g)
    l1 = float(matched_loss(noisy, label))
    l2 = float(matched_loss(noisy, label[:, torch.randperm(16, generator=g)]))
    assert l1 > 0 and math.isclose(l1, l2, rel_tol=1e-6)


def test_canonical_quaternions():
    q = torch.tensor([[0.0] * 3 + [-2.0, 0.0, 0.0, 0.0] + [0.0] * 5])
    c = canonical(q)
    assert np.allclose(c[0, 3:7].numpy(), [1.0, 0.0, 0.0, 0.0])


def test_proposal_report_merge(tmp_path):
    """Evaluation rows split over several runs merge into the report one run
    over all scenes writes: rows by seed, medians over the union."""
    import json

    from partgs_amd.proposal_train import merge_reports, summarize

    def row(seed, learned, first):
        return {"seed": seed, "iters_plain": 200, "iters_plain_first_hit": first, "iters_learned": learned,
                "fallback_to_plain": learned > 200}

    a = [row(3, 50, 100), row(1, 201, 150)]
    b = [row(2, 20, 200)]
    paths = []
    for i, rows in enumerate((a, b)):
        fn = tmp_path / f"r{i}.json"
        fn.write_text(json.dumps({"fit_iters": 200, "block_losses": True, "eval_seed": 0, "eval": rows}))
        paths.append(str(fn))
    m = merge_reports(paths, str(tmp_path / "m"))
    assert [r["seed"] for r in m["eval"]] == [1, 2, 3]
    assert m["eval_scenes"] == 3
    assert m["median_iters_ratio_vs_plain_total"] == 50 / 200
    assert m["median_iters_ratio_vs_plain_first_hit"] == 0.5
    assert abs(m["reached_within_1pct"] - 2 / 3) < 1e-12 and abs(m["fallback_fraction"] - 1 / 3) < 1e-12
    assert m == json.loads((tmp_path / "m" / "proposal_report.json").read_text())
    assert summarize(a + b, 200)["eval_scenes"] == 3
    import pytest
    with pytest.raises(ValueError):
        merge_reports([paths[0], paths[0]], str(tmp_path / "dup"))
