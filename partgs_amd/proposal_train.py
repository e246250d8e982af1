"""Train and evaluate the learned block proposal.

  python -m partgs_amd.proposal_train --train 256 --eval 16 --fit-iters 200 --out DIR

Labelling dominates the cost (one plain fit per training scene), so it can
be split over several runs: `--label-seed S --labels-out F --label-only`
labels scenes S, S+1, ... and saves their seeds and fitted blocks (the
points are regenerated from the seeds); `--labels-in F1 F2 ...` trains on
saved labels instead of labelling.  The evaluation splits the same way:
`--train-only` trains and saves the weights; `--weights W --eval N
--eval-seed S` evaluates held-out scenes S .. S+N-1; `--merge R1 R2 ...`
combines the evaluation rows of several reports into one.

1. Dataset: synthetic block scenes (partgs_amd.block_fit.random_block_scene,
   4-16 superquadrics, 20,000 noisy surface points each), each labelled by
   the plain pipeline: the k-means initialisation of partgs_amd.init with 16
   blocks, refined by the render-based plain optimiser (fit_blocks).
2. Training: the point-set MLP (partgs_amd.proposal) on the MFMA kernels,
   loss = per-scene Hungarian-matched weighted L2 between the 16 predicted
   and the 16 fitted blocks (q compared up to sign).
3. Evaluation on held-out scenes: the plain optimiser run from the plain
   initialisation and from the network's proposal with the same optimiser
   and loss; reported: iterations the proposal needs to reach 1.01 x the
   plain arm's final loss, against the plain arm's iteration count.  The
   plain initialisation stays the fallback: a scene where the proposal's
   final loss is worse keeps the plain result.

Writes DIR/proposal_weights.npz and DIR/proposal_report.json; prints one
progress line per stage item.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

from .block_fit import block_scene_points, fit_blocks, iterations_to_reach, logit, random_block_scene
from .block_losses import BlockLossWeights
from .init import plain_block_init
from .proposal import N_BLOCKS, ProposalNet

W_PARAM = torch.tensor([1.0] * 3 + [0.5] * 4 + [1.0] * 3 + [0.1] * 2)


def _log(msg):
    print(msg, flush=True)


def canonical(params: torch.Tensor) -> torch.Tensor:
    """Unit quaternions with w >= 0 (the block layer normalises q anyway)."""
    p = params.clone()
    q = p[..., 3:7]
    q = q / q.norm(dim=-1, keepdim=True).clamp_min(1e-8)
    q = torch.where(q[..., :1] < 0, -q, q)
    p[..., 3:7] = q
    return p


def plain_init(points: torch.Tensor, seed: int):
    params, occ = plain_block_init(points, n_blocks=N_BLOCKS, ratio_block_scene=0.25, seed=seed)
    return params, occ


def label_scene(seed: int, fit_iters: int, block_losses=None):
    scene = random_block_scene(seed)
    p0, o0 = plain_init(scene.points, seed)
    fit = fit_blocks(scene, p0, o0, iters=fit_iters, block_losses=block_losses, seed=seed)
    return scene, canonical(fit.params), fit.losses


def matched_loss(pred: torch.Tensor, label: torch.Tensor):
    """Hungarian-matched weighted L2 per scene; pred requires grad (B x 16 x 12)."""
    from scipy.optimize import linear_sum_assignment

    w = W_PARAM.to(pred.device)
    total = 0.0
    pd = pred.detach()
    for b in range(pred.shape[0]):
        d = pd[b][:, None, :] - label[b][None, :, :]
        dq_neg = pd[b][:, None, 3:7] + label[b][None, :, 3:7]
        cost_q = torch.minimum((d[..., 3:7] ** 2).sum(-1), (dq_neg ** 2).sum(-1))
        cost = ((d ** 2) * w).sum(-1) - (d[..., 3:7] ** 2 * w[3:7]).sum(-1) + w[3] * cost_q
        ri, ci = linear_sum_assignment(cost.cpu().numpy())
        ri = torch.as_tensor(ri, device=pred.device)
        ci = torch.as_tensor(ci, device=pred.device)
        tgt = label[b][ci].clone()
        flip = ((pred[b][ri, 3:7] - tgt[:, 3:7]) ** 2).sum(-1) > ((pred[b][ri, 3:7] + tgt[:, 3:7]) ** 2).sum(-1)
        tgt[:, 3:7] = torch.where(flip[:, None], -tgt[:, 3:7], tgt[:, 3:7])
        total = total + (((pred[b][ri] - tgt) ** 2) * w).sum() / N_BLOCKS
    return total / pred.shape[0]


def train(net: ProposalNet, points: torch.Tensor, labels: torch.Tensor, epochs: int, batch: int, lr: float,
          seed: int = 0):
    n = points.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g)
        ep_loss = 0.0
        for s in range(0, n, batch):
            idx = perm[s:s + batch].to(points.device)
            out, cache = net.forward(points[idx])
            o = out.detach().requires_grad_(True)
            loss = matched_loss(o, labels[idx])
            loss.backward()
            grads = net.backward(o.grad, cache)
            net.adam_step(grads, lr=lr)
            ep_loss += float(loss) * idx.numel()
        hist.append(ep_loss / n)
        if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
            _log(f"[proposal] epoch {ep} loss {hist[-1]:.4f}")
    return hist


def evaluate(net: ProposalNet, seeds, fit_iters: int, block_losses=None):
    rows = []
    for sd in seeds:
        scene = random_block_scene(sd)
        p0, o0 = plain_init(scene.points, sd)
        plain = fit_blocks(scene, p0, o0, iters=fit_iters, block_losses=block_losses, seed=sd)
        with torch.no_grad():
            out, _ = net.forward(scene.points[None])
        pl = canonical(out[0])
        ol = torch.full((N_BLOCKS,), logit(0.7), device=pl.device)
        learned = fit_blocks(scene, pl, ol, iters=fit_iters, block_losses=block_losses, seed=sd)
        final_plain = float(np.mean(plain.losses[-5:]))
        level = 1.01 * final_plain
        it_learned = iterations_to_reach(learned.losses, level)
        it_plain = iterations_to_reach(plain.losses, level)
        final_learned = float(np.mean(learned.losses[-5:]))
        rows.append({"seed": sd, "final_plain": final_plain, "final_learned": final_learned,
                     "iters_plain": fit_iters, "iters_plain_first_hit": it_plain, "iters_learned": it_learned,
                     "init_loss_plain": plain.losses[0], "init_loss_learned": learned.losses[0],
                     "fallback_to_plain": final_learned > final_plain})
        _log(f"[proposal] eval seed {sd}: plain final {final_plain:.4f} (first hit {it_plain}) learned hits at "
             f"{it_learned}, final {final_learned:.4f}")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--eval", type=int, default=16)
    ap.add_argument("--fit-iters", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default="proposal_out")
    ap.add_argument("--weights", default=None, help="skip training and load these weights")
    ap.add_argument("--label-seed", type=int, default=10_000, help="first training-scene seed")
    ap.add_argument("--labels-out", default=None, help="save the labelled seeds and blocks here")
    ap.add_argument("--labels-in", nargs="*", default=None, help="train on these saved labels")
    ap.add_argument("--label-only", action="store_true", help="label (and save), then stop")
    ap.add_argument("--eval-seed", type=int, default=90_000, help="first held-out scene seed")
    ap.add_argument("--train-only", action="store_true", help="train and save the weights, then stop")
    ap.add_argument("--merge", nargs="*", default=None, help="combine the evaluation rows of these reports")
    ap.add_argument("--no-block-losses", action="store_true",
                    help="fit with the image term alone (default: the block stage's full objective)")
    args = ap.parse_args(argv)
    if args.merge:
        merged = merge_reports(args.merge, args.out)
        _log(json.dumps({k: v for k, v in merged.items() if k != "eval"}))
        return 0
    os.makedirs(args.out, exist_ok=True)
    torch.manual_seed(0)
    net = ProposalNet(seed=0)
    # a run on saved weights reports no training-set size (the training run's report holds it)
    report = {"train_scenes": None if args.weights else args.train, "eval_scenes": args.eval,
              "fit_iters": args.fit_iters,
              "block_losses": not args.no_block_losses}
    bl = None if args.no_block_losses else BlockLossWeights()
    t0 = time.time()
    if args.weights:
        sd = np.load(args.weights)
        net.load_state_dict({k: torch.tensor(sd[k]) for k in sd.files})
    else:
        pts, labels, seeds = [], [], []
        if args.labels_in:
            for fn in args.labels_in:
                z = np.load(fn)
                for sd, lb in zip(z["seeds"], z["labels"]):
                    seeds.append(int(sd))
                    pts.append(block_scene_points(int(sd)))
                    labels.append(torch.tensor(lb, device=pts[-1].device))
                    if len(seeds) % 1000 == 0:
                        _log(f"[proposal] regenerated {len(seeds)} scenes ({time.time() - t0:.0f} s)")
            _log(f"[proposal] loaded {len(seeds)} labelled scenes from {len(args.labels_in)} file(s)")
        else:
            for i in range(args.train):
                sd = args.label_seed + i
                scene, label, losses = label_scene(sd, args.fit_iters, bl)
                seeds.append(sd)
                pts.append(scene.points)
                labels.append(label)
                if i % 16 == 0:
                    _log(f"[proposal] labelled {i + 1}/{args.train} scenes, fit loss {losses[0]:.4f} -> "
                         f"{losses[-1]:.4f} ({time.time() - t0:.0f} s)")
                if args.labels_out and (i + 1) % 64 == 0:  # kept if the run is cut short
                    np.savez(args.labels_out, seeds=np.array(seeds, np.int64),
                             labels=torch.stack(labels).cpu().numpy())
        points = torch.stack(pts)
        labels = torch.stack(labels)
        report["train_scenes"] = len(seeds)
        report["label_seconds"] = time.time() - t0
        if args.labels_out:
            np.savez(args.labels_out, seeds=np.array(seeds, np.int64), labels=labels.cpu().numpy())
        if args.label_only:
            _log(json.dumps(report))
            return 0
        t1 = time.time()
        hist = train(net, points, labels, args.epochs, args.batch, args.lr)
        report["train_seconds"] = time.time() - t1
        report["train_loss"] = [hist[0], hist[-1]]
        np.savez(os.path.join(args.out, "proposal_weights.npz"),
                 **{k: v.numpy() for k, v in net.state_dict().items()})
    if args.train_only:
        with open(os.path.join(args.out, "proposal_report.json"), "w") as f:
            json.dump(report, f, indent=1)
        _log(json.dumps(report))
        return 0
    rows = evaluate(net, range(args.eval_seed, args.eval_seed + args.eval), args.fit_iters, bl)
    report["eval_seed"] = args.eval_seed
    report["eval"] = rows
    report.update(summarize(rows, args.fit_iters))
    with open(os.path.join(args.out, "proposal_report.json"), "w") as f:
        json.dump(report, f, indent=1)
    _log(json.dumps({k: v for k, v in report.items() if k != "eval"}))
    return 0


def summarize(rows, fit_iters: int) -> dict:
    """Medians over evaluation rows: the proposal's iterations to 1.01 x the
    plain arm's final loss against the plain arm's total and first-hit
    iterations; the share of scenes reached within the budget and the share
    that falls back to the plain result."""
    return {
        "eval_scenes": len(rows),
        "median_iters_ratio_vs_plain_total": float(np.median([r["iters_learned"] / r["iters_plain"] for r in rows])),
        "median_iters_ratio_vs_plain_first_hit": float(
            np.median([r["iters_learned"] / r["iters_plain_first_hit"] for r in rows])),
        "reached_within_1pct": float(np.mean([r["iters_learned"] <= fit_iters for r in rows])),
        "fallback_fraction": float(np.mean([r["fallback_to_plain"] for r in rows])),
    }


def merge_reports(paths, out: str) -> dict:
    """One report from the evaluation rows of several (rows ordered by seed,
    duplicates rejected); the other fields come from the first report."""
    reports = [json.load(open(p)) for p in paths]
    rows = sorted((r for rep in reports for r in rep["eval"]), key=lambda r: r["seed"])
    seeds = [r["seed"] for r in rows]
    if len(set(seeds)) != len(seeds):
        raise ValueError("the reports evaluate some scene twice")
    merged = {k: v for k, v in reports[0].items() if k not in ("eval", "eval_seed")}
    merged["eval"] = rows
    merged.update(summarize(rows, merged["fit_iters"]))
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "proposal_report.json"), "w") as f:
        json.dump(merged, f, indent=1)
    return merged

if __name__ == "__main__":
    sys.exit(main())
