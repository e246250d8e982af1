"""Benchmark: forward + backward of one 800x600 view of 300,000 surfels.

Workload: surfels sampled on 8 random superquadrics (SH degree 3), one camera
at radius 3, tile 16, upstream gradients N(0, 1) on all 10 output channels;
a step is GaussianRasterizer.forward + backward through the drop-in API with
all inputs resident on the GPU.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+bwd ms/view (300k surfels, 800x600), %HBM peak; iters-to-loss vs plain"


def algorithmic_bytes(P_vis, M, L, N, T):
    """Compulsory HBM bytes per stage (each datum read or written once)."""
    return {
        "preprocess": (40 + 12 * M) * P_vis + 87 * P_vis,
        "tile_scan": 8 * T,
        "scatter": 20 * P_vis + 8 * L,
        "tile_sort": 8 * L + 4 * L,
        "blend_fwd": 8 * T + 76 * L + 60 * N,
        "grad_clear": 0,
        # the backward blend's gather and pixel inputs in its per-pixel pass,
        # its per-surfel gradient output in the per-pair pass
        "blend_bwd": 8 * T + 76 * L + 60 * N,
        "blend_bwd_grad": 76 * P_vis,
        "preprocess_bwd": (151 + 12 * M) * P_vis + (48 + 12 * M) * P_vis,
        # passes with no counterpart in the reference's byte model
        "accept": 0,
        "pair_store": 0,
    }


def pipeline_bytes(P, M, L, N, T):
    """B_alg(view) = (430 + 36 M) P + 196 L + 120 N + 24 T (fwd + bwd)."""
    return (430 + 36 * M) * P + 196 * L + 120 * N + 24 * T


ROOFLINE_STAGES = ("preprocess", "scatter", "blend_fwd", "blend_bwd", "blend_bwd_grad", "preprocess_bwd")

STAGE_KERNEL = {"preprocess": "k_preprocess", "tile_scan": "k_scan_tiles", "scatter": "k_scatter",
                "tile_sort": "k_sort_tiles", "blend_fwd": "k_blend_fwd", "blend_bwd": "k_blend_bwd",
                "preprocess_bwd": "k_preprocess_bwd", "accept": "k_accept", "blend_bwd_grad": "k_blend_bwd_grad",
                "pair_store": "k_pair_write"}


def measured_traffic(stage):
    """HBM bytes per launch of a stage's kernel from the committed counter
    summary (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes of this benchmark,
    each scaled by the same counter's reading for a copy of known size), or
    None when no summary covers it."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_c2b.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(STAGE_KERNEL.get(stage, stage))
    return None if k is None else int(k["bytes_per_launch"])


def cpu_baseline(scene, gc, go, budget_s=20.0):
    import numpy as np  # noqa: F401

    from oracle import oracle as O

    # the port compiled for this host; the shipped portable build otherwise
    native = O.use_native_build()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.set_threads(threads)
    cam = scene.camera
    kw = dict(means3D=scene.means3D, opacities=scene.opacities, viewmatrix=cam.world_view_transform,
              projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=scene.bg, W=cam.image_width,
              H=cam.image_height, tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5),
              shs=scene.shs, sh_degree=scene.sh_degree, scales=scene.scales, rotations=scene.rotations)
    gcn, gon = gc.numpy(), go.numpy()
    times = []
    o = None
    t_start = time.perf_counter()
    while len(times) < 5 and (time.perf_counter() - t_start) < budget_s:
        o = O.Oracle()
        t0 = time.perf_counter()
        o.forward(**kw)
        o.backward(gcn, gon)
        times.append((time.perf_counter() - t0) * 1e3)
        print(f"[bench] cpu oracle fwd+bwd run {len(times)}: {times[-1]:.1f} ms", file=sys.stderr, flush=True)
    base = {"value": statistics.median(times), "unit": "ms/view", "cores": O.get_threads(), "kind": "port",
            "sample": f"{len(times)} full fwd+bwd views of the same 300k-surfel 800x600 workload (median)",
            "build": "g++ -O3 -march=native -fopenmp" if native else "g++ -O3 -mavx2 -mfma -fopenmp",
            "note": "the C++ restatement of the reference's per-pixel logic: OpenMP over tiles, per-thread "
                    "FP64 gradient buffers, per-tile parallel sort; a reported baseline, not the target"}
    return base, o


GRAD_TOL = 1e-5


def oracle_parity(o, color, allmap, radii, grads):
    """The GPU step against the oracle run of the same workload: images,
    auxiliary maps and radii bit for bit, every gradient within GRAD_TOL of
    the oracle's FP64-accumulated sums per element (|a - b| <= tol (|b| +
    max|b|)).  grads maps the oracle's names to the GPU's arrays."""
    import numpy as np

    worst = 0.0
    for name, g in grads.items():
        ref = np.asarray(o.get(name), dtype=np.float64)
        a = np.asarray(g, dtype=np.float64).reshape(ref.shape)
        if ref.size:
            worst = max(worst, float((np.abs(a - ref) / (np.abs(ref) + np.abs(ref).max() + 1e-300)).max()))
    images = bool(np.array_equal(color, o.get("out_color")) and np.array_equal(allmap, o.get("out_others")))
    radii_ok = bool(np.array_equal(radii, o.get("radii")))
    return {"images_bitwise": images, "radii_bitwise": radii_ok, "grad_max_scaled_err": worst,
            "grad_tol": GRAD_TOL, "ok": images and radii_ok and worst <= GRAD_TOL}


def run_batched_groups(args, scene, cams, cfg):
    """The batched workload in view groups that share one workspace: each
    group's forward and backward run in turn inside one hipGraph and the
    per-surfel gradients are summed over the groups."""
    import torch

    from partgs_amd.batch import ViewGroups, capture_grouped_step

    dev = scene.means3D.device
    vg = ViewGroups(cams, scene.bg, scene.sh_degree, groups=args.groups, config=cfg)
    gen = torch.Generator(device="cpu").manual_seed(0)
    gc = torch.randn(3 * vg.N, generator=gen).to(dev)
    go = torch.randn(7 * vg.N, generator=gen).to(dev)
    step = capture_grouped_step(vg, scene.means3D, scene.shs, None, scene.opacities, scene.scales, scene.rotations,
                                gc, go)
    for _ in range(args.warmup):
        step.replay()
    torch.cuda.synchronize()
    if not step.ok():
        raise RuntimeError("capacity exceeded")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step.replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    L = sum(br.read_status()[0] for br in vg.parts)
    out = {
        "metric": "batched fwd+bwd ms/view (1M surfels, 32 ragged views, one hipGraph)",
        "value": round(ms / len(cams), 4), "unit": "ms/view", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (8 random superquadrics, 32 sphere cameras, seeded)",
        "config": {"workload": "1M superquadric surfels, SH degree 3, 32 views from {400x300, 640x480, 800x600}",
                   "views": len(cams), "groups": args.groups, "views_per_group": [br.V for br in vg.parts],
                   "pixels": vg.N, "tile": cfg.tile, "pair_store": cfg.pair_store, "num_rendered": L,
                   "workspace_gb": round(vg.workspace.numel() / 1e9, 2)},
    }
    print(json.dumps(out))


def run_batched(args):
    """32 cameras on the radius-3 sphere, resolutions drawn from {400x300,
    640x480, 800x600} (seed 0), 1,000,000 superquadric surfels (SH degree 3),
    forward + backward of all views replayed from one hipGraph."""
    import numpy as np
    import torch

    from partgs_amd import _lib
    from partgs_amd.batch import BatchConfig, BatchRasterizer, autotune, capture_step
    from partgs_amd.scenes import batched_views, superquadric_surfels

    dev = torch.device("cuda:0")
    scene = superquadric_surfels(0, P=1_000_000, width=800, height=600, sh_degree=3).to(dev)
    cams = [c.to(dev) for c in batched_views(0, 32)]
    print(f"[bench] batched: {len(cams)} views, {scene.P} surfels", file=sys.stderr, flush=True)
    cfg = BatchConfig()
    if args.groups > 1:
        return run_batched_groups(args, scene, cams, cfg)
    if args.autotune:
        cfg = autotune(cams, scene.bg, scene.sh_degree, scene.means3D, scene.shs, None, scene.opacities,
                       scene.scales, scene.rotations, tiles=(8, 16, 32), pair_stores=(0, 8), reps=2,
                       log=lambda m: print(m, file=sys.stderr, flush=True))
    br = BatchRasterizer(cams, scene.bg, scene.sh_degree, cfg)
    gen = torch.Generator(device="cpu").manual_seed(0)
    gc = torch.randn(3 * br.N, generator=gen).to(dev)
    go = torch.randn(7 * br.N, generator=gen).to(dev)
    step = capture_step(br, scene.means3D, scene.shs, None, scene.opacities, scene.scales, scene.rotations, gc, go)
    for _ in range(args.warmup):
        step.replay()
    torch.cuda.synchronize()
    if not step.ok():
        raise RuntimeError("capacity exceeded")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    L, _ = br.read_status()
    out = {
        "metric": "batched fwd+bwd ms/view (1M surfels, 32 ragged views, one hipGraph)",
        "value": round(ms / len(cams), 4), "unit": "ms/view", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic (8 random superquadrics, 32 sphere cameras, seeded)",
        "config": {"workload": "1M superquadric surfels, SH degree 3, 32 views from {400x300, 640x480, 800x600}",
                   "views": len(cams), "pixels": br.N, "tile": cfg.tile, "pair_store": cfg.pair_store,
                   "num_rendered": L, "workspace_gb": round(br.workspace.numel() / 1e9, 2)},
    }
    print(json.dumps(out))


def run_kernels(args):
    """Event timings of distCUDA2 (simple-knn) on superquadric point clouds of
    100k / 300k / 1M points, and of the superquadric signed distance of 16
    blocks (forward, forward + adjoint) at the overlap loss's 16 x 2048 points
    and the ray losses' 2048 x 128 samples."""
    import numpy as np
    import torch

    from partgs_amd import block_losses as BL
    from partgs_amd.knn import distCUDA2
    from partgs_amd.scenes import superquadric_surfels

    dev = torch.device("cuda:0")

    def ev_ms(fn, reps=20):
        fn()
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        torch.cuda.synchronize()
        return round(s.elapsed_time(e) / reps, 4)

    out = {"metric": "kernel timings (ms per call)", "n_gpus": 1, "data": "synthetic (seeded)"}
    for P in (100_000, 300_000, 1_000_000):
        pts = superquadric_surfels(0, P=P).means3D.to(dev)
        out[f"distCUDA2_{P}"] = ev_ms(lambda: distCUDA2(pts))
    rng = np.random.default_rng(0)
    prm = np.concatenate([np.log(rng.uniform(0.1, 0.5, (16, 3))), rng.normal(size=(16, 4)),
                          rng.normal(size=(16, 3)) * 0.3, rng.normal(size=(16, 2))], 1)
    params = torch.tensor(prm.astype(np.float32), device=dev)
    for N in (16 * 2048, 2048 * 128):
        pts = torch.rand(N, 3, device=dev) * 2 - 1
        out[f"sdf_fwd_16x{N}"] = ev_ms(lambda: BL.superquadric_sdf(params, pts))
        P = params.clone().requires_grad_(True)
        g = torch.randn(16, N, device=dev)
        out[f"sdf_fwd_bwd_16x{N}"] = ev_ms(lambda: torch.autograd.backward(BL.superquadric_sdf(P, pts), g))
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched", action="store_true",
                    help="32 ragged views of 1,000,000 surfels in one graph-captured batch instead")
    ap.add_argument("--autotune", action="store_true", help="with --batched: time tile / pair-store choices first")
    ap.add_argument("--groups", type=int, default=2,
                    help="with --batched: render the views in this many groups sharing one workspace "
                         "(1: all views in one pass)")
    ap.add_argument("--kernels", action="store_true",
                    help="time the simple-knn and superquadric-distance kernels instead (one JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 10 copies of a 1 GiB buffer (reference kernel for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only (one view per step on one device); "
                                   "sharding by view over several GPUs is not launched",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if args.copy_calibration:
        x = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        y = torch.empty_like(x)
        for _ in range(10):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": x.numel() * 4, "bytes_written": y.numel() * 4,
                          "launches": 10}))
        return

    if args.batched:
        run_batched(args)
        return
    if args.kernels:
        run_kernels(args)
        return

    from partgs_amd import _lib
    from partgs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    from partgs_amd.scenes import superquadric_surfels, upstream_grads

    dev = torch.device("cuda:0")
    scene_cpu = superquadric_surfels(0, P=300_000, width=800, height=600, sh_degree=3)
    scene = scene_cpu.to(dev)
    cam = scene.camera
    H, W = cam.image_height, cam.image_width
    gc_cpu, go_cpu = upstream_grads(0, H, W)
    gc, go = gc_cpu.to(dev), go_cpu.to(dev)
    rs = GaussianRasterizationSettings(image_height=H, image_width=W, tanfovx=math.tan(cam.FoVx * 0.5),
                                       tanfovy=math.tan(cam.FoVy * 0.5), bg=scene.bg, scale_modifier=1.0,
                                       viewmatrix=cam.world_view_transform, projmatrix=cam.full_proj_transform,
                                       sh_degree=scene.sh_degree, campos=cam.camera_center, prefiltered=False,
                                       debug=False, tile=args.tile)
    rast = GaussianRasterizer(rs)
    leaves = [t.clone().requires_grad_(True) for t in (scene.means3D, scene.shs, scene.opacities, scene.scales,
                                                       scene.rotations)]
    means2D = torch.zeros_like(scene.means3D, requires_grad=True)
    info = {}

    def step():
        for t in leaves + [means2D]:
            t.grad = None
        color, radii, allmap = rast(means3D=leaves[0], means2D=means2D, shs=leaves[1], opacities=leaves[2],
                                    scales=leaves[3], rotations=leaves[4])
        torch.autograd.backward((color, allmap), (gc, go))
        info["radii"] = radii
        info["color"], info["allmap"] = color, allmap

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-stage breakdown: K untimed steps with HIP events around every stage
    # (the events themselves cost a few microseconds each, so they stay out
    # of the timed region)
    _lib.reset_timing()
    _lib.set_timing(True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    _lib.set_timing(False)
    stages = _lib.get_timing()
    per_stage_ms = {k: v[0] / args.steps for k, v in stages.items() if v[1] > 0}
    kern_ms = sum(v[0] for v in stages.values()) / args.steps
    # the roofline object covers the slowest single-kernel stage with a byte
    # model; the pair-store passes (internal intermediates of the forward
    # blend) and the two concurrent tile sorts are timed in stages_ms and
    # covered by the pipeline figure and the traffic summary
    candidates = [k for k in per_stage_ms if k in ROOFLINE_STAGES]
    dominant = max(candidates, key=lambda k: per_stage_ms[k])
    # the timed region: K steps, HIP events around the dominant kernel only
    _lib.reset_timing()
    _lib.set_timing_stages([dominant])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _lib.set_timing(False)
    dom_ms, dom_launches = _lib.get_timing()[dominant]
    dom_avg_ms = dom_ms / max(dom_launches, 1)
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # sizes for the byte model, from one more forward
    from partgs_amd.rasterizer import rasterize_gaussians_fwd
    e = torch.empty(0, device=dev)
    L = rasterize_gaussians_fwd(scene.bg, scene.means3D, e, scene.opacities, scene.scales, scene.rotations, 1.0, e,
                                cam.world_view_transform, cam.full_proj_transform, rs.tanfovx, rs.tanfovy, H, W,
                                scene.shs, scene.sh_degree, cam.camera_center, False, False, args.tile)[0]
    P = scene.P
    P_vis = int((info["radii"] > 0).sum())
    M = scene.shs.shape[1]
    N = W * H
    T = ((W + args.tile - 1) // args.tile) * ((H + args.tile - 1) // args.tile)
    alg = algorithmic_bytes(P_vis, M, L, N, T)
    dom_bytes = alg[dominant]
    achieved = dom_bytes / (dom_avg_ms * 1e-3) / 1e9
    b_pipe = pipeline_bytes(P, M, L, N, T)
    cpu = None
    if not args.no_cpu:
        # the last timed step's outputs and gradients, checked against the
        # oracle run of the same workload that the baseline times
        last = {"color": info["color"].detach().cpu().numpy(), "allmap": info["allmap"].detach().cpu().numpy(),
                "radii": info["radii"].cpu().numpy(),
                "grads": {name: t.grad.cpu().numpy() for t, name in
                          zip(leaves + [means2D], ("dL_dmeans3D", "dL_dsh", "dL_dopacity", "dL_dscales",
                                                   "dL_drotations", "dL_dmeans2D"))}}
        cpu, o = cpu_baseline(scene_cpu, gc_cpu, go_cpu)
        cpu["parity"] = oracle_parity(o, last["color"], last["allmap"], last["radii"], last["grads"])
        print(f"[bench] parity against the oracle: {cpu['parity']}", file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": round(ms_per_step, 4),
        "unit": "ms/view",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (8 random superquadrics, seeded)",
        "config": {"workload": "300k superquadric surfels, SH degree 3, 800x600, 1 view fwd+bwd", "P": P,
                   "width": W, "height": H, "sh_degree": scene.sh_degree, "tile": args.tile,
                   "num_rendered": int(L), "visible": P_vis},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": measured_traffic(dominant),
                     "alg_bytes": int(dom_bytes), "avg_ms": round(dom_avg_ms, 5)},
        "pipeline": {"alg_bytes": int(b_pipe), "kernel_ms": round(kern_ms, 5),
                     "hbm_frac": round(b_pipe / (kern_ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 5),
                     "stages_ms": {k: round(v, 5) for k, v in per_stage_ms.items()},
                     "stages_note": "per step, from K further steps with HIP events around every stage"},
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(cpu["value"] / ms_per_step, 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
