"""Batched multi-view rasterisation: many cameras of uneven resolution, one
surfel set, one launch per pipeline stage, no host synchronisation, and an
optional hipGraph that replays a whole forward + backward.

The reference renders one view per call (renderer/gaussian_renderer/__init__.py:9-149
inside train.py's loop), each call paying a host round trip for the instance
count.  Here the views share a workspace sized for a capacity of (surfel,
tile) instances; a batch that needs more sets a status flag, and the wrapper
grows the workspace and runs again (outside graph capture).  Per view, the
results are those of :class:`partgs_amd.rasterizer.GaussianRasterizer`.

``autotune`` times the tile edge (8, 16, 32) and the accepted-pair store size
(the split between stored and recomputed pair evaluations) on the actual
batch and keeps the fastest.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from dataclasses import dataclass, field

import torch

from . import _lib


@dataclass
class BatchConfig:
    tile: int = 16
    # accepted-pair store per instance; 0 sizes it from the pairs the last
    # forward asked for (status words 4-5), with an eighth of headroom
    pair_store: int = 0


class BatchRasterizer:
    """Rasterises V <= 32 views of one surfel set together.

    The forward state lives in one workspace: run a forward and its backward
    before the next forward on the same object.

    cameras: objects with image_width, image_height, FoVx, FoVy,
    world_view_transform, full_proj_transform, camera_center (the reference's
    Camera / MiniCam fields).
    """

    def __init__(self, cameras, bg: torch.Tensor, sh_degree: int, config: BatchConfig | None = None,
                 scale_modifier: float = 1.0, capacity: int | None = None, device="cuda"):
        if not 1 <= len(cameras) <= _lib.SR_MAX_VIEWS:
            raise ValueError(f"1..{_lib.SR_MAX_VIEWS} views per batch (got {len(cameras)})")
        self.lib = _lib.load()
        self.device = torch.device(device)
        self.config = config or BatchConfig()
        self.V = len(cameras)
        self.widths = [int(c.image_width) for c in cameras]
        self.heights = [int(c.image_height) for c in cameras]
        self.tanx = [math.tan(c.FoVx * 0.5) for c in cameras]
        self.tany = [math.tan(c.FoVy * 0.5) for c in cameras]
        f32 = dict(dtype=torch.float32, device=self.device)
        self.viewmatrices = torch.stack([c.world_view_transform.to(**f32).reshape(16) for c in cameras]).contiguous()
        self.projmatrices = torch.stack([c.full_proj_transform.to(**f32).reshape(16) for c in cameras]).contiguous()
        self.campos = torch.stack([c.camera_center.to(**f32).reshape(3) for c in cameras]).contiguous()
        self.bg = bg.to(**f32).contiguous()
        self.sh_degree = int(sh_degree)
        self.scale_modifier = float(scale_modifier)
        self.pixels = [w * h for w, h in zip(self.widths, self.heights)]
        self.pix_base = [0]
        for n in self.pixels[:-1]:
            self.pix_base.append(self.pix_base[-1] + n)
        self.N = sum(self.pixels)
        self.capacity = int(capacity) if capacity else 0
        self.pair_capacity = 0
        self.slot_capacity = 0
        self.workspace = None
        self.status = torch.zeros(16, dtype=torch.int32, device=self.device)
        self._P = None
        self._host = dict(w=(C.c_int * self.V)(*self.widths), h=(C.c_int * self.V)(*self.heights),
                          tx=(C.c_float * self.V)(*self.tanx), ty=(C.c_float * self.V)(*self.tany))

    # -- descriptors ---------------------------------------------------------------
    def _views(self) -> _lib.SrViews:
        v = _lib.SrViews()
        v.V = self.V
        v.widths, v.heights = self._host["w"], self._host["h"]
        v.tanfovx, v.tanfovy = self._host["tx"], self._host["ty"]
        v.viewmatrices = self.viewmatrices.data_ptr()
        v.projmatrices = self.projmatrices.data_ptr()
        v.campos = self.campos.data_ptr()
        v.bg = self.bg.data_ptr()
        v.scale_modifier = self.scale_modifier
        v.sh_degree = self.sh_degree
        v.tile = self.config.tile
        v.pair_store = self.config.pair_store
        v.pair_capacity = self.pair_capacity
        v.slot_capacity = self.slot_capacity
        return v

    @staticmethod
    def _surfels(means3D, shs, colors, opacities, scales, rotations) -> _lib.SrSurfels:
        g = _lib.SrSurfels()
        g.P = means3D.shape[0]
        g.M = int(shs.shape[1]) if shs is not None else 0
        g.means3D = means3D.data_ptr()
        g.shs = shs.data_ptr() if shs is not None else None
        g.colors_precomp = colors.data_ptr() if colors is not None else None
        g.opacities = opacities.data_ptr()
        g.scales = scales.data_ptr()
        g.rotations = rotations.data_ptr()
        g.transMat_precomp = None
        return g

    def workspace_bytes(self, P: int, capacity: int, pair_capacity: int | None = None,
                        slot_capacity: int | None = None) -> int:
        saved = self.pair_capacity, self.slot_capacity
        if pair_capacity is not None:
            self.pair_capacity = int(pair_capacity)
        if slot_capacity is not None:
            self.slot_capacity = int(slot_capacity)
        try:
            return int(self.lib.sr_batch_workspace_bytes(C.byref(self._views()), P, capacity))
        finally:
            self.pair_capacity, self.slot_capacity = saved

    def _auto_pairs(self, capacity: int, pairs: int | None = None) -> int:
        """Pair-store size for a capacity: 0 (pair_store per instance) when
        the config fixes pair_store, else the given pair count (the caller
        scales a measured count by its headroom), or 9 pairs per instance
        before any count."""
        if self.config.pair_store > 0:
            return 0
        n = pairs if pairs is not None else 9 * capacity
        return min(int(n) + 4096, 0x7FFFFFFF)

    def ensure_workspace(self, P: int, capacity: int | None = None, pair_capacity: int | None = None,
                         slot_capacity: int | None = None) -> None:
        """Sizes the workspace for P surfels, a capacity of (surfel, tile)
        instances, an accepted-pair store and partial-gradient slots (0: one
        per instance and quadrant).  With no capacity known yet it starts at
        1: the first forward then stops after the tile histogram and reports
        the exact count, which the next call allocates for."""
        cap = int(capacity or self.capacity or 1)
        if pair_capacity is not None:
            pcap = int(pair_capacity)
        elif cap == self.capacity and self.pair_capacity > 0:
            pcap = self.pair_capacity  # keep a store sized from a measured count
        else:
            pcap = self._auto_pairs(cap)
        if slot_capacity is not None:
            scap = int(slot_capacity)
        elif cap == self.capacity:
            scap = self.slot_capacity
        else:
            scap = 0
        need = self.workspace_bytes(P, cap, pcap, scap)
        if need == 0:
            raise RuntimeError("invalid batch description")
        # reallocated when too small, and when over a quarter too large (the
        # sizes measured after a first forward are usually far below its
        # guesses); a captured graph keeps the workspace it was recorded on
        if self.workspace is None or not need <= self.workspace.numel() <= need + need // 4:
            self.workspace = None
            self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        self.capacity = cap
        self.pair_capacity = pcap
        self.slot_capacity = scap
        self._P = P

    # -- raw passes (graph-capturable) ---------------------------------------------
    def forward_raw(self, means3D, shs, colors, opacities, scales, rotations, out_color, out_others, radii):
        g = self._surfels(means3D, shs, colors, opacities, scales, rotations)
        _lib.check(self.lib.sr_forward_batched(
            C.byref(self._views()), C.byref(g), self.workspace.data_ptr(), self.workspace.numel(), self.capacity,
            out_color.data_ptr(), out_others.data_ptr(), radii.data_ptr(), self.status.data_ptr(),
            _lib.stream_handle(self.device)))

    def backward_raw(self, means3D, shs, colors, opacities, scales, rotations, radii, dL_dcolor, dL_dothers,
                     grads: dict):
        g = self._surfels(means3D, shs, colors, opacities, scales, rotations)
        gr = _lib.SrGrads()
        gr.dL_dmeans2D = grads["means2D"].data_ptr()
        gr.dL_dcolors = grads["colors"].data_ptr()
        gr.dL_dopacity = grads["opacities"].data_ptr()
        gr.dL_dmeans3D = grads["means3D"].data_ptr()
        gr.dL_dtransMat = None
        gr.dL_dsh = grads["shs"].data_ptr() if grads["shs"].numel() else None
        gr.dL_dscales = grads["scales"].data_ptr()
        gr.dL_drotations = grads["rotations"].data_ptr()
        _lib.check(self.lib.sr_backward_batched(
            C.byref(self._views()), C.byref(g), self.workspace.data_ptr(), self.workspace.numel(), self.capacity,
            radii.data_ptr(), dL_dcolor.data_ptr(), dL_dothers.data_ptr(), C.byref(gr),
            _lib.stream_handle(self.device)))

    def new_outputs(self, P: int):
        f32 = dict(dtype=torch.float32, device=self.device)
        return (torch.empty(3 * self.N, **f32), torch.empty(7 * self.N, **f32),
                torch.empty((self.V, P), dtype=torch.int32, device=self.device))

    def new_grads(self, P: int, M: int) -> dict:
        f32 = dict(dtype=torch.float32, device=self.device)
        return {"means2D": torch.empty((self.V, P, 3), **f32), "colors": torch.empty((P, 3), **f32),
                "opacities": torch.empty((P, 1), **f32), "means3D": torch.empty((P, 3), **f32),
                "shs": torch.empty((P, M, 3), **f32), "scales": torch.empty((P, 2), **f32),
                "rotations": torch.empty((P, 4), **f32)}

    def read_status(self):
        """(instances, flags) of the last forward; synchronises the stream."""
        st = self.status.cpu()
        return int(st[0]) & 0xFFFFFFFF, int(st[1])

    def read_pairs(self) -> int:
        """Accepted pairs the last forward asked the pair store for."""
        st = self.status.cpu()
        return (int(st[4]) & 0xFFFFFFFF) | ((int(st[5]) & 0xFFFFFFFF) << 32)

    def read_slots(self) -> int:
        """Partial-gradient slots the last forward needed."""
        return int(self.status.cpu()[8]) & 0xFFFFFFFF

    def forward_checked(self, means3D, shs, colors, opacities, scales, rotations):
        """Forward that grows the workspace until the batch fits (host sync)."""
        P = means3D.shape[0]
        self.ensure_workspace(P)
        out = self.new_outputs(P)
        for _ in range(4):
            self.forward_raw(means3D, shs, colors, opacities, scales, rotations, *out)
            need, flags = self.read_status()
            if not flags & _lib.SR_STATUS_CAPACITY:
                return out
            if need <= self.capacity:  # the instances fit, the slots did not
                self.ensure_workspace(P, self.capacity, self.pair_capacity, int(self.read_slots() * 1.1) + 1024)
            else:  # the scan counted the slots too
                self.ensure_workspace(P, int(need * 1.1) + 1024, None, int(self.read_slots() * 1.1) + 1024)
        raise RuntimeError("batched rasteriser: capacity did not converge")

    # -- views of the flat outputs ---------------------------------------------------
    def split(self, flat: torch.Tensor, channels: int):
        """Per-view (channels, H, W) views of a flat batched image buffer."""
        return [flat[channels * b: channels * (b + n)].view(channels, h, w)
                for b, n, w, h in zip(self.pix_base, self.pixels, self.widths, self.heights)]

    def pack(self, images, channels: int) -> torch.Tensor:
        """Inverse of split: per-view (channels, H, W) tensors -> flat buffer."""
        return torch.cat([im.reshape(-1) for im in images]).contiguous()

    def __call__(self, means3D, opacities, scales, rotations, shs=None, colors_precomp=None, means2D=None):
        """Autograd forward over all views: (colors list, allmaps list, radii V x P)."""
        if (shs is None) == (colors_precomp is None):
            raise Exception("Please provide excatly one of either SHs or precomputed colors!")
        if means2D is None:
            means2D = torch.zeros((self.V,) + tuple(means3D.shape), dtype=means3D.dtype, device=means3D.device,
                                  requires_grad=True)
        e = torch.empty(0, device=means3D.device)
        color, others, radii = _RasterizeBatch.apply(means3D, means2D, shs if shs is not None else e,
                                                     colors_precomp if colors_precomp is not None else e,
                                                     opacities, scales, rotations, self)
        return self.split(color, 3), self.split(others, 7), radii


def _c(t):
    return None if t is None or t.numel() == 0 else t.contiguous().float()


class _RasterizeBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, means3D, means2D, shs, colors, opacities, scales, rotations, br: BatchRasterizer):
        args = (_c(means3D), _c(shs), _c(colors), _c(opacities), _c(scales), _c(rotations))
        color, others, radii = br.forward_checked(*args)
        ctx.br = br
        ctx.args = args
        ctx.capacity = (br.capacity, br.pair_capacity)
        ctx.save_for_backward(radii)
        ctx.mark_non_differentiable(radii)
        return color, others, radii

    @staticmethod
    def backward(ctx, g_color, g_others, g_radii):
        br = ctx.br
        (radii,) = ctx.saved_tensors
        means3D, shs, colors, opac, scales, rots = ctx.args
        P = means3D.shape[0]
        if (br.capacity, br.pair_capacity) != ctx.capacity:
            raise RuntimeError("the batch workspace was reallocated between forward and backward")
        M = shs.shape[1] if shs is not None else 0
        grads = br.new_grads(P, M)
        g_color = torch.zeros(3 * br.N, device=means3D.device) if g_color is None else g_color.contiguous()
        g_others = torch.zeros(7 * br.N, device=means3D.device) if g_others is None else g_others.contiguous()
        br.backward_raw(means3D, shs, colors, opac, scales, rots, radii, g_color, g_others, grads)
        return (grads["means3D"], grads["means2D"], grads["shs"] if shs is not None else None,
                grads["colors"] if colors is not None else None, grads["opacities"].view_as(opac),
                grads["scales"], grads["rotations"], None)


@dataclass
class GraphedStep:
    """A captured forward + backward over fixed buffers; replay() reruns it on
    whatever the input buffers hold.  Check ``ok()`` after a replay: a batch
    that outgrew the capacity leaves its outputs undefined."""

    br: BatchRasterizer
    inputs: tuple
    upstream: tuple
    outputs: tuple = field(default_factory=tuple)
    grads: dict = field(default_factory=dict)
    graph: torch.cuda.CUDAGraph | None = None

    def replay(self):
        self.graph.replay()

    def ok(self) -> bool:
        return not (self.br.read_status()[1] & _lib.SR_STATUS_CAPACITY)


def capture_step(br: BatchRasterizer, means3D, shs, colors, opacities, scales, rotations, dL_dcolor, dL_dothers,
                 headroom: float = 1.25) -> GraphedStep:
    """Captures forward + backward of the batch into one hipGraph.

    Runs once eagerly to size the capacity (with ``headroom``), then records
    the raw passes on a side stream.  Inputs and upstream gradients are read
    from the given buffers at every replay.
    """
    P = means3D.shape[0]
    br.forward_checked(means3D, shs, colors, opacities, scales, rotations)
    need, _ = br.read_status()
    cap = max(br.capacity, int(need * headroom) + 1024)
    pairs = br.read_pairs()
    br.ensure_workspace(P, cap, br._auto_pairs(cap, int(pairs * headroom)), int(br.read_slots() * headroom) + 1024)
    outs = br.new_outputs(P)
    M = shs.shape[1] if shs is not None else 0
    grads = br.new_grads(P, M)
    _lib.set_timing(False)  # no event records inside a capture
    s = torch.cuda.Stream(device=br.device)
    s.wait_stream(torch.cuda.current_stream(br.device))
    with torch.cuda.stream(s):  # warm-up on the capture stream
        br.forward_raw(means3D, shs, colors, opacities, scales, rotations, *outs)
        br.backward_raw(means3D, shs, colors, opacities, scales, rotations, outs[2], dL_dcolor, dL_dothers, grads)
    torch.cuda.current_stream(br.device).wait_stream(s)
    torch.cuda.synchronize(br.device)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        br.forward_raw(means3D, shs, colors, opacities, scales, rotations, *outs)
        br.backward_raw(means3D, shs, colors, opacities, scales, rotations, outs[2], dL_dcolor, dL_dothers, grads)
    return GraphedStep(br, (means3D, shs, colors, opacities, scales, rotations), (dL_dcolor, dL_dothers), outs,
                       grads, g)


# ---- view groups sharing one workspace ---------------------------------------------
class ViewGroups:
    """The views of a batch rendered in consecutive groups that share one
    workspace.  Each group's forward and backward run before the next group's
    forward, so the workspace holds one group's state at a time (the largest
    group's size instead of the whole batch's); per-surfel gradients are
    summed over the groups, screen-space statistics and images stay per view.
    Per view, every output equals the one-batch result; the per-surfel
    gradient sums are regrouped (one addition per group)."""

    def __init__(self, cameras, bg: torch.Tensor, sh_degree: int, groups: int = 2,
                 config: BatchConfig | None = None, scale_modifier: float = 1.0, device="cuda"):
        V = len(cameras)
        if not 1 <= groups <= V:
            raise ValueError(f"1..{V} groups for {V} views (got {groups})")
        bounds = [V * g // groups for g in range(groups + 1)]
        self.view_base = bounds[:-1]
        self.parts = [BatchRasterizer(cameras[a:b], bg, sh_degree, config, scale_modifier, device=device)
                      for a, b in zip(bounds[:-1], bounds[1:])]
        self.V = V
        self.device = self.parts[0].device
        self.pix_base = []
        n = 0
        for br in self.parts:
            self.pix_base.append(n)
            n += br.N
        self.N = n
        self.workspace = None

    def share_workspace(self, P: int, sizes) -> None:
        """One workspace for every group: sizes[g] = (capacity, pair capacity,
        slot capacity)."""
        need = max(br.workspace_bytes(P, c, pc, sc) for br, (c, pc, sc) in zip(self.parts, sizes))
        if need == 0:
            raise RuntimeError("invalid batch description")
        for br in self.parts:
            br.workspace = None
        if self.workspace is None or self.workspace.numel() < need:
            self.workspace = None
            self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        for br, (c, pc, sc) in zip(self.parts, sizes):
            br.workspace = self.workspace
            br.capacity = int(c)
            br.pair_capacity = int(pc)
            br.slot_capacity = int(sc)
            br._P = P

    def ok(self) -> bool:
        return all(not (br.read_status()[1] & _lib.SR_STATUS_CAPACITY) for br in self.parts)


@dataclass
class GroupedStep:
    """capture_grouped_step's graph and buffers: outputs hold every view
    (flat colour and aux planes in view order, radii V x P), grads the
    per-surfel sums over all views and dL_dmeans2D per view."""

    vg: ViewGroups
    outputs: tuple
    grads: dict
    graph: torch.cuda.CUDAGraph

    def replay(self):
        self.graph.replay()

    def ok(self) -> bool:
        return self.vg.ok()


def capture_grouped_step(vg: ViewGroups, means3D, shs, colors, opacities, scales, rotations, dL_dcolor,
                         dL_dothers, headroom: float = 1.25) -> GroupedStep:
    """capture_step over view groups: every group's forward and backward in
    one hipGraph, on one shared workspace.  dL_dcolor / dL_dothers are laid
    out like the whole batch's outputs (each group reads its own slice)."""
    P = means3D.shape[0]
    M = shs.shape[1] if shs is not None else 0
    sizes = []
    for br in vg.parts:  # size each group once, eagerly, one workspace at a time
        br.forward_checked(means3D, shs, colors, opacities, scales, rotations)
        need, _ = br.read_status()
        cap = max(br.capacity, int(need * headroom) + 1024)
        sizes.append((cap, br._auto_pairs(cap, int(br.read_pairs() * headroom)), int(br.read_slots() * headroom) + 1024))
        br.workspace = None
    vg.share_workspace(P, sizes)
    f32 = dict(dtype=torch.float32, device=vg.device)
    color = torch.empty(3 * vg.N, **f32)
    others = torch.empty(7 * vg.N, **f32)
    radii = torch.empty((vg.V, P), dtype=torch.int32, device=vg.device)
    total = {"means2D": torch.empty((vg.V, P, 3), **f32), "colors": torch.empty((P, 3), **f32),
             "opacities": torch.empty((P, 1), **f32), "means3D": torch.empty((P, 3), **f32),
             "shs": torch.empty((P, M, 3), **f32), "scales": torch.empty((P, 2), **f32),
             "rotations": torch.empty((P, 4), **f32)}
    summed = [k for k in total if k != "means2D"]
    extra = {k: torch.empty_like(total[k]) for k in summed}

    def run():
        for g, br in enumerate(vg.parts):
            a, n = vg.pix_base[g], br.N
            v0 = vg.view_base[g]
            outs = (color[3 * a: 3 * (a + n)], others[7 * a: 7 * (a + n)], radii[v0: v0 + br.V])
            br.forward_raw(means3D, shs, colors, opacities, scales, rotations, *outs)
            grads = dict(total if g == 0 else extra)
            grads["means2D"] = total["means2D"][v0: v0 + br.V]
            br.backward_raw(means3D, shs, colors, opacities, scales, rotations, outs[2],
                            dL_dcolor[3 * a: 3 * (a + n)], dL_dothers[7 * a: 7 * (a + n)], grads)
            if g > 0:
                for k in summed:
                    if total[k].numel():
                        total[k].add_(extra[k])

    _lib.set_timing(False)  # no event records inside a capture
    s = torch.cuda.Stream(device=vg.device)
    s.wait_stream(torch.cuda.current_stream(vg.device))
    with torch.cuda.stream(s):  # warm-up on the capture stream
        run()
    torch.cuda.current_stream(vg.device).wait_stream(s)
    torch.cuda.synchronize(vg.device)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        run()
    return GroupedStep(vg, (color, others, radii), total, graph)


# ---- autotuner ---------------------------------------------------------------------
_CACHE: dict = {}


def _key(br: BatchRasterizer, P: int, M: int) -> str:
    res = sorted(zip(br.widths, br.heights))
    return f"P{P}-M{M}-V{br.V}-" + ",".join(f"{w}x{h}" for w, h in res)


def autotune(cameras, bg, sh_degree, means3D, shs, colors, opacities, scales, rotations,
             tiles=(8, 16, 32), pair_stores=(0, 8, 48), reps: int = 3, cache_path: str | None = None,
             log=None) -> BatchConfig:
    """Picks the fastest (tile, pair_store) for this batch by timing forward +
    backward with HIP events; results are cached per (P, M, resolutions) in
    memory and, when cache_path is given, in a JSON file."""
    probe = BatchRasterizer(cameras, bg, sh_degree)
    P = means3D.shape[0]
    M = shs.shape[1] if shs is not None else 0
    key = _key(probe, P, M)
    if cache_path and os.path.exists(cache_path) and key not in _CACHE:
        with open(cache_path) as f:
            _CACHE.update(json.load(f))
    if key in _CACHE:
        return BatchConfig(**_CACHE[key]["config"])
    results = []
    for tile in tiles:
        for ps in pair_stores:
            cfg = BatchConfig(tile=tile, pair_store=ps)
            br = BatchRasterizer(cameras, bg, sh_degree, cfg)
            out = br.forward_checked(means3D, shs, colors, opacities, scales, rotations)
            grads = br.new_grads(P, M)
            gc = torch.ones(3 * br.N, device=br.device)
            go = torch.ones(7 * br.N, device=br.device)
            br.backward_raw(means3D, shs, colors, opacities, scales, rotations, out[2], gc, go, grads)
            torch.cuda.synchronize(br.device)
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(reps):
                br.forward_raw(means3D, shs, colors, opacities, scales, rotations, *out)
                br.backward_raw(means3D, shs, colors, opacities, scales, rotations, out[2], gc, go, grads)
            ev[1].record()
            torch.cuda.synchronize(br.device)
            ms = ev[0].elapsed_time(ev[1]) / reps
            results.append((ms, cfg))
            if log:
                log(f"[autotune] tile {tile} pair_store {ps}: {ms:.3f} ms")
            del br, out, grads
    ms, best = min(results, key=lambda r: r[0])
    _CACHE[key] = {"config": {"tile": best.tile, "pair_store": best.pair_store}, "ms": ms,
                   "all": [{"tile": c.tile, "pair_store": c.pair_store, "ms": m} for m, c in results]}
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(_CACHE, f, indent=1)
    return best
