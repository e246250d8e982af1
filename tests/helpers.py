"""Shared helpers for the parity tests (GPU library vs the CPU oracle)."""
from __future__ import annotations

import math

import numpy as np
import torch

from oracle import oracle as O
from partgs_amd import _lib
from partgs_amd.rasterizer import rasterize_gaussians_bwd, rasterize_gaussians_fwd

REC_QUADS = 5


def gpu_forward(scene, *, tile=16, scale_modifier=1.0, colors=None, transmat=None, prefiltered=False, debug=False,
                pair_store=0, capacity=0, pair_capacity=0, slot_capacity=0):
    """Runs the HIP forward on a SurfelScene; returns a dict of outputs and state."""
    dev = torch.device("cuda")
    cam = scene.camera
    empty = torch.empty(0, device=dev)
    shs = scene.shs.to(dev) if colors is None else empty
    col = colors.to(dev) if colors is not None else empty
    sc = scene.scales.to(dev) if transmat is None else empty
    rot = scene.rotations.to(dev) if transmat is None else empty
    tm = transmat.to(dev) if transmat is not None else empty
    args = (scene.bg.to(dev), scene.means3D.to(dev), col, scene.opacities.to(dev), sc, rot, scale_modifier, tm,
            cam.world_view_transform.to(dev), cam.full_proj_transform.to(dev), math.tan(cam.FoVx * 0.5),
            math.tan(cam.FoVy * 0.5), cam.image_height, cam.image_width, shs, scene.sh_degree,
            cam.camera_center.to(dev), prefiltered, debug, tile)
    L, color, others, radii, geom, binb, img = rasterize_gaussians_fwd(*args, pair_store=pair_store,
                                                                       capacity=capacity, pair_capacity=pair_capacity,
                                                                       slot_capacity=slot_capacity)
    torch.cuda.synchronize()
    out = dict(L=L, color=color, others=others, radii=radii, geom=geom, bin=binb, img=img, args=args, tile=tile,
               pair_store=pair_store, capacity=capacity, pair_capacity=pair_capacity, slot_capacity=slot_capacity)
    return out


def read_state(res, scene):
    """Decodes internal state arrays from the three state buffers."""
    P = scene.P
    W, H, B, L = scene.camera.image_width, scene.camera.image_height, res["tile"], res["L"]
    gx, gy = (W + B - 1) // B, (H + B - 1) // B
    T = gx * gy
    geom = res["geom"].cpu().numpy()
    img = res["img"].cpu().numpy()
    binb = res["bin"].cpu().numpy()

    def view(buf, name, dtype, count):  # the binning arrays are laid out for the capacity
        off = _lib.state_offset(name, P, W, H, B, res.get("capacity") or L)
        return np.frombuffer(buf, dtype=dtype, count=count, offset=off)

    st = {}
    if P > 0:
        rec = view(geom, "rec", np.float32, P * REC_QUADS * 4).reshape(P, REC_QUADS * 4)
        st["transMat"] = rec[:, 0:9]
        st["means2D"] = rec[:, 9:11]
        st["normal_opacity"] = np.concatenate([rec[:, 12:15], rec[:, 11:12]], axis=1)
        st["rgb"] = np.concatenate([rec[:, 15:16], rec[:, 16:18]], axis=1)
        st["flip"] = rec[:, 19].astype(np.int8)  # camera-facing sign, x2 for the FP64 preprocess
        st["depths"] = view(geom, "depths", np.float32, P)
        st["tiles_touched"] = view(geom, "tiles_touched", np.uint32, P)
        st["clamped"] = view(geom, "clamped", np.uint8, 4 * P).reshape(P, 4)[:, :3]
        st["ranges"] = view(img, "ranges", np.uint32, 2 * T).reshape(T, 2)
        st["final_T"] = view(img, "final_T", np.float32, 3 * W * H).reshape(3, H, W)
        st["n_contrib"] = view(img, "n_contrib", np.uint32, 2 * W * H).reshape(2, H, W)
        st["point_list"] = view(binb, "point_list", np.uint32, L) if L > 0 else np.zeros(0, np.uint32)
    return st


def gpu_backward(res, dL_dcolor, dL_dothers):
    a = res["args"]
    dev = torch.device("cuda")
    out = rasterize_gaussians_bwd(a[0], a[1], res["radii"], a[2], a[4], a[5], a[6], a[7], a[8], a[9], a[10], a[11],
                                  dL_dcolor.to(dev), dL_dothers.to(dev), a[14], a[15], a[16], res["geom"], res["L"],
                                  res["bin"], res["img"], a[19], res["tile"], res["pair_store"],
                                  res.get("capacity", 0), res.get("pair_capacity", 0))
    torch.cuda.synchronize()
    names = ("dL_dmeans2D", "dL_dcolors", "dL_dopacity", "dL_dmeans3D", "dL_dtransMat", "dL_dsh", "dL_dscales",
             "dL_drotations")
    return {n: t.cpu().numpy() for n, t in zip(names, out)}


def oracle_run(scene, *, tile=16, scale_modifier=1.0, colors=None, transmat=None, f64=False, prefiltered=False):
    cam = scene.camera
    o = O.Oracle()
    kw = dict(means3D=scene.means3D, opacities=scene.opacities, viewmatrix=cam.world_view_transform,
              projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=scene.bg, W=cam.image_width,
              H=cam.image_height, tanfovx=math.tan(cam.FoVx * 0.5), tanfovy=math.tan(cam.FoVy * 0.5),
              scale_modifier=scale_modifier, tile=tile, f64=f64, prefiltered=prefiltered)
    if colors is None:
        kw.update(shs=scene.shs, sh_degree=scene.sh_degree)
    else:
        kw.update(colors_precomp=colors)
    if transmat is None:
        kw.update(scales=scene.scales, rotations=scene.rotations)
    else:
        kw.update(transMat_precomp=transmat)
    rc = o.forward(**kw)
    return o, rc


def rel_err(a, b):
    """max |a - b| / max |b| (0 when both are zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max() if b.size else 0.0
    num = np.abs(a - b).max() if b.size else 0.0
    if den == 0.0:
        return num
    return num / den


def assert_close_scaled(a, b, tol, name=""):
    """Per-element |a - b| <= tol * (|b| + max|b|): tolerance relative to the
    element, with an absolute floor of tol * max|b| for the tensor."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    if b.size == 0:
        return
    floor = np.abs(b).max()
    bad = np.abs(a - b) > tol * (np.abs(b) + floor)
    assert not bad.any(), f"{name}: {bad.sum()} of {b.size} elements outside {tol} (max rel {rel_err(a, b):.3e})"
