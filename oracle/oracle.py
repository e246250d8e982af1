"""ctypes wrapper of the CPU oracle (oracle/libsr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  See
oracle/sr_oracle.cpp for what each stage restates and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsr_oracle.so")


class SroScene(C.Structure):
    _fields_ = [
        ("P", C.c_int), ("D", C.c_int), ("M", C.c_int), ("W", C.c_int), ("H", C.c_int), ("tile", C.c_int),
        ("prefiltered", C.c_int),
        ("means3D", C.c_void_p), ("shs", C.c_void_p), ("colors_precomp", C.c_void_p), ("opacities", C.c_void_p),
        ("scales", C.c_void_p), ("rotations", C.c_void_p), ("transMat_precomp", C.c_void_p),
        ("viewmatrix", C.c_void_p), ("projmatrix", C.c_void_p), ("campos", C.c_void_p), ("bg", C.c_void_p),
        ("scale_modifier", C.c_double), ("tan_fovx", C.c_double), ("tan_fovy", C.c_double),
    ]


def build() -> str:
    """Compiles the oracle with its Makefile (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def use_native_build() -> bool:
    """Compiles the oracle for the host it runs on (-march=native, the
    Makefile's flags otherwise) into a temporary directory and makes it the
    library this module calls; False (and the portable build stays) when no
    compiler is at hand.  For timing the CPU baseline on the machine that
    runs it; the shipped build targets any AVX2 host."""
    global _lib
    import shutil
    import tempfile

    cxx = shutil.which("g++")
    if not cxx:
        return False
    d = tempfile.mkdtemp(prefix="sr_oracle_native_")
    out = os.path.join(d, "libsr_oracle_native.so")
    cmd = [cxx, "-O3", "-std=c++17", "-march=native", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-fPIC",
           "-shared", "-o", out, os.path.join(HERE, "sr_oracle.cpp")]
    if subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL).returncode != 0:
        return False
    _lib = _bind(C.CDLL(out))
    return True


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def _bind(lib):
    """Argument and result types of the oracle's C functions."""
    lib.sro_create.restype = C.c_void_p
    lib.sro_destroy.argtypes = [C.c_void_p]
    lib.sro_forward.argtypes = [C.c_void_p, C.POINTER(SroScene), C.c_int]
    lib.sro_forward.restype = C.c_int
    lib.sro_num_rendered.argtypes = [C.c_void_p]
    lib.sro_num_rendered.restype = C.c_int
    lib.sro_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.sro_backward.restype = C.c_int
    lib.sro_get.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
    lib.sro_get.restype = C.c_int
    lib.sro_expf.argtypes = [C.c_float]
    lib.sro_expf.restype = C.c_float
    lib.sro_expf_array.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    lib.sro_sh_rgb.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                               C.c_void_p, C.c_int]
    lib.sro_higher_msb.argtypes = [C.c_uint32]
    lib.sro_higher_msb.restype = C.c_uint32
    lib.sro_set_threads.argtypes = [C.c_int]
    lib.sro_get_threads.restype = C.c_int
    lib.sro_mark_visible.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.sro_forward_record.argtypes = [C.c_void_p, C.POINTER(SroScene), C.c_int]
    lib.sro_forward_record.restype = C.c_int
    lib.sro_replay.argtypes = [C.c_void_p, C.POINTER(SroScene)]
    lib.sro_replay.restype = C.c_int
    lib.sro_set_fp64_degenerate.argtypes = [C.c_int]
    lib.sro_forward_part.argtypes = [C.c_void_p, C.POINTER(SroScene), C.c_int, C.c_int, C.c_void_p]
    lib.sro_forward_part.restype = C.c_int
    lib.sro_backward_part.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.sro_backward_part.restype = C.c_int
    lib.sro_alpha_capped.argtypes = [C.c_void_p]
    lib.sro_alpha_capped.restype = C.c_int
    return lib


def set_threads(n: int) -> None:
    load().sro_set_threads(int(n))


def get_threads() -> int:
    return int(load().sro_get_threads())


def expf(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32).ravel())
    y = np.empty_like(x)
    load().sro_expf_array(x.ctypes.data, y.ctypes.data, x.size)
    return y


def sh_rgb(deg: int, sh, means, campos, f64: bool = False):
    """Colour of n points from SH (n, M, 3) seen from campos; returns (rgb, clamped)."""
    dt = np.float64 if f64 else np.float32
    sh = np.ascontiguousarray(np.asarray(sh, dtype=dt))
    means = np.ascontiguousarray(np.asarray(means, dtype=dt))
    campos = np.ascontiguousarray(np.asarray(campos, dtype=dt))
    n, M = sh.shape[0], sh.shape[1]
    rgb = np.empty((n, 3), dtype=dt)
    cl = np.empty((n, 3), dtype=np.uint8)
    load().sro_sh_rgb(deg, M, sh.ctypes.data, means.ctypes.data, campos.ctypes.data, n, rgb.ctypes.data,
                      cl.ctypes.data, 1 if f64 else 0)
    return rgb, cl.astype(bool)


def set_fp64_degenerate(on: bool) -> None:
    """True (the default): near-degenerate surfels take the FP64 preprocess
    like the HIP kernel; False: all FP32, the reference's own arithmetic."""
    load().sro_set_fp64_degenerate(1 if on else 0)


def higher_msb(n: int) -> int:
    return int(load().sro_higher_msb(n))


# (name, dtype kind, shape as a function of sizes); 'R' = real of the run's precision
_SHAPES = {
    "radii": ("i4", lambda d: (d["P"],)),
    "depths": ("R", lambda d: (d["P"],)),
    "means2D": ("R", lambda d: (d["P"], 2)),
    "transMat": ("R", lambda d: (d["P"], 9)),
    "normal_opacity": ("R", lambda d: (d["P"], 4)),
    "rgb": ("R", lambda d: (d["P"], 3)),
    "clamped": ("u1", lambda d: (d["P"], 3)),
    "flip": ("i1", lambda d: (d["P"],)),
    "tiles_touched": ("u4", lambda d: (d["P"],)),
    "point_offsets": ("u4", lambda d: (d["P"],)),
    "keys": ("u8", lambda d: (d["L"],)),
    "point_list": ("u4", lambda d: (d["L"],)),
    "ranges": ("u4", lambda d: (d["T"], 2)),
    "out_color": ("R", lambda d: (3, d["H"], d["W"])),
    "out_others": ("R", lambda d: (d.get("O", 7), d["H"], d["W"])),
    "out_semantic": ("R", lambda d: (d.get("S", 0), d["H"], d["W"])),
    "dL_dsemantics": ("R", lambda d: (d["P"], d.get("S", 0))),
    "final_T": ("R", lambda d: (3, d["H"], d["W"])),
    "n_contrib": ("u4", lambda d: (2, d["H"], d["W"])),
    "acc_dL_dtransMat": ("f8", lambda d: (d["P"], 9)),
    "acc_dL_dmean2D": ("f8", lambda d: (d["P"], 3)),
    "acc_dL_dnormal": ("f8", lambda d: (d["P"], 3)),
    "acc_dL_dopacity": ("f8", lambda d: (d["P"], 1)),
    "acc_dL_dcolors": ("f8", lambda d: (d["P"], 3)),
    "dL_dmeans2D": ("R", lambda d: (d["P"], 3)),
    "dL_dcolors": ("R", lambda d: (d["P"], 3)),
    "dL_dopacity": ("R", lambda d: (d["P"], 1)),
    "dL_dmeans3D": ("R", lambda d: (d["P"], 3)),
    "dL_dtransMat": ("R", lambda d: (d["P"], 9)),
    "dL_dsh": ("R", lambda d: (d["P"], d["M"], 3)),
    "dL_dscales": ("R", lambda d: (d["P"], 2)),
    "dL_drotations": ("R", lambda d: (d["P"], 4)),
    "dL_dnormal": ("R", lambda d: (d["P"], 3)),
}


def _arr(a, dtype):
    if a is None:
        return None
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


class Oracle:
    """One oracle run: forward (and optionally backward) on host arrays."""

    def __init__(self):
        self.lib = load()
        self.h = self.lib.sro_create()
        self._keep = []
        self.dims = None
        self.f64 = False

    def __del__(self):
        try:
            self.lib.sro_destroy(self.h)
        except Exception:
            pass

    def forward(self, *, means3D, opacities, viewmatrix, projmatrix, campos, bg, W, H, tanfovx, tanfovy,
                shs=None, sh_degree=0, colors_precomp=None, scales=None, rotations=None, transMat_precomp=None,
                scale_modifier=1.0, tile=16, prefiltered=False, f64=False, record=False) -> int:
        """Preprocess, binning and blend.  record=True keeps every discrete
        decision for replay()."""
        s = self._scene(means3D=means3D, opacities=opacities, viewmatrix=viewmatrix, projmatrix=projmatrix,
                        campos=campos, bg=bg, W=W, H=H, tanfovx=tanfovx, tanfovy=tanfovy, shs=shs,
                        sh_degree=sh_degree, colors_precomp=colors_precomp, scales=scales, rotations=rotations,
                        transMat_precomp=transMat_precomp, scale_modifier=scale_modifier, tile=tile,
                        prefiltered=prefiltered, f64=f64)
        fn = self.lib.sro_forward_record if record else self.lib.sro_forward
        rc = fn(self.h, C.byref(s), 1 if f64 else 0)
        if rc < 0:
            raise ValueError("oracle: bad arguments")
        gx, gy = (W + tile - 1) // tile, (H + tile - 1) // tile
        self.f64 = f64
        self.dims = dict(P=s.P, M=s.M, W=int(W), H=int(H), T=gx * gy, L=self.lib.sro_num_rendered(self.h))
        return rc

    def forward_part(self, *, semantics, f64=False, **kw) -> int:
        """The _part variant (diff-surfel-rasterization_part): same keywords
        as forward plus semantics (P x S, S <= 16); 8 aux channels."""
        s = self._scene(f64=f64, **kw)
        sem = _arr(semantics, np.float64 if f64 else np.float32)
        self._keep.append(sem)
        S = int(sem.shape[1]) if sem.ndim == 2 else 0
        rc = self.lib.sro_forward_part(self.h, C.byref(s), 1 if f64 else 0, S,
                                       sem.ctypes.data if sem.size else None)
        if rc < 0:
            raise ValueError("oracle: bad arguments")
        W, H, tile = int(kw["W"]), int(kw["H"]), int(kw.get("tile", 16))
        gx, gy = (W + tile - 1) // tile, (H + tile - 1) // tile
        self.f64 = f64
        self.dims = dict(P=s.P, M=s.M, W=W, H=H, T=gx * gy, L=self.lib.sro_num_rendered(self.h), S=S, O=8)
        return rc

    def backward_part(self, dL_dcolor, dL_dsemantic, dL_dothers):
        dt = np.float64 if self.f64 else np.float32
        gc, gs, go = _arr(dL_dcolor, dt), _arr(dL_dsemantic, dt), _arr(dL_dothers, dt)
        self.lib.sro_backward_part(self.h, gc.ctypes.data, gs.ctypes.data if gs.size else None, go.ctypes.data)

    def replay(self, **kw) -> None:
        """Re-evaluates the recorded forward on new inputs (same keywords as
        forward) with its discrete decisions held: a smooth function of the
        inputs whose derivative is the gradient the backward computes."""
        kw.setdefault("f64", self.f64)
        s = self._scene(**kw)
        if self.lib.sro_replay(self.h, C.byref(s)) != 0:
            raise ValueError("oracle: replay needs a recorded forward of the same shape")

    @property
    def alpha_capped(self) -> int:
        """Blended pairs of the recorded forward whose alpha hit the 0.99 cap."""
        return int(self.lib.sro_alpha_capped(self.h))

    def _scene(self, *, means3D, opacities, viewmatrix, projmatrix, campos, bg, W, H, tanfovx, tanfovy,
               shs=None, sh_degree=0, colors_precomp=None, scales=None, rotations=None, transMat_precomp=None,
               scale_modifier=1.0, tile=16, prefiltered=False, f64=False):
        dt = np.float64 if f64 else np.float32
        arrs = {k: _arr(v, dt) for k, v in dict(
            means3D=means3D, shs=shs, colors_precomp=colors_precomp, opacities=opacities, scales=scales,
            rotations=rotations, transMat_precomp=transMat_precomp, viewmatrix=viewmatrix, projmatrix=projmatrix,
            campos=campos, bg=bg).items()}
        self._keep = list(arrs.values())
        P = int(arrs["means3D"].shape[0])
        M = int(arrs["shs"].shape[1]) if arrs["shs"] is not None and arrs["shs"].size else 0
        s = SroScene()
        s.P, s.D, s.M, s.W, s.H, s.tile, s.prefiltered = P, int(sh_degree), M, int(W), int(H), int(tile), int(
            bool(prefiltered))
        for k, v in arrs.items():
            setattr(s, k, None if v is None or v.size == 0 else v.ctypes.data)
        s.scale_modifier, s.tan_fovx, s.tan_fovy = float(scale_modifier), float(tanfovx), float(tanfovy)
        return s

    @property
    def num_rendered(self) -> int:
        return self.dims["L"]

    def backward(self, dL_dcolor, dL_dothers):
        dt = np.float64 if self.f64 else np.float32
        gc = _arr(dL_dcolor, dt)
        go = _arr(dL_dothers, dt)
        self.lib.sro_backward(self.h, gc.ctypes.data, go.ctypes.data)

    def get(self, name: str) -> np.ndarray:
        kind, shape = _SHAPES[name]
        if kind == "R":
            kind = "f8" if self.f64 else "f4"
        out = np.empty(shape(self.dims), dtype=np.dtype(kind))
        rc = self.lib.sro_get(self.h, name.encode(), out.ctypes.data if out.size else None, out.nbytes)
        if rc != 0:
            raise KeyError(f"oracle get({name}) failed: {rc}")
        return out


def run_scene(scene, *, f64=False, tile=16, grads=None, scale_modifier=1.0, use_colors=False):
    """Convenience: runs a partgs_amd.scenes.SurfelScene through the oracle."""
    cam = scene.camera
    o = Oracle()
    kw = dict(means3D=scene.means3D, opacities=scene.opacities, viewmatrix=cam.world_view_transform,
              projmatrix=cam.full_proj_transform, campos=cam.camera_center, bg=scene.bg, W=cam.image_width,
              H=cam.image_height, tanfovx=np.tan(cam.FoVx * 0.5), tanfovy=np.tan(cam.FoVy * 0.5),
              scales=scene.scales, rotations=scene.rotations, scale_modifier=scale_modifier, tile=tile, f64=f64)
    if use_colors:
        kw["colors_precomp"] = use_colors
    else:
        kw["shs"] = scene.shs
        kw["sh_degree"] = scene.sh_degree
    o.forward(**kw)
    if grads is not None:
        o.backward(*grads)
    return o


def mark_visible(means3D, viewmatrix) -> np.ndarray:
    m = _arr(means3D, np.float32)
    v = _arr(viewmatrix, np.float32)
    out = np.zeros(m.shape[0], dtype=np.uint8)
    load().sro_mark_visible(m.shape[0], m.ctypes.data, v.ctypes.data, out.ctypes.data)
    return out.astype(bool)
