"""ctypes binding of ``libsurfel_rast.so`` (the C ABI in include/surfel_rast.h).

The library must have been built (``python -m partgs_amd.build``); there is no
fallback implementation: importing the rasteriser without it raises.
torch is imported first so that the HIP runtime torch already loaded is the
one the library binds to (both resolve the soname ``libamdhip64.so.7``).
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime before the library)

_HERE = os.path.dirname(os.path.abspath(__file__))
# PARTGS_AMD_LIB selects another build of the library (a variant built with
# other compile-time constants, for timing it against the default)
LIB_PATH = os.environ.get("PARTGS_AMD_LIB") or os.path.join(_HERE, "lib", "libsurfel_rast.so")

SR_OK = 0
SR_ERR_INVALID = 1
SR_ERR_HIP = 2
SR_ERR_ALLOC = 3
SR_ERR_PREFILTERED = 4
SR_ERR_CAPACITY = 5
SR_NUM_STAGES = 11
STAGE_NAMES = (
    "preprocess",
    "tile_scan",
    "scatter",
    "tile_sort",
    "blend_fwd",
    "grad_clear",
    "blend_bwd",
    "preprocess_bwd",
    "accept",
    "blend_bwd_grad",
    "pair_store",
)

EXPORTED = (
    "sr_forward",
    "sr_backward",
    "sr_mark_visible",
    "sr_geom_bytes",
    "sr_image_bytes",
    "sr_binning_bytes",
    "sr_state_offset",
    "sr_set_timing",
    "sr_set_timing_stages",
    "sr_get_timing",
    "sr_reset_timing",
    "sr_last_error",
    "sr_version",
    "sb_generate_fwd",
    "sb_generate_bwd",
    "sb_bwd_workspace_bytes",
    "sb_generate_bwd_noalloc",
    "sk_dist2",
    "sb_sdf_fwd",
    "sb_sdf_bwd",
    "sb_sdf_bwd_workspace_bytes",
    "sk_dist2_workspace_bytes",
    "sr_batch_workspace_bytes",
    "sr_forward_batched",
    "sr_backward_batched",
    "sr_pair_count",
    "sr_slot_count",
    "sr_forward_part",
    "sr_backward_part",
)

SR_MAX_VIEWS = 32
SR_STATUS_CAPACITY = 2


class SrSettings(C.Structure):
    _fields_ = [
        ("image_height", C.c_int),
        ("image_width", C.c_int),
        ("tanfovx", C.c_float),
        ("tanfovy", C.c_float),
        ("bg", C.c_void_p),
        ("scale_modifier", C.c_float),
        ("viewmatrix", C.c_void_p),
        ("projmatrix", C.c_void_p),
        ("sh_degree", C.c_int),
        ("campos", C.c_void_p),
        ("prefiltered", C.c_int),
        ("debug", C.c_int),
        ("tile", C.c_int),
        ("pair_store", C.c_int),
        ("capacity", C.c_int),
        ("pair_capacity", C.c_int),
        ("slot_capacity", C.c_int),
    ]


class SrSurfels(C.Structure):
    _fields_ = [
        ("P", C.c_int),
        ("M", C.c_int),
        ("means3D", C.c_void_p),
        ("shs", C.c_void_p),
        ("colors_precomp", C.c_void_p),
        ("opacities", C.c_void_p),
        ("scales", C.c_void_p),
        ("rotations", C.c_void_p),
        ("transMat_precomp", C.c_void_p),
    ]


class SrGrads(C.Structure):
    _fields_ = [
        ("dL_dmeans2D", C.c_void_p),
        ("dL_dcolors", C.c_void_p),
        ("dL_dopacity", C.c_void_p),
        ("dL_dmeans3D", C.c_void_p),
        ("dL_dtransMat", C.c_void_p),
        ("dL_dsh", C.c_void_p),
        ("dL_dscales", C.c_void_p),
        ("dL_drotations", C.c_void_p),
    ]


class SrViews(C.Structure):
    _fields_ = [
        ("V", C.c_int),
        ("widths", C.POINTER(C.c_int)),
        ("heights", C.POINTER(C.c_int)),
        ("tanfovx", C.POINTER(C.c_float)),
        ("tanfovy", C.POINTER(C.c_float)),
        ("viewmatrices", C.c_void_p),
        ("projmatrices", C.c_void_p),
        ("campos", C.c_void_p),
        ("bg", C.c_void_p),
        ("scale_modifier", C.c_float),
        ("sh_degree", C.c_int),
        ("tile", C.c_int),
        ("pair_store", C.c_int),
        ("pair_capacity", C.c_int),
        ("slot_capacity", C.c_int),
    ]


ALLOC_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t, C.c_int)

_lib = None
_lock = threading.Lock()


def load(path: str = LIB_PATH) -> C.CDLL:
    """Loads the library once; raises when it is missing or incomplete."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"partgs_amd: HIP library not found at {path}; build it with "
                "`python -m partgs_amd.build` (there is no CPU fallback)"
            )
        lib = C.CDLL(path)
        missing = [n for n in EXPORTED if not hasattr(lib, n)]
        if missing:
            raise ImportError(f"partgs_amd: {path} lacks symbols {missing}")
        vp, i, f, sz = C.c_void_p, C.c_int, C.c_float, C.c_size_t
        lib.sr_forward.argtypes = [C.POINTER(SrSettings), C.POINTER(SrSurfels), ALLOC_FN, vp, vp, vp, vp,
                                   C.POINTER(C.c_int), vp]
        lib.sr_forward.restype = i
        lib.sr_backward.argtypes = [C.POINTER(SrSettings), C.POINTER(SrSurfels), i, vp, vp, vp, vp, vp, vp,
                                    C.POINTER(SrGrads), vp]
        lib.sr_backward.restype = i
        lib.sr_mark_visible.argtypes = [i, vp, vp, vp, vp, vp]
        lib.sr_mark_visible.restype = i
        lib.sr_geom_bytes.argtypes = [i]
        lib.sr_geom_bytes.restype = sz
        lib.sr_image_bytes.argtypes = [i, i, i]
        lib.sr_image_bytes.restype = sz
        lib.sr_binning_bytes.argtypes = [i, i, i, i, i, i, i]
        lib.sr_binning_bytes.restype = sz
        lib.sr_state_offset.argtypes = [C.c_char_p, i, i, i, i, i]
        lib.sr_state_offset.restype = sz
        lib.sr_set_timing.argtypes = [i]
        lib.sr_set_timing.restype = i
        lib.sr_set_timing_stages.argtypes = [C.c_uint]
        lib.sr_set_timing_stages.restype = i
        lib.sr_get_timing.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int), i]
        lib.sr_get_timing.restype = i
        lib.sr_reset_timing.argtypes = []
        lib.sr_reset_timing.restype = None
        lib.sr_last_error.argtypes = []
        lib.sr_last_error.restype = C.c_char_p
        lib.sb_generate_fwd.argtypes = [vp, vp, vp, vp, vp, vp]
        lib.sb_generate_fwd.restype = i
        lib.sb_generate_bwd.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, sz, vp]
        lib.sb_generate_bwd_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
        lib.sb_generate_bwd_noalloc.restype = i
        lib.sb_bwd_workspace_bytes.argtypes = [vp]
        lib.sb_bwd_workspace_bytes.restype = sz
        lib.sb_generate_bwd.restype = i
        lib.sr_batch_workspace_bytes.argtypes = [C.POINTER(SrViews), i, i]
        lib.sr_batch_workspace_bytes.restype = sz
        lib.sr_forward_batched.argtypes = [C.POINTER(SrViews), C.POINTER(SrSurfels), vp, sz, i, vp, vp, vp, vp, vp]
        lib.sr_forward_batched.restype = i
        lib.sr_backward_batched.argtypes = [C.POINTER(SrViews), C.POINTER(SrSurfels), vp, sz, i, vp, vp, vp,
                                            C.POINTER(SrGrads), vp]
        lib.sr_backward_batched.restype = i
        lib.sr_version.argtypes = []
        lib.sr_version.restype = C.c_char_p
        lib.sr_pair_count.argtypes = [C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint)]
        lib.sr_slot_count.argtypes = [C.POINTER(C.c_uint), C.POINTER(C.c_uint)]
        lib.sr_slot_count.restype = i
        lib.sr_forward_part.argtypes = [C.POINTER(SrSettings), C.POINTER(SrSurfels), i, vp, ALLOC_FN, vp, vp, vp,
                                        vp, vp, C.POINTER(C.c_int), vp]
        lib.sr_forward_part.restype = i
        lib.sr_backward_part.argtypes = [C.POINTER(SrSettings), C.POINTER(SrSurfels), i, vp, i, vp, vp, vp, vp, vp,
                                         vp, vp, C.POINTER(SrGrads), vp, vp]
        lib.sr_backward_part.restype = i
        lib.sr_pair_count.restype = i
        lib.sb_sdf_fwd.argtypes = [vp, i, f, f, vp, i, vp, vp]
        lib.sb_sdf_fwd.restype = i
        lib.sb_sdf_bwd_workspace_bytes.argtypes = [i, i]
        lib.sb_sdf_bwd_workspace_bytes.restype = sz
        lib.sb_sdf_bwd.argtypes = [vp, i, f, f, vp, i, vp, vp, vp, sz, vp]
        lib.sb_sdf_bwd.restype = i
        lib.sk_dist2_workspace_bytes.argtypes = [i]
        lib.sk_dist2_workspace_bytes.restype = sz
        lib.sk_dist2.argtypes = [i, vp, vp, vp, sz, vp]
        lib.sk_dist2.restype = i
        _lib = lib
        return lib


class CapacityError(RuntimeError):
    """More tile instances than the binning capacity the caller gave, or,
    with the instances within it (slots True), more partial-gradient slots
    than its slot capacity."""

    def __init__(self, needed: int, capacity: int, slots: bool = False):
        if slots:
            super().__init__(f"the partial-gradient slots exceed the slot capacity ({needed} tile instances)")
        else:
            super().__init__(f"{needed} tile instances exceed the binning capacity {capacity}")
        self.needed = needed
        self.capacity = capacity
        self.slots = slots


def check(rc: int) -> None:
    """Raises RuntimeError carrying the library's message for a non-zero code."""
    if rc != SR_OK:
        msg = load().sr_last_error().decode(errors="replace")
        raise RuntimeError(f"surfel_rast error {rc}: {msg}")


def ptr(t) -> int | None:
    """Device pointer of a tensor, or None (NULL) for None / empty tensors."""
    if t is None or t.numel() == 0:
        return None
    return t.data_ptr()


def stream_handle(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def state_offset(name: str, P: int, W: int, H: int, tile: int, L: int) -> int:
    off = load().sr_state_offset(name.encode(), P, W, H, tile, L)
    if off == (1 << 64) - 1:
        raise KeyError(name)
    return off


def pair_count():
    """(accepted pairs, tile instances) of the latest completed forward of
    this thread on the current device, or None before one completed (never
    waits)."""
    n = C.c_ulonglong(0)
    L = C.c_uint(0)
    if load().sr_pair_count(C.byref(n), C.byref(L)) != SR_OK:
        return None
    return int(n.value), int(L.value)


def slot_count():
    """(partial-gradient slots, tile instances) of the same forward as
    pair_count, or None before one completed (never waits)."""
    n = C.c_uint(0)
    L = C.c_uint(0)
    if load().sr_slot_count(C.byref(n), C.byref(L)) != SR_OK:
        return None
    return int(n.value), int(L.value)


def set_timing(enabled: bool) -> None:
    check(load().sr_set_timing(1 if enabled else 0))


def set_timing_stages(names) -> None:
    """Time only the named stages (names from STAGE_NAMES)."""
    mask = 0
    for n in names:
        mask |= 1 << STAGE_NAMES.index(n)
    check(load().sr_set_timing_stages(mask))


def get_timing() -> dict:
    """Per-stage {name: (total_ms, launches)} accumulated since the last reset."""
    ms = (C.c_double * SR_NUM_STAGES)()
    cnt = (C.c_int * SR_NUM_STAGES)()
    check(load().sr_get_timing(ms, cnt, SR_NUM_STAGES))
    return {STAGE_NAMES[k]: (ms[k], cnt[k]) for k in range(SR_NUM_STAGES)}


def reset_timing() -> None:
    load().sr_reset_timing()
