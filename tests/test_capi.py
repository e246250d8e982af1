"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import glob
import os
import re

import pytest
import torch

from partgs_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b((?:sr|sb|sp)_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("sr_forward", "sr_backward", "sr_mark_visible", "sr_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_layout_queries_without_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.sr_geom_bytes(1000) > 1000 * 80
    assert lib.sr_image_bytes(800, 600, 16) > 800 * 600 * 20
    assert lib.sr_binning_bytes(5000, 800, 600, 16, 0, 0, 0) > 5000 * 20
    # the pair store follows pair_capacity when one is given
    assert lib.sr_binning_bytes(5000, 800, 600, 16, 0, 1000, 0) < lib.sr_binning_bytes(5000, 800, 600, 16, 0, 0, 0)
    # and the partial-gradient slots follow slot_capacity (80 bytes each)
    full = lib.sr_binning_bytes(5000, 800, 600, 16, 0, 1000, 0)
    assert full - lib.sr_binning_bytes(5000, 800, 600, 16, 0, 1000, 6000) >= (4 * 5000 - 6000) * 80 - 256
    off = _lib.state_offset("point_list", 1000, 800, 600, 16, 5000)
    assert off % 256 == 0
    with pytest.raises(KeyError):
        _lib.state_offset("no_such_array", 1000, 800, 600, 16, 5000)
    assert lib.sr_version().decode().startswith("surfel_rast")


def test_invalid_arguments_are_reported_without_gpu():
    """Argument validation runs before any HIP call."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    s = _lib.SrSettings()
    s.image_height, s.image_width, s.tile = 48, 64, 12
    g = _lib.SrSurfels()
    g.P = 10
    nr = ctypes.c_int(0)
    fn = _lib.ALLOC_FN(lambda c, n, w: None)
    rc = lib.sr_forward(ctypes.byref(s), ctypes.byref(g), fn, None, 1, 1, 1, ctypes.byref(nr), None)
    assert rc == _lib.SR_ERR_INVALID
    assert b"tile" in lib.sr_last_error()
    s.tile = 16
    rc = lib.sr_forward(ctypes.byref(s), ctypes.byref(g), fn, None, 1, 1, 1, ctypes.byref(nr), None)
    assert rc == _lib.SR_ERR_INVALID
    assert b"SHs or precomputed colors" in lib.sr_last_error()


def test_part_arguments_are_reported_without_gpu():
    """The _part entry points reject more than 16 semantic channels and NULL
    semantics before any HIP call; the Python layer checks shapes first."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    s = _lib.SrSettings()
    s.image_height, s.image_width, s.tile = 48, 64, 16
    g = _lib.SrSurfels()
    g.P = 10
    g.means3D = g.colors_precomp = g.opacities = g.scales = g.rotations = 16  # never dereferenced
    s.bg = s.viewmatrix = s.projmatrix = s.campos = 16
    nr = ctypes.c_int(0)
    fn = _lib.ALLOC_FN(lambda c, n, w: None)
    rc = lib.sr_forward_part(ctypes.byref(s), ctypes.byref(g), 17, 16, fn, None, 1, 1, 1, 1, ctypes.byref(nr), None)
    assert rc == _lib.SR_ERR_INVALID
    assert b"semantic channels" in lib.sr_last_error()
    rc = lib.sr_forward_part(ctypes.byref(s), ctypes.byref(g), 4, None, fn, None, 1, 1, 1, 1, ctypes.byref(nr), None)
    assert rc == _lib.SR_ERR_INVALID
    assert b"semantics must not be NULL" in lib.sr_last_error()
    gr = _lib.SrGrads()
    rc = lib.sr_backward_part(ctypes.byref(s), ctypes.byref(g), 4, 16, 0, 16, 16, 16, 16, 16, None, 16,
                              ctypes.byref(gr), 16, None)
    assert rc == _lib.SR_ERR_INVALID
    from partgs_amd import rasterizer_part as RP
    with pytest.raises(RuntimeError, match="semantic types"):
        RP._semantics(torch.zeros(10, 17), 10, torch.device("cpu"))
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        RP._semantics(torch.zeros(10, 3), 10, torch.device("cpu"))
    assert RP._semantics(None, 10, torch.device("cpu"))[1] == 0


def _views(V=2, tile=16):
    import ctypes as C
    v = _lib.SrViews()
    v.V = V
    w = (C.c_int * max(V, 1))(*([64] * max(V, 1)))
    h = (C.c_int * max(V, 1))(*([48] * max(V, 1)))
    t = (C.c_float * max(V, 1))(*([0.5] * max(V, 1)))
    v.widths, v.heights, v.tanfovx, v.tanfovy = w, h, t, t
    v.viewmatrices = v.projmatrices = v.campos = v.bg = 16  # never dereferenced: validation fails first
    v.tile = tile
    v._keep = (w, h, t)
    return v


def test_batched_workspace_and_validation_without_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    v = _views()
    small = lib.sr_batch_workspace_bytes(ctypes.byref(v), 1000, 10)
    big = lib.sr_batch_workspace_bytes(ctypes.byref(v), 1000, 100000)
    assert 0 < small < big
    assert lib.sr_batch_workspace_bytes(ctypes.byref(_views(V=0)), 1000, 10) == 0
    assert lib.sr_batch_workspace_bytes(ctypes.byref(_views(V=_lib.SR_MAX_VIEWS + 1)), 1000, 10) == 0
    g = _lib.SrSurfels()
    g.P = 10
    g.means3D = g.opacities = g.scales = g.rotations = g.colors_precomp = 16
    g.transMat_precomp = 16
    rc = lib.sr_forward_batched(ctypes.byref(v), ctypes.byref(g), 16, big, 100000, 16, 16, 16, 16, None)
    assert rc == _lib.SR_ERR_INVALID and b"precomputed T" in lib.sr_last_error()
    g.transMat_precomp = None
    rc = lib.sr_forward_batched(ctypes.byref(_views(tile=12)), ctypes.byref(g), 16, big, 100000, 16, 16, 16, 16,
                                None)
    assert rc == _lib.SR_ERR_INVALID and b"tile" in lib.sr_last_error()
    rc = lib.sr_forward_batched(ctypes.byref(v), ctypes.byref(g), 16, small, 100000, 16, 16, 16, 16, None)
    assert rc == _lib.SR_ERR_INVALID and b"workspace" in lib.sr_last_error()
