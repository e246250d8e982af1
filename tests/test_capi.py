"""The C-ABI library loads without a GPU and exports every declared symbol."""
import ctypes
import glob
import os
import re

import pytest

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
    assert lib.sr_binning_bytes(5000, 800, 600, 16, 0) > 5000 * 20
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
