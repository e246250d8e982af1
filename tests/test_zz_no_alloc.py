"""sb_generate_bwd_noalloc (the block-layer adjoint on the library's fixed
scratch, no allocation) against sb_generate_bwd (the same adjoint on a
caller workspace), through the C ABI: bit-identical gradients on two calls
in a row, the second on fewer blocks (so stale scratch contents from the
first call would show), and a clear error beyond the scratch's size."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import block_oracle as BO
from partgs_amd import _lib
from partgs_amd.blocks import _desc, make_grid

pytestmark = pytest.mark.gpu


def _inputs(seed, B, n_eta, n_omega):
    rng = np.random.default_rng(seed)
    params, occ = BO.random_blocks(rng, B)
    grid = make_grid(n_eta, n_omega)
    P = B * n_eta * n_omega
    dev = torch.device("cuda")
    f = lambda a: torch.tensor(a, dtype=torch.float32, device=dev).contiguous()  # noqa: E731
    p, o = f(params), f(occ)
    ups = [f(rng.normal(size=(P, k))) for k in (3, 2, 4, 1)]
    return grid, p, o, ups


def _with_workspace(lib, d, ups, B):
    dp = torch.empty((B, 12), dtype=torch.float32, device="cuda")
    do = torch.empty((B,), dtype=torch.float32, device="cuda")
    ws = torch.empty(int(lib.sb_bwd_workspace_bytes(C.byref(d))), dtype=torch.uint8, device="cuda")
    _lib.check(lib.sb_generate_bwd(C.byref(d), *(u.data_ptr() for u in ups), dp.data_ptr(), do.data_ptr(),
                                   ws.data_ptr(), ws.numel(), _lib.stream_handle(torch.device("cuda"))))
    return dp, do


def _no_alloc(lib, d, ups, B):
    dp = torch.full((B, 12), float("nan"), dtype=torch.float32, device="cuda")
    do = torch.full((B,), float("nan"), dtype=torch.float32, device="cuda")
    _lib.check(lib.sb_generate_bwd_noalloc(C.byref(d), *(u.data_ptr() for u in ups), dp.data_ptr(),
                                           do.data_ptr(), _lib.stream_handle(torch.device("cuda"))))
    return dp, do


def test_no_alloc_block_backward_matches_workspace_variant():
    lib = _lib.load()
    # 64 blocks x 64 x 64 = 262144 surfels (the scratch's whole size), then
    # 8 blocks x 24 x 40
    for seed, (B, ne, no) in enumerate(((64, 64, 64), (8, 24, 40))):
        print(f"[no_alloc] call {seed}: B={B} grid {ne}x{no}", flush=True)
        grid, p, o, ups = _inputs(seed, B, ne, no)
        d = _desc(grid, p, o)
        want = _with_workspace(lib, d, ups, B)
        got = _no_alloc(lib, d, ups, B)
        torch.cuda.synchronize()
        for g, w, name in zip(got, want, ("params", "occupancy")):
            assert torch.isfinite(g).all(), name
            assert torch.equal(g, w), name
    # one block more than the scratch holds: a clear error, nothing launched
    grid, p, o, ups = _inputs(3, 1, 64, 64)
    d = _desc(grid, p, o)
    d.B = 65
    rc = lib.sb_generate_bwd_noalloc(C.byref(d), *(u.data_ptr() for u in ups), p.data_ptr(), o.data_ptr(),
                                     _lib.stream_handle(torch.device("cuda")))
    assert rc == _lib.SR_ERR_INVALID
    assert "262144" in lib.sr_last_error().decode()
