"""distCUDA2 (submodules/simple-knn/spatial.cu:16-27, simple_knn.cu:185-220).

CPU: the oracle pinned by closed forms (a regular grid: spacing squared for
interior points; duplicates: 0; P = 1, 2, 3 with the reference's FLT_MAX
padding) and by brute force against the KD-tree path.  GPU: the HIP kernel
bit for bit against the oracle, on uniform, clustered, planar (the origin
inside the reference's bounding box) and duplicated clouds up to 1,000,000
points, and run to run.
"""
import numpy as np
import pytest
import torch

from oracle import knn_oracle as K


def test_grid_closed_form():
    h = np.float32(0.125)
    g = np.stack(np.meshgrid(*[np.arange(6)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32) * h
    out = K.dist2_brute(g)
    interior = np.all((g > 0) & (g < 5 * h), axis=1)
    assert np.all(out[interior] == h * h)
    # a corner has 3 neighbours at h
    corner = np.all(g == 0, axis=1)
    assert out[corner][0] == h * h


def test_small_counts_and_duplicates():
    assert K.dist2_brute(np.zeros((1, 3))).tolist() == [np.inf]
    two = K.dist2_brute(np.array([[0, 0, 0], [1, 0, 0]], np.float32))
    assert np.all(np.isinf(two))  # (1 + FLT_MAX) + FLT_MAX overflows
    three = K.dist2_brute(np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0]], np.float32))
    want = np.float32((np.float32(np.float32(1.0) + np.float32(4.0)) + K.FLT_MAX) / np.float32(3.0))
    assert three[0] == want
    dup = np.array([[0.5, 0.5, 0.5]] * 4 + [[2, 2, 2]], np.float32)
    assert np.all(K.dist2_brute(dup)[:4] == 0.0)


@pytest.mark.parametrize("seed", [0, 1])
def test_kdtree_equals_brute_force(seed):
    rng = np.random.default_rng(seed)
    pts = np.concatenate([rng.normal(size=(1500, 3)), rng.normal(size=(500, 3)) * 0.01 + 3]).astype(np.float32)
    pts[:10] = pts[10:20]  # duplicates
    np.testing.assert_array_equal(K.dist2_kdtree(pts), K.dist2_brute(pts))


def _clouds():
    rng = np.random.default_rng(7)
    yield "uniform_5k", rng.uniform(-1, 1, (5000, 3))
    c = rng.normal(size=(20, 3)) * 2
    yield "clustered_40k", (c[rng.integers(0, 20, 40000)] + rng.normal(size=(40000, 3)) * 0.02)
    plane = rng.uniform(1, 3, (20000, 3))
    plane[:, 2] = 0.0
    yield "plane_20k", plane
    d = rng.uniform(-1, 1, (3000, 3))
    yield "duplicates_6k", np.concatenate([d, d])
    yield "tiny_3", rng.uniform(-1, 1, (3, 3))
    yield "tiny_1", rng.uniform(-1, 1, (1, 3))


@pytest.mark.gpu
@pytest.mark.parametrize("name,pts", list(_clouds()), ids=[n for n, _ in _clouds()])
def test_gpu_matches_oracle(name, pts):
    from simple_knn._C import distCUDA2

    pts = pts.astype(np.float32)
    print(f"[knn] {name} P={len(pts)}", flush=True)
    got = distCUDA2(torch.from_numpy(pts).cuda()).cpu().numpy()
    np.testing.assert_array_equal(got, K.dist2_kdtree(pts))


@pytest.mark.gpu
def test_gpu_one_million_and_determinism():
    from partgs_amd.knn import distCUDA2
    from partgs_amd.scenes import superquadric_surfels

    pts = superquadric_surfels(0, P=1_000_000).means3D.numpy()
    print("[knn] superquadric 1M", flush=True)
    t = torch.from_numpy(pts).cuda()
    a = distCUDA2(t)
    b = distCUDA2(t)
    assert torch.equal(a, b)
    np.testing.assert_array_equal(a.cpu().numpy(), K.dist2_kdtree(pts))
    assert distCUDA2(t[:0]).shape == (0,)


def test_c_abi_validation_without_gpu():
    import ctypes
    import os

    from partgs_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    assert lib.sk_dist2_workspace_bytes(1000) > 1000 * 20
    assert lib.sk_dist2(-1, None, None, None, 0, None) == _lib.SR_ERR_INVALID
    assert lib.sk_dist2(0, None, None, None, 0, None) == _lib.SR_OK
    assert lib.sk_dist2(10, 16, 16, 16, ctypes.c_size_t(8), None) == _lib.SR_ERR_INVALID
    assert b"workspace" in lib.sr_last_error()
