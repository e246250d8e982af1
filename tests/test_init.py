"""The plain block initialisation (dataset_readers.py:487-503, superquadric.py:68-120,
block_gaussian_model.py:146-166), pinned by its defining properties: k-means
recovers well-separated clusters and is a fixed point of Lloyd's update,
the outlier filter keeps exactly the points under mean + 0.5 std of the
brute-force 20-NN mean distance, and the parameters take the reference's
constants."""
import math

import numpy as np
import torch

from partgs_amd.init import kmeans, knn_mean_distance, plain_block_init, remove_statistical_outlier


def _blobs(seed, k=8, n=300, spread=0.05):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-3, 3, size=(k, 3))
    while np.min([np.linalg.norm(a - b) for i, a in enumerate(centres) for b in centres[i + 1:]]) < 1.0:
        centres = rng.uniform(-3, 3, size=(k, 3))
    pts = np.concatenate([c + spread * rng.normal(size=(n, 3)) for c in centres])
    truth = np.repeat(np.arange(k), n)
    return torch.tensor(pts, dtype=torch.float32), truth, centres


def test_knn_mean_distance_brute_force():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(257, 3))
    got = knn_mean_distance(torch.tensor(x), 20, chunk=64).numpy()
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    d.sort(axis=1)
    want = d[:, 1:21].mean(axis=1)
    assert np.allclose(got, want, rtol=1e-10, atol=1e-12)


def test_statistical_outlier_filter():
    pts, _, _ = _blobs(2, k=2, n=400)
    far = torch.tensor([[20.0, 0, 0], [0, -25.0, 3.0]])
    allp = torch.cat([pts, far])
    kept, mask = remove_statistical_outlier(allp.double(), 20, 0.5)
    assert not mask[-1] and not mask[-2]
    x = allp.double().numpy()
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)
    d.sort(axis=1)
    md = d[:, 1:21].mean(axis=1)
    want = md <= md.mean() + 0.5 * md.std(ddof=1)
    assert np.array_equal(mask.numpy(), want)
    assert kept.shape[0] == want.sum()


def test_kmeans_recovers_separated_clusters():
    pts, truth, _ = _blobs(3)
    labels, c = kmeans(pts, 8, seed=0)
    labels = labels.numpy()
    # a bijection between found and true clusters
    for t in range(8):
        assert len(set(labels[truth == t])) == 1
    assert len(set(labels)) == 8
    # Lloyd fixed point: every centre is the mean of its points, every point is nearest its centre
    for j in range(8):
        assert torch.allclose(c[j], pts[torch.tensor(labels == j)].mean(0), atol=1e-5)
    assert np.array_equal(torch.cdist(pts, c).argmin(1).numpy(), labels)


def test_plain_init_constants_and_translations():
    pts, truth, centres = _blobs(4)
    params, occ = plain_block_init(pts, n_blocks=8, ratio_block_scene=0.25, seed=0, remove_outliers=False)
    assert params.shape == (8, 12) and occ.shape == (8,)
    assert torch.allclose(params[:, 0:3], torch.full((8, 3), math.log(0.25)))
    assert torch.equal(params[:, 3:7], torch.tensor([[1.0, 0, 0, 0]] * 8))
    assert torch.equal(params[:, 10:12], torch.zeros(8, 2))
    assert torch.allclose(torch.sigmoid(occ), torch.full((8,), 0.7))
    # translation = a cluster mean plus a jitter of at most half the scene extent per axis
    ext = (pts.max(0).values - pts.min(0).values).numpy()
    t = params[:, 7:10].numpy()
    off = np.abs(t[:, None, :] - centres[None, :, :]) / ext  # (block, true cluster, axis)
    best = off.max(axis=2).min(axis=1)
    assert (best <= 0.5 + 0.02).all()
