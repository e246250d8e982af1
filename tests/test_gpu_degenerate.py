"""GPU parity of the FP64 preprocess for near-degenerate surfels: on scenes
built to be degenerate (grazing normals, footprints down to 1e-7 of a
pixel, radii at their ceil steps) the HIP kernel selects the same surfels
as the oracle (flip = +-2 in the record) and every preprocess output,
binning array, image and gradient meets the parity bars of
test_gpu_parity."""
import numpy as np
import pytest

from tests.helpers import gpu_forward, oracle_run, read_state
from tests.test_gpu_parity import _check_backward, _check_forward
from tests.test_oracle_degenerate import degenerate_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tile", [16, 8])
def test_degenerate_scene_matches_oracle(tile):
    scene = degenerate_scene(5, P=60000, W=400, H=300)
    o, rc = oracle_run(scene, tile=tile)
    assert rc == 0
    n_deg = int((np.abs(o.get("flip")) == 2).sum())
    print(f"[degenerate] tile {tile}: {n_deg} surfels in FP64, L={o.num_rendered}", flush=True)
    assert n_deg > 1000
    res = gpu_forward(scene, tile=tile)
    st = read_state(res, scene)
    vis = res["radii"].cpu().numpy() > 0
    np.testing.assert_array_equal(st["flip"][vis], o.get("flip")[vis])
    _check_forward(scene, res, o)
    _check_backward(scene, res, o, seed=12)
