"""Writes golden vectors from the reference's own Python functions.

Run once in a container that has the reference tree mounted at
/root/reference; the vectors are committed as tests/golden/*.npz and the
tests read only those files (the reference does not travel).

Functions covered, and why a vector is needed beyond the mathematical checks:

* utils/sh_utils.py:eval_sh -- SH basis orthonormality pins the magnitudes of
  the band constants but not the sign convention of each basis function.
* utils/general_utils.py:rot_to_quat_batch -- the quaternion round trip pins
  well-conditioned matrices; the branch choice at ties of the four candidate
  magnitudes, the 0.1 floor and the w >= 0 standardisation at w == 0 need the
  reference's own outputs.
"""
import importlib.util
import math
import os

import numpy as np
import torch

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def _load(name, rel):
    spec = importlib.util.spec_from_file_location(name, os.path.join(REF, rel))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def sh_vectors():
    sh_utils = _load("ref_sh_utils", "utils/sh_utils.py")
    g = torch.Generator().manual_seed(0)
    n = 64
    dirs = torch.randn(n, 3, generator=g, dtype=torch.float64)
    dirs = dirs / dirs.norm(dim=1, keepdim=True)
    # axis-aligned and diagonal directions as well
    extra = torch.tensor([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
                          [1, 1, 1], [-1, 1, -1]], dtype=torch.float64)
    dirs = torch.cat([dirs, extra / extra.norm(dim=1, keepdim=True)])
    coeffs = torch.randn(dirs.shape[0], 3, 16, generator=g, dtype=torch.float64)
    out = {"dirs": dirs.numpy(), "coeffs": coeffs.numpy()}
    for deg in range(4):
        out[f"deg{deg}"] = sh_utils.eval_sh(deg, coeffs, dirs).numpy()
    np.savez(os.path.join(OUT, "eval_sh.npz"), **out)


def quat_vectors():
    gu = _load("ref_general_utils", "utils/general_utils.py")
    g = torch.Generator().manual_seed(1)
    q = torch.randn(256, 4, generator=g, dtype=torch.float32)
    q = q / q.norm(dim=1, keepdim=True)
    w, x, y, z = q.unbind(1)
    R = torch.stack([
        torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1),
    ], 1)
    special = []
    # rotations by pi about axes and diagonals (w = 0, trace = -1), identity,
    # and matrices with tied candidate magnitudes
    for ax in ([1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]):
        a = torch.tensor(ax, dtype=torch.float32)
        a = a / a.norm()
        for ang in (math.pi, math.pi / 2, 2 * math.pi / 3, 0.0):
            K = torch.tensor([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]], dtype=torch.float32)
            special.append(torch.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * (K @ K))
    Rall = torch.cat([R, torch.stack(special)])
    quat = gu.rot_to_quat_batch(Rall)
    np.savez(os.path.join(OUT, "rot_to_quat.npz"), R=Rall.numpy(), quat=quat.numpy())


if __name__ == "__main__":
    sh_vectors()
    quat_vectors()
    print("golden vectors written to", OUT)
