"""On-disk formats of the reference, so its tooling reads what this package
writes and the other way round.

* The surfel PLY: one ``vertex`` element of float32 properties in the order of
  ``construct_list_of_attributes`` (scene/gaussian_model.py:280-292; the
  semantic variant games/block_mesh_splatting/scene/two_gaussian_model.py:
  257-271 inserts ``semantic_*`` between ``opacity`` and ``scale_*``), written
  as binary little-endian PLY 1.0, the layout plyfile (environment.yml:26,
  unpinned) writes for ``PlyData([PlyElement.describe(elements, 'vertex')])``
  (gaussian_model.py:294-310).  SH features are stored channel-major:
  ``f_dc_0..2`` then ``f_rest_{c * (M - 1) + k}`` for colour c and band
  coefficient k (the ``transpose(1, 2).flatten`` of gaussian_model.py:300-301).
  Reading follows ``_load_ply`` (gaussian_model.py:322-362,
  two_gaussian_model.py:293-330): properties found by prefix and sorted by
  their numeric suffix.
* Checkpoint tuples: ``capture_block`` / ``restore_block``
  (games/block_mesh_splatting/scene/block_gaussian_model.py:49-96, saved by
  train.py:142 as ``(tuple, iteration)`` in ``chkpnt<it>.pth``) and
  ``capture`` / ``restore`` of the semantic model (two_gaussian_model.py:
  201-235, ``part_chkpnt<it>.pth``, train.py:315).  The fields are given as a
  dict keyed by the reference's attribute names; the tuple is positional, as
  on disk.

The PLY reader accepts binary little/big-endian and ASCII files with scalar
properties of any PLY numeric type; list properties are rejected.
"""
from __future__ import annotations

import os

import numpy as np
import torch

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2", "ushort": "u2",
    "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4",
    "double": "f8", "float64": "f8",
}

# capture_block order, block_gaussian_model.py:50-69
BLOCK_CHECKPOINT_FIELDS = (
    "active_sh_degree", "_features_dc", "_features_rest", "_alpha", "_scale", "sq_r", "sq_s", "sq_t", "sq_occ",
    "sq_eps", "faces", "sq_eta", "sq_omega", "max_radii2D", "xyz_gradient_accum", "denom", "optimizer",
    "spatial_lr_scale",
)
# capture order of the semantic model, two_gaussian_model.py:202-216
PART_CHECKPOINT_FIELDS = (
    "active_sh_degree", "_xyz", "_features_dc", "_features_rest", "_scaling", "_rotation", "_opacity", "_semantic",
    "max_radii2D", "xyz_gradient_accum", "denom", "optimizer", "spatial_lr_scale",
)


def ply_attributes(n_dc: int = 3, n_rest: int = 45, n_scale: int = 2, n_rot: int = 4, n_semantic: int | None = None):
    """Property names in file order (construct_list_of_attributes);
    n_semantic None is the plain model, an int the semantic one."""
    names = ["x", "y", "z", "nx", "ny", "nz"]
    names += [f"f_dc_{i}" for i in range(n_dc)]
    names += [f"f_rest_{i}" for i in range(n_rest)]
    names.append("opacity")
    if n_semantic is not None:
        names += [f"semantic_{i}" for i in range(n_semantic)]
    names += [f"scale_{i}" for i in range(n_scale)]
    names += [f"rot_{i}" for i in range(n_rot)]
    return names


def _np(t) -> np.ndarray:
    if isinstance(t, torch.Tensor):
        t = t.detach().cpu().numpy()
    return np.asarray(t)


def write_ply(path: str, names, columns: np.ndarray) -> None:
    """Binary little-endian PLY with one vertex element of float32 properties."""
    columns = np.ascontiguousarray(columns, dtype="<f4")
    if columns.ndim != 2 or columns.shape[1] != len(names):
        raise ValueError("columns must be (N, len(names))")
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    header = ["ply", "format binary_little_endian 1.0", f"element vertex {columns.shape[0]}"]
    header += [f"property float {n}" for n in names]
    header.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(header) + "\n").encode("ascii"))
        f.write(columns.tobytes())


def read_ply(path: str) -> dict:
    """The vertex element of a PLY file as {property name: float64 array}."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{path}: not a PLY file")
        fmt, elements, cur = None, [], None
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: header without end_header")
            tok = line.decode("ascii").split()
            if not tok or tok[0] in ("comment", "obj_info"):
                continue
            if tok[0] == "format":
                fmt = tok[1]
            elif tok[0] == "element":
                cur = {"name": tok[1], "count": int(tok[2]), "props": []}
                elements.append(cur)
            elif tok[0] == "property":
                if tok[1] == "list":
                    raise ValueError(f"{path}: list properties are not supported")
                cur["props"].append((tok[2], _PLY_TYPES[tok[1]]))
            elif tok[0] == "end_header":
                break
        body = f.read()
    if fmt not in ("binary_little_endian", "binary_big_endian", "ascii"):
        raise ValueError(f"{path}: unknown format {fmt}")
    out, off = None, 0
    text = body.decode("ascii").split() if fmt == "ascii" else None
    for el in elements:
        if fmt == "ascii":
            n = el["count"] * len(el["props"])
            vals = np.array(text[off:off + n], dtype=np.float64).reshape(el["count"], len(el["props"]))
            off += n
            data = {name: vals[:, i] for i, (name, _) in enumerate(el["props"])}
        else:
            endian = "<" if fmt == "binary_little_endian" else ">"
            dt = np.dtype([(name, endian + t) for name, t in el["props"]])
            arr = np.frombuffer(body, dtype=dt, count=el["count"], offset=off)
            off += dt.itemsize * el["count"]
            data = {name: arr[name].astype(np.float64) for name, _ in el["props"]}
        if el["name"] == "vertex":
            out = data
            break
    if out is None:
        raise ValueError(f"{path}: no vertex element")
    return out


def save_surfel_ply(path: str, xyz, features_dc, features_rest, opacity, scaling, rotation, semantic=None) -> None:
    """_save_ply (gaussian_model.py:294-310; with semantic,
    two_gaussian_model.py:273-291).  features_dc is (P, 1, 3) and
    features_rest (P, M - 1, 3), as the models hold them; opacity is the
    stored logit (the reference writes inverse_sigmoid(get_opacity), which is
    the logit up to rounding); normals are written as zeros."""
    xyz = _np(xyz).reshape(-1, 3)
    P = xyz.shape[0]
    f_dc = np.transpose(_np(features_dc).reshape(P, -1, 3), (0, 2, 1)).reshape(P, -1)
    f_rest = np.transpose(_np(features_rest).reshape(P, -1, 3), (0, 2, 1)).reshape(P, -1)
    cols = [xyz, np.zeros_like(xyz), f_dc, f_rest, _np(opacity).reshape(P, 1)]
    n_sem = None
    if semantic is not None:
        sem = _np(semantic).reshape(P, -1)
        n_sem = sem.shape[1]
        cols.append(sem)
    scal = _np(scaling).reshape(P, -1)
    rot = _np(rotation).reshape(P, -1)
    cols += [scal, rot]
    names = ply_attributes(f_dc.shape[1], f_rest.shape[1], scal.shape[1], rot.shape[1], n_sem)
    write_ply(path, names, np.concatenate([c.astype(np.float32) for c in cols], axis=1))


def _by_prefix(v: dict, prefix: str) -> np.ndarray:
    names = sorted((n for n in v if n.startswith(prefix)), key=lambda n: int(n.split("_")[-1]))
    P = len(v["x"])
    return np.stack([v[n] for n in names], axis=1) if names else np.zeros((P, 0))


def load_surfel_ply(path: str, max_sh_degree: int = 3, device="cpu") -> dict:
    """_load_ply (gaussian_model.py:322-362; semantic channels as in
    two_gaussian_model.py:293-330).  Returns float32 tensors: _xyz (P, 3),
    _features_dc (P, 1, 3), _features_rest (P, M - 1, 3), _opacity (P, 1),
    _scaling, _rotation and _semantic (P, S; S = 0 without semantic_*)."""
    v = read_ply(path)
    P = len(v["x"])
    xyz = np.stack([v["x"], v["y"], v["z"]], axis=1)
    f_dc = np.stack([v["f_dc_0"], v["f_dc_1"], v["f_dc_2"]], axis=1).reshape(P, 3, 1)
    rest = _by_prefix(v, "f_rest_")
    M = (max_sh_degree + 1) ** 2
    if rest.shape[1] != 3 * M - 3:
        raise ValueError(f"{path}: {rest.shape[1]} f_rest properties, expected {3 * M - 3} for degree {max_sh_degree}")
    rest = rest.reshape(P, 3, M - 1)
    t = lambda a: torch.tensor(a, dtype=torch.float32, device=device)  # noqa: E731
    return {
        "_xyz": t(xyz),
        "_features_dc": t(f_dc).transpose(1, 2).contiguous(),
        "_features_rest": t(rest).transpose(1, 2).contiguous(),
        "_opacity": t(np.asarray(v["opacity"])[:, None]),
        "_scaling": t(_by_prefix(v, "scale_")),
        "_rotation": t(_by_prefix(v, "rot")),
        "_semantic": t(_by_prefix(v, "semantic_")),
    }


def _capture(fields: dict, order) -> tuple:
    missing = [k for k in order if k not in fields]
    if missing:
        raise KeyError(f"checkpoint fields missing: {missing}")
    return tuple(fields[k] for k in order)


def _restore(model_args, order) -> dict:
    if len(model_args) != len(order):
        raise ValueError(f"checkpoint tuple has {len(model_args)} fields, expected {len(order)}")
    return dict(zip(order, model_args))


def capture_block(fields: dict) -> tuple:
    """The 18-tuple of capture_block (block_gaussian_model.py:49-69)."""
    return _capture(fields, BLOCK_CHECKPOINT_FIELDS)


def restore_block(model_args) -> dict:
    """Fields of a capture_block tuple (block_gaussian_model.py:71-96)."""
    return _restore(model_args, BLOCK_CHECKPOINT_FIELDS)


def capture_part(fields: dict) -> tuple:
    """The 13-tuple of the semantic model's capture (two_gaussian_model.py:201-216)."""
    return _capture(fields, PART_CHECKPOINT_FIELDS)


def restore_part(model_args) -> dict:
    """Fields of a semantic-model capture tuple, by the capture order.  (The
    reference's restore, two_gaussian_model.py:218-231, unpacks positions 7
    and 8 as max_radii2D and _semantic, the reverse of its capture; the
    capture order is what the file holds.)"""
    return _restore(model_args, PART_CHECKPOINT_FIELDS)


def save_checkpoint(path: str, model_args: tuple, iteration: int) -> None:
    """torch.save((model_args, iteration), path), as train.py:142,315."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    torch.save((model_args, int(iteration)), path)


def load_checkpoint(path: str):
    """(model_args, iteration) from a checkpoint file, with the loader that
    executes nothing from the file (tensors, dicts, lists and numbers only)."""
    model_args, iteration = torch.load(path, weights_only=True, map_location="cpu")
    return tuple(model_args), int(iteration)

