"""On-disk formats (scene/gaussian_model.py:280-362, two_gaussian_model.py:
201-330, block_gaussian_model.py:49-96, train.py:142,315): the PLY property
list and byte layout, the channel-major SH order, round trips, foreign PLY
encodings, and the positional checkpoint tuples."""
import struct

import numpy as np
import pytest
import torch

from partgs_amd import formats as F


def _model(P=7, deg=3, S=None, seed=0):
    g = torch.Generator().manual_seed(seed)
    M = (deg + 1) ** 2
    m = dict(xyz=torch.randn(P, 3, generator=g), features_dc=torch.randn(P, 1, 3, generator=g),
             features_rest=torch.randn(P, M - 1, 3, generator=g), opacity=torch.randn(P, 1, generator=g),
             scaling=torch.randn(P, 2, generator=g), rotation=torch.randn(P, 4, generator=g))
    if S is not None:
        m["semantic"] = torch.randn(P, S, generator=g)
    return m


def test_attribute_list_of_the_reference():
    names = F.ply_attributes()
    assert len(names) == 61
    assert names[:9] == ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    assert names[9] == "f_rest_0" and names[53] == "f_rest_44" and names[54] == "opacity"
    assert names[55:] == ["scale_0", "scale_1", "rot_0", "rot_1", "rot_2", "rot_3"]
    sem = F.ply_attributes(n_semantic=3)
    assert sem[54:60] == ["opacity", "semantic_0", "semantic_1", "semantic_2", "scale_0", "scale_1"]


def test_byte_layout_and_channel_major_sh(tmp_path):
    m = _model(P=3)
    path = str(tmp_path / "pc" / "point_cloud.ply")
    F.save_surfel_ply(path, **m)
    raw = open(path, "rb").read()
    head, body = raw.split(b"end_header\n", 1)
    lines = head.decode().splitlines()
    assert lines[:3] == ["ply", "format binary_little_endian 1.0", "element vertex 3"]
    assert lines[3:] == [f"property float {n}" for n in F.ply_attributes()]
    assert len(body) == 3 * 61 * 4
    rec = struct.unpack("<61f", body[61 * 4:2 * 61 * 4])  # second vertex
    fr = m["features_rest"][1]
    # f_rest_{c * 15 + k} = features_rest[k, c]: colour-major, as transpose(1, 2).flatten
    assert rec[9 + 0 * 15 + 4] == np.float32(fr[4, 0])
    assert rec[9 + 2 * 15 + 14] == np.float32(fr[14, 2])
    assert rec[6:9] == tuple(np.float32(m["features_dc"][1, 0]))
    assert rec[3:6] == (0.0, 0.0, 0.0)
    assert rec[54] == np.float32(m["opacity"][1, 0])


@pytest.mark.parametrize("S", [None, 0, 5])
def test_round_trip(tmp_path, S):
    m = _model(P=50, S=S)
    path = str(tmp_path / "point_cloud.ply")
    F.save_surfel_ply(path, **m)
    back = F.load_surfel_ply(path, max_sh_degree=3)
    for k, ref in (("_xyz", m["xyz"]), ("_features_dc", m["features_dc"]), ("_features_rest", m["features_rest"]),
                   ("_opacity", m["opacity"]), ("_scaling", m["scaling"]), ("_rotation", m["rotation"])):
        assert torch.equal(back[k], ref), k
    assert back["_semantic"].shape == (50, S or 0)
    if S:
        assert torch.equal(back["_semantic"], m["semantic"])
    with pytest.raises(ValueError, match="f_rest"):
        F.load_surfel_ply(path, max_sh_degree=2)


def test_foreign_encodings(tmp_path):
    """ASCII and big-endian files, a comment, a double and an extra element."""
    names = F.ply_attributes(n_rest=0)
    vals = np.arange(2 * len(names), dtype=np.float64).reshape(2, len(names)) * 0.5
    head = ["ply", "format ascii 1.0", "comment made by hand", "element vertex 2"]
    head += [f"property {'double' if n == 'x' else 'float'} {n}" for n in names]
    head += ["element face 0", "property uchar red", "end_header"]
    body = "\n".join(" ".join(repr(float(v)) for v in row) for row in vals)
    p1 = tmp_path / "a.ply"
    p1.write_text("\n".join(head) + "\n" + body + "\n")
    v = F.read_ply(str(p1))
    assert np.array_equal(v["x"], vals[:, 0]) and np.array_equal(v["rot_3"], vals[:, -1])
    big = ["ply", "format binary_big_endian 1.0", "element vertex 2"] + [f"property float {n}" for n in names]
    p2 = tmp_path / "b.ply"
    p2.write_bytes(("\n".join(big + ["end_header"]) + "\n").encode() + vals.astype(">f4").tobytes())
    back = F.load_surfel_ply(str(p2), max_sh_degree=0)
    assert np.array_equal(back["_rotation"].numpy(), vals[:, -4:].astype(np.float32))
    p3 = tmp_path / "c.ply"
    p3.write_text("ply\nformat ascii 1.0\nelement vertex 1\nproperty list uchar int idx\nend_header\n1 2\n")
    with pytest.raises(ValueError, match="list"):
        F.read_ply(str(p3))


def test_checkpoint_tuples(tmp_path):
    P, B = 12, 3
    params = [torch.nn.Parameter(torch.randn(B, 3)), torch.nn.Parameter(torch.randn(B, 4))]
    opt = torch.optim.Adam(params, lr=1e-3)
    (params[0].sum() + params[1].square().sum()).backward()
    opt.step()
    fields = {
        "active_sh_degree": 3, "_features_dc": torch.randn(P, 1, 3), "_features_rest": torch.randn(P, 15, 3),
        "_alpha": torch.rand(B, 4, 3), "_scale": torch.rand(B, 4, 1), "sq_r": params[1].detach(),
        "sq_s": params[0].detach(), "sq_t": torch.randn(B, 3), "sq_occ": torch.randn(B), "sq_eps": torch.zeros(B, 2),
        "faces": torch.randint(0, 10, (B, 4, 3)), "sq_eta": torch.rand(B, 10), "sq_omega": torch.rand(B, 10),
        "max_radii2D": torch.zeros(P), "xyz_gradient_accum": torch.zeros(P, 1), "denom": torch.zeros(P, 1),
        "optimizer": opt.state_dict(), "spatial_lr_scale": 2.5,
    }
    tup = F.capture_block(fields)
    assert len(tup) == 18 and tup[8] is fields["sq_occ"] and tup[16] is fields["optimizer"]
    path = str(tmp_path / "chkpnt7000.pth")
    F.save_checkpoint(path, tup, 7000)
    back, it = F.load_checkpoint(path)
    assert it == 7000
    r = F.restore_block(back)
    for k in F.BLOCK_CHECKPOINT_FIELDS:
        if isinstance(fields[k], torch.Tensor):
            assert torch.equal(r[k], fields[k]), k
    assert r["spatial_lr_scale"] == 2.5 and r["active_sh_degree"] == 3
    opt2 = torch.optim.Adam([torch.nn.Parameter(p.detach().clone()) for p in params], lr=1e-3)
    opt2.load_state_dict(r["optimizer"])
    assert torch.equal(opt2.state_dict()["state"][0]["exp_avg"], opt.state_dict()["state"][0]["exp_avg"])
    part = {k: fields.get(k, torch.full((2,), float(i))) for i, k in enumerate(F.PART_CHECKPOINT_FIELDS)}
    ptup = F.capture_part(part)
    assert len(ptup) == 13 and ptup[7] is part["_semantic"] and ptup[8] is part["max_radii2D"]
    assert F.restore_part(ptup)["_semantic"] is part["_semantic"]
    with pytest.raises(ValueError):
        F.restore_block(ptup)
    with pytest.raises(KeyError):
        F.capture_part({"_xyz": 1})
