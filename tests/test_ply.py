"""3DGS .ply bridge (gray_amd/ply.py): property mapping of the reference
converter (convert/safetensors_ply_conversion.py), round trips, and files
with the extra properties 3DGS writes (normals)."""
import os

import numpy as np
import torch

from gray_amd import ply


def sample(n=37, k=15, seed=0):
    g = torch.Generator().manual_seed(seed)
    return {
        "mean": torch.randn(n, 3, generator=g), "rotation": torch.randn(n, 4, generator=g),
        "scale": torch.randn(n, 3, generator=g), "opacity": torch.randn(n, 1, generator=g),
        "sh_coeffs_dc": torch.randn(n, 1, 3, generator=g), "sh_coeffs_rest": torch.randn(n, k, 3, generator=g),
    }


def test_round_trip_and_channel_major_rest(tmp_path):
    t = sample()
    p = str(tmp_path / "g.ply")
    ply.tensors_to_ply(t, p)
    v = ply.read_vertices(p)
    assert v.dtype.names[:4] == ("x", "y", "z", "opacity") and len(v) == 37
    # f_rest_j = channel * K + k (3DGS layout)
    assert np.array_equal(v["f_rest_16"], t["sh_coeffs_rest"][:, 1, 1].numpy())
    assert np.array_equal(v["f_rest_44"], t["sh_coeffs_rest"][:, 14, 2].numpy())
    back = ply.ply_to_tensors(p)
    for key, val in t.items():
        assert torch.equal(back[key], val), key
    assert int(back["current_sh_degree"]) == 3


def test_safetensors_files_round_trip(tmp_path):
    import safetensors.torch

    t = sample(k=8)
    s1, p, s2 = (str(tmp_path / n) for n in ("a.safetensors", "b.ply", "c.safetensors"))
    safetensors.torch.save_file(t, s1)
    ply.safetensors_to_ply(s1, p)
    ply.ply_to_safetensors(p, s2)
    back = safetensors.torch.load_file(s2)
    for key, val in t.items():
        assert torch.equal(back[key], val), key
    assert int(back["current_sh_degree"]) == 2


def test_reads_3dgs_style_file_with_normals_and_doubles(tmp_path):
    n = 5
    fields = [("x", "f4"), ("y", "f4"), ("z", "f4"), ("nx", "f4"), ("ny", "f4"), ("nz", "f4"),
              ("f_dc_0", "f4"), ("f_dc_1", "f4"), ("f_dc_2", "f4"), ("opacity", "f8"),
              ("scale_0", "f4"), ("scale_1", "f4"), ("scale_2", "f4"),
              ("rot_0", "f4"), ("rot_1", "f4"), ("rot_2", "f4"), ("rot_3", "f4"), ("flags", "u1")]
    v = np.zeros(n, dtype=fields)
    for i, (name, _) in enumerate(fields):
        v[name] = np.arange(n) + 10 * i
    p = str(tmp_path / "s.ply")
    ply.write_vertices(p, v)
    t = ply.ply_to_tensors(p)
    assert "channels" in t and "sh_coeffs_rest" not in t
    assert np.array_equal(t["mean"][:, 2].numpy(), np.arange(n) + 20)
    assert np.array_equal(t["opacity"][:, 0].numpy(), np.arange(n) + 90)
    assert np.array_equal(t["rotation"][:, 3].numpy(), np.arange(n) + 160)
    # ascii variant of the same vertex element
    a = str(tmp_path / "s_ascii.ply")
    with open(a, "w") as f:
        f.write("ply\nformat ascii 1.0\ncomment made by a test\nelement vertex %d\n" % n)
        for name, t_ in fields:
            f.write("property %s %s\n" % ({"f4": "float", "f8": "double", "u1": "uchar"}[t_], name))
        f.write("end_header\n")
        for row in v:
            f.write(" ".join(str(x) for x in row) + "\n")
    assert np.array_equal(ply.read_vertices(a), v)
