"""3DGS `.ply` <-> GRay safetensors bridge (reference
convert/safetensors_ply_conversion.py:21-146), without plyfile: a small
reader/writer for the PLY vertex element (binary little endian, or ASCII on
read).

Tensor <-> property mapping (kept identical to the reference so files move
between the two):
  mean [N,3]            x, y, z
  opacity [N,1]         opacity (logit)
  rotation [N,4]        rot_0..rot_3 (r, x, y, z)
  scale [N,3]           scale_0..scale_2 (log)
  sh_coeffs_dc [N,1,3]  f_dc_0..f_dc_2
  sh_coeffs_rest [N,K,3] f_rest_j, j = channel * K + k (channel-major, as 3DGS)
  current_sh_degree [1] sqrt(K + 1) - 1 on import
  channels [N,3]        f_dc_0..f_dc_2 when there are no f_rest properties
"""
from __future__ import annotations

import math
import os
from typing import Dict

import numpy as np

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2", "ushort": "u2",
    "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4",
    "double": "f8", "float64": "f8",
}


def read_vertices(path: str) -> np.ndarray:
    """The vertex element of a PLY file as a numpy structured array."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{path}: not a PLY file")
        fmt, count, props, in_vertex, elements_before = None, None, [], False, []
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: header without end_header")
            words = line.decode("ascii").split()
            if not words or words[0] in ("comment", "obj_info"):
                continue
            if words[0] == "end_header":
                break
            if words[0] == "format":
                fmt = words[1]
            elif words[0] == "element":
                in_vertex = words[1] == "vertex"
                if in_vertex:
                    count = int(words[2])
                elif count is None:
                    elements_before.append(words[1])
            elif words[0] == "property" and in_vertex:
                if words[1] == "list":
                    raise ValueError(f"{path}: list properties in the vertex element are not supported")
                props.append((words[2], _PLY_TYPES[words[1]]))
        if count is None:
            raise ValueError(f"{path}: no vertex element")
        if elements_before:
            raise ValueError(f"{path}: elements before the vertex element are not supported")
        if fmt == "binary_little_endian":
            dtype = np.dtype([(n, "<" + t) for n, t in props])
            data = np.frombuffer(f.read(dtype.itemsize * count), dtype=dtype, count=count)
            return data.copy()
        if fmt == "ascii":
            dtype = np.dtype([(n, t) for n, t in props])
            out = np.empty(count, dtype=dtype)
            for i in range(count):
                vals = f.readline().split()
                out[i] = tuple(np.array(v, dtype=t) for v, (_, t) in zip(vals, props))
            return out
        raise ValueError(f"{path}: unsupported PLY format {fmt}")


def write_vertices(path: str, vertex: np.ndarray) -> None:
    """Writes a structured array as the vertex element of a binary PLY."""
    names = {"f4": "float", "f8": "double", "i4": "int", "u4": "uint", "u1": "uchar", "i1": "char",
             "i2": "short", "u2": "ushort"}
    lines = ["ply", "format binary_little_endian 1.0", f"element vertex {len(vertex)}"]
    le = []
    for name in vertex.dtype.names:
        t = vertex.dtype.fields[name][0]
        lines.append(f"property {names[t.str[1:]]} {name}")
        le.append((name, "<" + t.str[1:]))
    lines.append("end_header")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as f:
        f.write(("\n".join(lines) + "\n").encode("ascii"))
        f.write(vertex.astype(np.dtype(le)).tobytes())


def ply_to_tensors(path: str) -> Dict[str, "object"]:
    """Reference ply_to_safetensors (:21-70) without the file write."""
    import torch

    v = read_vertices(path)
    n = len(v)
    f32 = lambda *keys: np.stack([v[k] for k in keys], axis=1).astype(np.float32)  # noqa: E731
    out = {
        "mean": torch.from_numpy(f32("x", "y", "z")),
        "opacity": torch.from_numpy(f32("opacity")),
        "rotation": torch.from_numpy(f32("rot_0", "rot_1", "rot_2", "rot_3")),
        "scale": torch.from_numpy(f32("scale_0", "scale_1", "scale_2")),
    }
    rest = sorted((k for k in v.dtype.names if k.startswith("f_rest_")), key=lambda k: int(k.rsplit("_", 1)[1]))
    if rest:
        if len(rest) % 3:
            raise ValueError(f"{path}: {len(rest)} f_rest properties is not a multiple of 3")
        k = len(rest) // 3
        deg = math.isqrt(k + 1) - 1
        if (deg + 1) ** 2 != k + 1:
            raise ValueError(f"{path}: {k} SH rest terms per channel is not (d+1)^2 - 1")
        sh_rest = np.ascontiguousarray(f32(*rest).reshape(n, 3, k).transpose(0, 2, 1))
        out["sh_coeffs_dc"] = torch.from_numpy(f32("f_dc_0", "f_dc_1", "f_dc_2")[:, None, :].copy())
        out["sh_coeffs_rest"] = torch.from_numpy(sh_rest)
        out["current_sh_degree"] = torch.tensor([deg], dtype=torch.int32)
    else:
        out["channels"] = torch.from_numpy(f32("f_dc_0", "f_dc_1", "f_dc_2"))
    return out


def tensors_to_ply(tensors: Dict[str, "object"], path: str) -> None:
    """Reference safetensors_to_ply (:73-146) from in-memory tensors."""
    missing = sorted({"mean", "rotation", "scale", "opacity"} - set(tensors))
    if missing:
        raise ValueError("missing Gaussian tensors: " + ", ".join(missing))
    a = lambda k: tensors[k].detach().cpu().numpy().astype(np.float32)  # noqa: E731
    has_sh = "sh_coeffs_dc" in tensors and "sh_coeffs_rest" in tensors
    if not has_sh and "channels" not in tensors:
        raise ValueError("neither SH coefficients nor channels present")
    mean, opacity, rotation, scale = a("mean"), a("opacity"), a("rotation"), a("scale")
    fields = [(n, "f4") for n in ("x", "y", "z", "opacity", "rot_0", "rot_1", "rot_2", "rot_3", "scale_0",
                                  "scale_1", "scale_2", "f_dc_0", "f_dc_1", "f_dc_2")]
    if has_sh:
        dc, rest = a("sh_coeffs_dc"), a("sh_coeffs_rest")
        fields += [(f"f_rest_{j}", "f4") for j in range(rest.shape[1] * rest.shape[2])]
    v = np.empty(mean.shape[0], dtype=fields)
    for i, n in enumerate(("x", "y", "z")):
        v[n] = mean[:, i]
    v["opacity"] = opacity[:, 0]
    for i in range(4):
        v[f"rot_{i}"] = rotation[:, i]
    for i in range(3):
        v[f"scale_{i}"] = scale[:, i]
    if has_sh:
        for i in range(3):
            v[f"f_dc_{i}"] = dc[:, 0, i]
        flat = np.ascontiguousarray(rest.transpose(0, 2, 1).reshape(rest.shape[0], -1))
        for j in range(flat.shape[1]):
            v[f"f_rest_{j}"] = flat[:, j]
    else:
        ch = a("channels")
        for i in range(3):
            v[f"f_dc_{i}"] = ch[:, i]
    write_vertices(path, v)


def ply_to_safetensors(input_path: str, output_path: str) -> None:
    import safetensors.torch

    os.makedirs(os.path.dirname(os.path.abspath(output_path)), exist_ok=True)
    safetensors.torch.save_file(ply_to_tensors(input_path), output_path)


def safetensors_to_ply(input_path: str, output_path: str) -> None:
    import safetensors.torch

    tensors_to_ply(safetensors.torch.load_file(input_path), output_path)
