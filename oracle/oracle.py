"""ctypes wrapper of the C oracle (oracle/gray_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product (gray_amd/).
Each wrapper names the reference function it restates (see gray_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgray_oracle.so")

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


class Knobs(C.Structure):
    _fields_ = [
        ("exp_power", C.c_float), ("alpha_threshold", C.c_float), ("t_threshold", C.c_float),
        ("eps_fwd", C.c_float), ("eps_scale_grad", C.c_float), ("bg", C.c_float * 3),
        ("znear", C.c_float), ("zfar", C.c_float), ("render_depth", C.c_int), ("fp16_hits", C.c_int),
    ]


def make_knobs(exp_power=2.0, alpha_threshold=0.01, t_threshold=0.01, eps_fwd=1e-12, eps_scale_grad=1e-12,
               bg=(0.0, 0.0, 0.0), znear=0.0, zfar=99999.9, render_depth=False, fp16_hits=False) -> Knobs:
    k = Knobs()
    k.exp_power, k.alpha_threshold, k.t_threshold = exp_power, alpha_threshold, t_threshold
    k.eps_fwd, k.eps_scale_grad = eps_fwd, eps_scale_grad
    for i in range(3):
        k.bg[i] = bg[i]
    k.znear, k.zfar, k.render_depth = znear, zfar, int(bool(render_depth))
    k.fp16_hits = int(bool(fp16_hits))
    return k


def ellipsoids(geo, rgb, ro, rd, min_opacity=0.0, znear=0.0, zfar=99999.9, bg=(0.0, 0.0, 0.0)):
    """ellipsoid_viewer.cu:1-94 (display mode): colours [R,3] and depth [R]."""
    geo, rgb, ro, rd = _f(geo), _f(rgb), _f(ro), _f(rd)
    r = ro.shape[0]
    out = np.zeros((r, 3), np.float32)
    depth = np.zeros(r, np.float32)
    lib().or_ellipsoids(C.c_int(geo.shape[0]), _p(geo), _p(rgb), C.c_int(r), _p(ro), _p(rd),
                        C.c_float(min_opacity), C.c_float(znear), C.c_float(zfar), _p(_f(bg)), _p(out), _p(depth))
    return out, depth


def fp16_round(x) -> np.ndarray:
    """float -> binary16 (round to nearest even) -> float, as the reference's
    __float2half_rn storage of hit distances and alphas
    (per_pixel_linked_list.h:28-29)."""
    x = _f(x)
    y = np.empty_like(x)
    lib().or_fp16_round_array(C.c_int64(x.size), _p(x), _p(y))
    return y


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _lib.or_num_threads.restype = C.c_int
    return _lib


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a, t=f32p):
    return a.ctypes.data_as(t) if a is not None else None


def snapshot(mean, rotation, scale, opacity, alpha_threshold=0.01, exp_power=2.0, gsf=1.0):
    """bvh_wrapper.cu:8-53 + kernel.cu:3-22: geo [N,16], l2w [N,12]."""
    mean, rotation, scale, opacity = _f(mean), _f(rotation), _f(scale), _f(opacity)
    n = mean.shape[0]
    geo = np.zeros((n, 16), np.float32)
    l2w = np.zeros((n, 12), np.float32)
    lib().or_snapshot(C.c_int(n), _p(mean), _p(rotation), _p(scale), _p(opacity), C.c_float(alpha_threshold),
                      C.c_float(exp_power), C.c_float(gsf), _p(geo), _p(l2w))
    return geo, l2w


def pinhole_rays(width, height, fov, origin, c2w):
    """camera.h:17-36: ray origins/directions [H*W, 3] (row-major pixels)."""
    ro = np.zeros((height * width, 3), np.float32)
    rd = np.zeros((height * width, 3), np.float32)
    lib().or_pinhole_rays(C.c_int(width), C.c_int(height), C.c_int(width), C.c_float(fov), _p(_f(origin)),
                          _p(_f(c2w)), _p(ro), _p(rd))
    return ro, rd


def pinhole_rays_jittered(width, height, fov, origin, c2w, calls):
    """camera.h:17-36 with the sub-pixel jitter seeded by (pixel, calls)."""
    ro = np.zeros((height * width, 3), np.float32)
    rd = np.zeros((height * width, 3), np.float32)
    lib().or_pinhole_rays_jittered(C.c_int(width), C.c_int(height), C.c_int(width), C.c_float(fov),
                                   _p(_f(origin)), _p(_f(c2w)), C.c_uint32(calls), _p(ro), _p(rd))
    return ro, rd


def tea(v0, v1, rounds, key) -> int:
    lib().or_tea.restype = C.c_uint32
    k = (C.c_uint32 * 4)(*key)
    return int(lib().or_tea(C.c_uint32(v0), C.c_uint32(v1), C.c_int(rounds), k))


def lcg(state) -> int:
    lib().or_lcg.restype = C.c_uint32
    return int(lib().or_lcg(C.c_uint32(state)))


def hitlists(geo, ro, rd, alpha_threshold=0.01, max_hits=None):
    """All hits of each ray sorted by (distance, id) (shaders.cu:22-47)."""
    geo, ro, rd = _f(geo), _f(ro), _f(rd)
    n, r = geo.shape[0], ro.shape[0]
    maxh = max_hits or max(n, 1)
    ids = np.zeros((r, maxh), np.uint32)
    dists = np.zeros((r, maxh), np.float32)
    counts = np.zeros(r, np.int32)
    lib().or_hitlists(C.c_int(n), _p(geo), C.c_float(alpha_threshold), C.c_int(r), _p(ro), _p(rd), C.c_int(maxh),
                      _p(ids, u32p), _p(dists), _p(counts, i32p))
    return ids, dists, counts


def render(geo, rgb, ro, rd, knobs: Knobs, mode: int = 0, cap: int = 4096):
    """forward_pass.cu:6-151 for explicit rays; returns a dict of per-ray arrays."""
    geo, rgb, ro, rd = _f(geo), _f(rgb), _f(ro), _f(rd)
    n, r = geo.shape[0], ro.shape[0]
    cap = max(1, min(cap, max(n, 1)))
    out = {
        "C": np.zeros((r, 3), np.float32), "D": np.zeros(r, np.float32), "T": np.zeros(r, np.float32),
        "F": np.zeros(r, np.float32), "est": np.zeros((r, 3), np.float32), "ncomp": np.zeros(r, np.int32),
        "ids": np.zeros((r, cap), np.uint32), "Tb": np.zeros((r, cap), np.float32), "nhits": np.zeros(r, np.int32),
    }
    lib().or_render(C.c_int(n), _p(geo), _p(rgb), C.c_int(r), _p(ro), _p(rd), C.byref(knobs), C.c_int(mode),
                    _p(out["C"]), _p(out["D"]), _p(out["T"]), _p(out["F"]), _p(out["est"]), _p(out["ncomp"], i32p),
                    C.c_int(cap), _p(out["ids"], u32p), _p(out["Tb"]), _p(out["nhits"], i32p))
    out["cap"] = cap
    return out


def backward(geo, l2w, rgb, rotation, scale, opacity, ro, rd, fwd, grad_out, knobs: Knobs, pixels: float):
    """backward_pass.cu:15-141: returns gradients (mean, rotation, scale, opacity, rgb, pruning)."""
    geo, l2w, rgb = _f(geo), _f(l2w), _f(rgb)
    rotation, scale, opacity = _f(rotation), _f(scale), _f(opacity)
    ro, rd, grad_out = _f(ro), _f(rd), _f(grad_out)
    n, r = geo.shape[0], ro.shape[0]
    g = {
        "mean": np.zeros((n, 3), np.float32), "rotation": np.zeros((n, 4), np.float32),
        "scale": np.zeros((n, 3), np.float32), "opacity": np.zeros((n, 1), np.float32),
        "rgb": np.zeros((n, 3), np.float32), "pruning": np.zeros((n, 1), np.float32),
    }
    lib().or_backward(C.c_int(n), _p(geo), _p(l2w), _p(rgb), _p(rotation), _p(scale), _p(opacity), C.c_int(r),
                      _p(ro), _p(rd), _p(grad_out), _p(fwd["T"]), _p(fwd["F"]), _p(fwd["est"]),
                      _p(fwd["ncomp"], i32p), C.c_int(fwd["cap"]), _p(fwd["ids"], u32p), C.byref(knobs),
                      C.c_float(pixels), _p(g["mean"]), _p(g["rotation"]), _p(g["scale"]), _p(g["opacity"]),
                      _p(g["rgb"]), _p(g["pruning"]))
    return g


def sh_basis(degree, x, y, z):
    """Real SH basis (15 rest coefficients) and partials by the reference's complex recursion."""
    b, bx, by, bz = (np.zeros(15, np.float32) for _ in range(4))
    lib().or_sh_basis(C.c_int(degree), C.c_float(x), C.c_float(y), C.c_float(z), _p(b), _p(bx), _p(by), _p(bz))
    return b, bx, by, bz


def sh_forward(mean, sh_dc, sh_rest, origin, degree):
    """sh.cu:224-247: per-Gaussian colour for a camera origin."""
    mean, dc, rest = _f(mean), _f(sh_dc), _f(sh_rest)
    n, k = mean.shape[0], rest.reshape(mean.shape[0], -1, 3).shape[1]
    rgb = np.zeros((n, 3), np.float32)
    lib().or_sh_forward(C.c_int(n), C.c_int(k), C.c_int(degree), _p(mean), _p(dc), _p(rest), _p(_f(origin)), _p(rgb))
    return rgb


def sh_backward(mean, sh_rest, origin, rgb, grad_rgb, visible, degree, grad_dc=None, grad_rest=None,
                grad_mean=None, accumulate=False):
    """sh.cu:249-287 (assignment of dc/rest grads unless accumulate)."""
    mean, rest, rgb, grad_rgb = _f(mean), _f(sh_rest), _f(rgb), _f(grad_rgb)
    n = mean.shape[0]
    k = rest.reshape(n, -1, 3).shape[1]
    gdc = np.zeros((n, 1, 3), np.float32) if grad_dc is None else _f(grad_dc).copy()
    grest = np.zeros((n, k, 3), np.float32) if grad_rest is None else _f(grad_rest).copy()
    gmean = np.zeros((n, 3), np.float32) if grad_mean is None else _f(grad_mean).copy()
    vis = None if visible is None else np.ascontiguousarray(visible, dtype=np.uint8)
    lib().or_sh_backward(C.c_int(n), C.c_int(k), C.c_int(degree), _p(mean), _p(rest), _p(_f(origin)), _p(rgb),
                         _p(grad_rgb), _p(vis, u8p), _p(gdc), _p(grest), _p(gmean), C.c_int(int(accumulate)))
    return gdc, grest, gmean


def adam(p, g, m, v, lr, beta1, beta2, eps, step, zero_grads=True):
    """adam.cu:13-22 in place on float32 arrays; step = steps after the increment."""
    for a in (p, g, m, v):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    lib().or_adam(C.c_int64(p.size), _p(p), _p(g), _p(m), _p(v), C.c_float(lr), C.c_float(beta1), C.c_float(beta2),
                  C.c_float(eps), C.c_int(step), C.c_int(int(zero_grads)))


def num_threads() -> int:
    return int(lib().or_num_threads())
