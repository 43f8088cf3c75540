"""Shared set-up for the GPU parity tests (test infrastructure)."""
from __future__ import annotations

import math

import numpy as np
import torch

from gray_amd.camera import CameraInfo
from gray_amd.native import NativeRaytracer


def setup_native(scene, cam: CameraInfo, alpha_threshold=0.01, t_threshold=0.03, exp_power=2.0, bg=(0, 0, 0),
                 sh=True, sh_degree=3, channels=None, hit_capacity=None, znear=0.0, zfar=99999.9,
                 list_capacity=None, collect_budget=0, wave_list_min=0, collect_mode=0):
    n = scene.count if scene is not None else 0
    W, H = cam.image_width, cam.image_height
    nr = NativeRaytracer(W, H, n, 3, 300_000_000, 120_000_000, hit_capacity=hit_capacity)
    if list_capacity is not None:
        nr.list_capacity = list_capacity
    nr.collect_budget = collect_budget
    nr.wave_list_min = wave_list_min
    nr.collect_mode = collect_mode
    g = nr.get_gaussians()
    cfg = nr.get_config()
    cfg.alpha_threshold.fill_(alpha_threshold)
    cfg.t_threshold.fill_(t_threshold)
    cfg.exp_power.fill_(exp_power)
    cfg.background_channels.copy_(torch.tensor(bg, dtype=torch.float32))
    cfg.enable_sh.fill_(sh)
    cfg.update_channels.fill_(not sh)
    if n > 0:
        with torch.no_grad():
            g.mean.copy_(scene.mean)
            g.rotation.copy_(scene.rotation)
            g.scale.copy_(scene.scale)
            g.opacity.copy_(scene.opacity)
            g.sh_coeffs_dc.copy_(scene.sh_dc)
            g.sh_coeffs_rest.copy_(scene.sh_rest)
            if channels is not None:
                g.channels.copy_(channels)
            g.current_sh_degree.fill_(sh_degree)
    set_camera(nr, cam, znear, zfar)
    nr.rebuild_bvh()
    return nr


def set_camera(nr: NativeRaytracer, cam: CameraInfo, znear=0.0, zfar=99999.9):
    c = nr.get_camera()
    c.znear.fill_(znear)
    c.zfar.fill_(zfar)
    c.vertical_fov_radians.fill_(float(cam.fov_y))
    c.set_pose(cam.origin_cuda(), cam.rotation_c2w_blender_cuda())


def mock_camera(n=32) -> CameraInfo:
    """The reference tests' camera: origin (0,0,1), R = -I, fov_y 1.5."""
    return CameraInfo(uid=0, R=-np.eye(3), T=np.zeros(3), origin=np.array([0.0, 0.0, 1.0]), fov_y=1.5,
                      fov_x=1.5, image_width=n, image_height=n)


def rays_of(nr: NativeRaytracer):
    fb = nr.get_framebuffer()
    h, w = nr.height, nr.width
    ro = fb.ray_origin[:h, :w].reshape(-1, 3).cpu().numpy()
    rd = fb.ray_direction[:h, :w].reshape(-1, 3).cpu().numpy()
    return ro, rd


def rel_err(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-12)


def deg(x):
    return math.radians(x)


def per_element_err(a, b, atol_frac=1e-3):
    """Worst per-element relative error: max over elements of
    |a - b| / (|b| + atol_frac * max|b|).  A bound of 1e-3 on it means
    |a - b| <= 1e-3 |b| + 1e-6 max|b| element by element, so small gradient
    components are held to the relative bar, not just the largest one."""
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    floor = atol_frac * max(np.abs(b).max(), 1e-30)
    return float((np.abs(a - b) / (np.abs(b) + floor)).max())


def oracle_knobs(nr):
    """Oracle knobs from a NativeRaytracer's config and camera holders."""
    from oracle import oracle as O

    c = nr.get_config()
    cam = nr.get_camera()
    return O.make_knobs(exp_power=float(c.exp_power), alpha_threshold=float(c.alpha_threshold),
                        t_threshold=float(c.t_threshold), eps_fwd=float(c.eps_forward_normalization),
                        eps_scale_grad=float(c.eps_scale_grad), bg=tuple(c.background_channels.tolist()),
                        znear=float(cam.znear), zfar=float(cam.zfar), render_depth=bool(c.render_depth))


def comparable_gaussians(ncomp, fwd, geo, ro, rd, alpha_threshold=0.01, max_rays=4):
    """Mask of the Gaussians whose gradients compare element by element:
    those hit by no ray whose composited count differs between the GPU
    (ncomp, per ray) and the oracle forward `fwd`.  The stop test
    T < t_threshold compares alphas from two exp implementations, so a ray
    can composite one hit more or less (a handful per frame at most); the
    Gaussians such a ray hits are held to the frame-level bound only."""
    from oracle import oracle as O

    n = geo.shape[0]
    diff = np.nonzero(np.asarray(ncomp).reshape(-1) != fwd["ncomp"])[0]
    ok = np.ones(n, bool)
    if diff.size:
        assert diff.size <= max_rays, f"{diff.size} rays composite differently"
        ids, _, counts = O.hitlists(geo, ro[diff], rd[diff], alpha_threshold, n)
        for r in range(diff.size):
            ok[ids[r, :counts[r]]] = False
    return ok


def assert_grads_close(pairs, ok=None, bound=1e-3):
    """Every (name, gpu, reference) gradient within `bound` element by
    element (per_element_err) over the comparable Gaussians `ok` (rows), and
    within `bound` of the largest component over all of them."""
    for name, a, b in pairs:
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64).reshape(a.shape)
        sel = slice(None) if ok is None else ok
        assert per_element_err(a[sel], b[sel]) < bound, (name, per_element_err(a[sel], b[sel]))
        assert rel_err(a, b) < bound, (name, rel_err(a, b))
