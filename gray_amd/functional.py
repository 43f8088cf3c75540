"""Functional multi-camera rendering with gradients for every parameter.

`MultiViewRenderer.render(cameras, mean, rotation, scale, opacity, sh_dc,
sh_rest)` renders a batch of pinhole cameras in launches of up to 8 views
(the C ABI's `gray_render_args.views`) and returns images [V, H, W, 3].  It is
a `torch.autograd.Function`: backward replays the stored composited hit
lists (rays whose list exceeded the store are recomputed by re-traversal)
and returns gradients for all six parameter tensors, the SH gradients summed
over the cameras (the reference's single-camera `sh_backward_kernel`
assigns, sh.cu:249-287; across several cameras only the sum is a gradient).

The reference renders one camera per call (gray/raytracer.py:116-171); the
per-hit semantics are those of cuda/forward_pass.cu and cuda/backward_pass.cu.
Hit-list memory (collect lists of one launch, their task buffers and the
stored composited lists of every view) is held under `memory_cap_bytes`.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import torch

from . import _lib as L
from .camera import CameraInfo

F32 = torch.float32
MAX_VIEWS_PER_LAUNCH = 8
TASK_BYTES_PER_LANE = 256  # two task buffers of 16 {lane, ref} pairs per lane
MAX_STORED_HITS = 64
MAX_LIST_HITS = 512


@dataclass
class RenderSettings:
    """Knobs of the reference config (cuda/core/config.h; same defaults)."""
    exp_power: float = 2.0
    alpha_threshold: float = 0.01
    t_threshold: float = 0.01
    global_scale_factor: float = 1.0
    eps_forward_normalization: float = 1e-12
    eps_scale_grad: float = 1e-12
    background: Sequence[float] = (0.0, 0.0, 0.0)
    znear: float = 0.0
    zfar: float = 99999.9
    sh_degree: int = 3
    memory_cap_bytes: int = 4 << 30
    list_capacity: Optional[int] = None   # None: from the memory cap
    fp16_hits: bool = False               # the reference's fp16 hit storage (distances, compositing alphas)
    hit_capacity: Optional[int] = None    # None: from the memory cap
    deterministic: bool = False           # bit-reproducible gradients (ordered sums instead of atomics; slower)


def capacities(views: int, width: int, height: int, cap_bytes: int):
    """(stored hits per ray, collected hits per ray) that keep the hit-list
    memory of a batch under cap_bytes: 8 B per stored entry for every ray of
    every view, 16 B per list entry and TASK_BYTES_PER_LANE per lane of the
    largest launch."""
    tiles = -(-width // 8) * -(-height // 8)
    lanes = min(views, MAX_VIEWS_PER_LAUNCH) * tiles * 64
    rays = views * width * height
    avail = cap_bytes - lanes * TASK_BYTES_PER_LANE
    if avail <= 0:
        return 0, 0
    stored = int(min(MAX_STORED_HITS, 0.2 * avail / (8 * rays)))
    lists = int(min(MAX_LIST_HITS, (avail - 8 * stored * rays) / (16 * lanes)))
    return max(stored, 0), max(lists, 0)


class MultiViewRenderer:
    """Owns a library context (snapshot, BVH, hit lists) and the per-view
    buffers of one image size."""

    def __init__(self, width: int, height: int, settings: Optional[RenderSettings] = None, device=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.W, self.H = int(width), int(height)
        self.settings = settings or RenderSettings()
        self.ctx = L.Context()
        s = self.settings
        t = lambda v, dt=F32: torch.tensor([v], dtype=dt, device=self.device)  # noqa: E731
        self._k = {
            "exp_power": t(s.exp_power), "alpha_threshold": t(s.alpha_threshold), "t_threshold": t(s.t_threshold),
            "global_scale_factor": t(s.global_scale_factor),
            "eps_forward_normalization": t(s.eps_forward_normalization), "eps_scale_grad": t(s.eps_scale_grad),
            "background": torch.tensor(list(s.background), dtype=F32, device=self.device),
            "render_depth": t(False, torch.bool), "enable_sh": t(True, torch.bool),
        }
        self._sh_degree = torch.tensor([s.sh_degree], dtype=torch.int32, device=self.device)
        self._V = 0
        self._built_n = -1
        self.last_capacities = (0, 0)
        self.generation = 0  # forwards so far; a backward must follow its own forward

    # ---- plumbing ---------------------------------------------------------
    def knobs(self) -> L.Knobs:
        k = self._k
        return L.Knobs(k["exp_power"].data_ptr(), k["alpha_threshold"].data_ptr(), k["t_threshold"].data_ptr(),
                       k["global_scale_factor"].data_ptr(), k["eps_forward_normalization"].data_ptr(),
                       k["eps_scale_grad"].data_ptr(), k["background"].data_ptr(), k["render_depth"].data_ptr(),
                       k["enable_sh"].data_ptr())

    def _alloc(self, V: int, n: int) -> None:
        dev, H, W = self.device, self.H, self.W
        s = self.settings
        stored, lists = capacities(V, W, H, s.memory_cap_bytes)
        if s.hit_capacity is not None:
            stored = int(s.hit_capacity)
        if s.list_capacity is not None:
            lists = int(s.list_capacity)
        self.last_capacities = (stored, lists)
        z = lambda *shape, dt=F32: torch.zeros(*shape, dtype=dt, device=dev)  # noqa: E731
        self.cam_origin = z(V, 3)
        self.cam_c2w = z(V, 3, 3)
        self.cam_fov = z(V)
        self.cam_znear = torch.full((V,), float(s.znear), dtype=F32, device=dev)
        self.cam_zfar = torch.full((V,), float(s.zfar), dtype=F32, device=dev)
        self.out = z(V, H, W, 3)
        self.T = z(V, H, W)
        self.F = z(V, H, W)
        self.est = z(V, H, W, 3)
        self.ro = z(V, H, W, 3)
        self.rd = z(V, H, W, 3)
        self.hit_ids = z(V, max(stored, 1), H * W, dt=torch.int32)
        self.hit_T = z(V, max(stored, 1), H * W)
        self.n_comp = z(V, H * W, dt=torch.int32)
        self.comp_rgb = z(V, H * W, 3)
        self.grad_out = z(V, H, W, 3)
        self.rgb = z(V, n, 3)
        self.grad_rgb = z(V, n, 3)
        self._V, self._n = V, n

    def _views(self, V: int) -> "C.Array":
        arr = (L.View * V)()
        for i in range(V):
            v = arr[i]
            v.origin = self.cam_origin[i].data_ptr()
            v.rotation_c2w = self.cam_c2w[i].data_ptr()
            v.vertical_fov = self.cam_fov[i:i + 1].data_ptr()
            v.znear = self.cam_znear[i:i + 1].data_ptr()
            v.zfar = self.cam_zfar[i:i + 1].data_ptr()
            v.width, v.height, v.stride, v.alloc_height = self.W, self.H, self.W, self.H
            v.rgb = self.rgb[i].data_ptr()
            v.out_channels = self.out[i].data_ptr()
            v.out_transmittance = self.T[i].data_ptr()
            v.out_full_transmittance = self.F[i].data_ptr()
            v.out_remaining_est = self.est[i].data_ptr()
            v.out_ray_origin = self.ro[i].data_ptr()
            v.out_ray_direction = self.rd[i].data_ptr()
            v.hit_ids = self.hit_ids[i].data_ptr()
            v.hit_trans = self.hit_T[i].data_ptr()
            v.n_comp = self.n_comp[i].data_ptr()
            v.comp_rgb = self.comp_rgb[i].data_ptr()
            v.grad_out = self.grad_out[i].data_ptr()
            v.grad_rgb = self.grad_rgb[i].data_ptr()
        return arr

    def _args(self, views, V: int, save: bool) -> L.RenderArgs:
        a = L.RenderArgs()
        a.knobs = self.knobs()
        a.num_views = V
        a.views = C.cast(views, C.POINTER(L.View))
        a.save = 1 if save else 0
        a.hit_capacity, a.list_capacity = self.last_capacities
        a.fp16_hits = 1 if self.settings.fp16_hits else 0
        a.deterministic = 1 if self.settings.deterministic else 0
        a.use_bvh = 1
        return a

    def _params(self, mean, rotation, scale, opacity, sh_dc, sh_rest) -> L.Params:
        p = L.Params()
        p.count = mean.shape[0]
        p.num_sh_coeffs = sh_rest.shape[1]  # non-DC coefficients per Gaussian (15 at degree 3)
        p.sh_is_fp16 = 0
        p.current_sh_degree = self._sh_degree.data_ptr()
        p.mean, p.rotation, p.scale, p.opacity = (t.data_ptr() for t in (mean, rotation, scale, opacity))
        p.sh_dc, p.sh_rest = sh_dc.data_ptr(), sh_rest.data_ptr()
        return p

    def set_cameras(self, cameras: Sequence[CameraInfo]) -> None:
        for i, cam in enumerate(cameras):
            if (cam.image_width, cam.image_height) != (self.W, self.H):
                raise ValueError(f"camera {i} is {cam.image_width}x{cam.image_height}, renderer is {self.W}x{self.H}")
            self.cam_origin[i].copy_(cam.origin_cuda())
            self.cam_c2w[i].copy_(cam.rotation_c2w_blender_cuda())
            self.cam_fov[i] = float(cam.fov_y)

    # ---- passes -----------------------------------------------------------
    def forward(self, cameras: Sequence[CameraInfo], mean, rotation, scale, opacity, sh_dc, sh_rest,
                save: bool = True, rebuild: bool = False) -> torch.Tensor:
        V, n = len(cameras), mean.shape[0]
        if V == 0:
            raise ValueError("no cameras")
        if V != self._V or n != self._n:
            self._alloc(V, n)
        self.set_cameras(cameras)
        params = self._params(mean, rotation, scale, opacity, sh_dc, sh_rest)
        knobs = self.knobs()
        stream = L.stream_ptr()
        rebuild = rebuild or n != self._built_n
        L.check(L.LIB.gray_bvh_update(self.ctx.handle, C.byref(params), C.byref(knobs), 1 if rebuild else 0, stream))
        self._built_n = n
        views = self._views(V)
        args = self._args(views, V, save)
        # the launch computes each view's SH colours itself, beside its collect pass
        args.sh_params = C.cast(C.pointer(params), C.c_void_p)
        L.check(L.LIB.gray_render_forward(self.ctx.handle, C.byref(args), stream))
        self.generation += 1
        return self.out

    def backward(self, grad_images: torch.Tensor, mean, rotation, scale, opacity, sh_dc, sh_rest):
        V, n = self._V, self._n
        self.grad_out.copy_(grad_images.reshape(V, self.H, self.W, 3))
        self.grad_rgb.zero_()
        g_mean = torch.zeros_like(mean)
        g_rot = torch.zeros_like(rotation)
        g_scale = torch.zeros_like(scale)
        g_op = torch.zeros_like(opacity)
        g_dc = torch.zeros_like(sh_dc)
        g_rest = torch.zeros_like(sh_rest)
        visible = torch.zeros(n, dtype=torch.uint8, device=self.device)
        params = self._params(mean, rotation, scale, opacity, sh_dc, sh_rest)
        grads = L.Grads(g_mean.data_ptr(), g_rot.data_ptr(), g_scale.data_ptr(), g_op.data_ptr(), None,
                        visible.data_ptr())
        stream = L.stream_ptr()
        views = self._views(V)
        L.check(L.LIB.gray_render_backward(self.ctx.handle, C.byref(self._args(views, V, True)), C.byref(params),
                                           C.byref(grads), stream))
        origins = (C.c_void_p * V)(*[self.cam_origin[i].data_ptr() for i in range(V)])
        rgb = (C.c_void_p * V)(*[self.rgb[i].data_ptr() for i in range(V)])
        grgb = (C.c_void_p * V)(*[self.grad_rgb[i].data_ptr() for i in range(V)])
        L.check(L.LIB.gray_sh_backward(self.ctx.handle, C.byref(params), C.byref(self.knobs()), V, origins, rgb, grgb,
                                       visible.data_ptr(), g_dc.data_ptr(), g_rest.data_ptr(), g_mean.data_ptr(), 1,
                                       stream))
        return g_mean, g_rot, g_scale, g_op, g_dc, g_rest

    def debug_snapshot(self):
        """Geometry snapshot of the last forward: geo [N,16], aabb [N,6]."""
        geo = torch.zeros(self._n, 16, dtype=F32, device=self.device)
        aabb = torch.zeros(self._n, 6, dtype=F32, device=self.device)
        L.check(L.LIB.gray_debug_snapshot(self.ctx.handle, geo.data_ptr(), aabb.data_ptr(), L.stream_ptr()))
        return geo, aabb

    def device_status(self) -> None:
        L.check(L.LIB.gray_device_status(self.ctx.handle, L.stream_ptr()))

    def render(self, cameras: Sequence[CameraInfo], mean, rotation, scale, opacity, sh_dc, sh_rest,
               rebuild: bool = False) -> torch.Tensor:
        """Images [V, H, W, 3]; differentiable in all six parameter tensors."""
        return _RenderViews.apply(self, tuple(cameras), rebuild, mean, rotation, scale, opacity, sh_dc, sh_rest)


def _prep(t: torch.Tensor, shape_tail) -> torch.Tensor:
    t = t.detach()
    if t.dtype != F32 or not t.is_contiguous():
        t = t.to(F32).contiguous()
    if tuple(t.shape[1:]) != tuple(shape_tail):
        raise ValueError(f"expected [N, {', '.join(map(str, shape_tail))}], got {list(t.shape)}")
    return t


class _RenderViews(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer: MultiViewRenderer, cameras, rebuild, mean, rotation, scale, opacity, sh_dc, sh_rest):
        p = (_prep(mean, (3,)), _prep(rotation, (4,)), _prep(scale, (3,)), _prep(opacity, (1,)),
             _prep(sh_dc, (1, 3)), _prep(sh_rest, (sh_rest.shape[1], 3)))
        out = renderer.forward(cameras, *p, save=torch.is_grad_enabled() or any(
            t.requires_grad for t in (mean, rotation, scale, opacity, sh_dc, sh_rest)), rebuild=rebuild)
        ctx.renderer = renderer
        ctx.params = p
        ctx.generation = renderer.generation
        return out.clone()

    @staticmethod
    def backward(ctx, grad_images):
        if ctx.renderer.generation != ctx.generation:
            raise RuntimeError("MultiViewRenderer: backward of an older render (its hit lists were overwritten by a "
                               "later forward); call backward before rendering again")
        grads = ctx.renderer.backward(grad_images.contiguous(), *ctx.params)
        return (None, None, None) + tuple(grads)
