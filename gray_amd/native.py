"""Native facade with the interface of the reference's torch custom class.

`NativeRaytracer` is what `Raytracer.cuda_module` holds.  It mirrors
`torch.classes.gray.Raytracer` (reference cuda/raytracer.h:29-184,
raytracer.cpp:3-13): same constructor arguments, the same pass methods and
getters returning holders whose attributes are the same device tensors
(reference cuda/core/*.h bind() tables).  Python mutates them in place
(fill_/copy_); unlike the reference no raw pointers are baked into a device
struct, every call passes the current tensors to the C ABI.

Differences by design: there are no per-pixel linked lists (the k-buffer
replaces them); `get_ppll_forward_data` / `get_ppll_backward_data` return the
hit store used by the backward instead.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L

F32 = torch.float32
# hits per ray kept by the forward's collect pass (16 bytes each): 512 covers
# the grazing rays of dense surface scenes; longer lists re-traverse
DEFAULT_LIST_CAPACITY = 512


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("gray_amd needs a ROCm GPU (torch.cuda is unavailable); there is no CPU fallback")
    return torch.device("cuda", torch.cuda.current_device())


class CameraDataHolder:
    """Reference cuda/core/camera.h:167-201."""

    def __init__(self, device):
        self.origin = torch.zeros(3, dtype=F32, device=device)
        self.vertical_fov_radians = torch.zeros(1, dtype=F32, device=device)
        self.rotation_c2w = torch.zeros(3, 3, dtype=F32, device=device)
        self.rotation_w2c = torch.zeros(3, 3, dtype=F32, device=device)
        self.znear = torch.zeros(1, dtype=F32, device=device)
        self.zfar = torch.zeros(1, dtype=F32, device=device)

    def set_pose(self, c2w_origin: torch.Tensor, c2w_rotation: torch.Tensor) -> None:
        if tuple(c2w_rotation.shape) != (3, 3):
            raise RuntimeError("c2w_rotation must be 3x3")
        if tuple(c2w_origin.shape) != (3,):
            raise RuntimeError("c2w_origin must be 3")
        self.rotation_c2w.copy_(c2w_rotation)
        self.origin.copy_(c2w_origin)
        self.rotation_w2c.copy_(c2w_rotation.transpose(0, 1))


class ConfigDataHolder:
    """Reference cuda/core/config.h:229-288 (same knobs and defaults)."""

    def __init__(self, device):
        t = lambda v, dt=F32: torch.tensor([v], dtype=dt, device=device)  # noqa: E731
        b = lambda v: torch.tensor([v], dtype=torch.bool, device=device)  # noqa: E731
        self.exp_power = t(2.0)
        self.alpha_threshold = t(0.01)
        self.t_threshold = t(0.01)
        self.jitter_primary_rays = b(False)
        self.global_scale_factor = t(1.0)
        self.eps_forward_normalization = t(1e-12)
        self.eps_scale_grad = t(1e-12)
        self.background_channels = torch.zeros(L.NUM_CHANNELS, dtype=F32, device=device)
        self.render_depth = b(False)
        # read only by the host (the display switch and the no-gradient check
        # of Raytracer.__call__), so it lives in host memory: reading it costs
        # no device synchronisation; fill_/copy_ work on it as on the others
        self.render_ellipsoids = torch.tensor([False], dtype=torch.bool)
        self.ellipsoid_min_opacity = t(0.0)
        self.rays_from_python = b(False)
        self.zero_grads = b(True)
        self.enable_sh = b(True)
        self.has_pre_mlp = b(False)
        self.needs_ray_output = b(False)
        self.update_channels = b(True)

    def knobs(self, call_index=None) -> L.Knobs:
        """call_index: the metadata's total_num_calls (seed of the ray jitter)."""
        return L.Knobs(
            self.exp_power.data_ptr(), self.alpha_threshold.data_ptr(), self.t_threshold.data_ptr(),
            self.global_scale_factor.data_ptr(), self.eps_forward_normalization.data_ptr(),
            self.eps_scale_grad.data_ptr(), self.background_channels.data_ptr(), self.render_depth.data_ptr(),
            self.enable_sh.data_ptr(), self.jitter_primary_rays.data_ptr(),
            call_index.data_ptr() if call_index is not None else None, self.needs_ray_output.data_ptr(),
        )


class FramebufferDataHolder:
    """Reference cuda/core/framebuffer.h:72-121 ([H, W, C] row-major, allocated size)."""

    def __init__(self, width: int, height: int, device):
        z = lambda c: torch.zeros(height, width, c, dtype=F32, device=device)  # noqa: E731
        self.output_channels = z(L.NUM_CHANNELS)
        self.output_depth = z(1)
        self.transmittance = z(1)
        self.full_transmittance = z(1)
        self.remaining_channels_estimate = z(L.NUM_CHANNELS)
        self.ray_origin = z(3)
        self.ray_direction = z(3)
        self.grad_output_channels = z(L.NUM_CHANNELS)


class GaussianDataHolder:
    """Reference cuda/core/gaussians.h:70-419: parameters, gradients (attached as
    .grad), Adam moments, per-tensor learning rates and pruning statistics."""

    PARAMS = ("mean", "rotation", "scale", "opacity", "channels", "sh_coeffs_dc", "sh_coeffs_rest")
    LRS = ("lr_mean", "lr_rotation", "lr_scale", "lr_opacity", "lr_channels", "lr_sh_dc", "lr_sh_rest")

    def __init__(self, max_sh_degree: int, inference_only: bool, device):
        self.device = device
        self.count = 1
        self.max_sh_degree = int(max_sh_degree)
        self.num_sh_coeffs = (self.max_sh_degree + 1) ** 2 - 1
        self.inference_only = bool(inference_only)
        self.current_sh_degree = torch.tensor([0], dtype=torch.int32, device=device)
        sh_dtype = torch.float16 if inference_only else F32
        z = lambda *s, dt=F32: torch.zeros(*s, dtype=dt, device=device)  # noqa: E731
        self.mean = z(1, 3)
        self.rotation = z(1, 4)
        self.scale = z(1, 3)
        self.opacity = z(1, 1)
        self.channels = z(1, L.NUM_CHANNELS)
        self.sh_coeffs_dc = z(1, 1, L.NUM_CHANNELS, dt=sh_dtype)
        self.sh_coeffs_rest = z(1, self.num_sh_coeffs, L.NUM_CHANNELS, dt=sh_dtype)
        for name in self.LRS:
            setattr(self, name, torch.ones(1, dtype=F32, device=device))
        self.beta_1 = torch.tensor([0.9], dtype=F32, device=device)
        self.beta_2 = torch.tensor([0.999], dtype=F32, device=device)
        self.epsilon = torch.tensor([1e-15], dtype=F32, device=device)
        self.sh_update_laziness = torch.tensor([1], dtype=torch.int32, device=device)
        # gradients, moments and pruning statistics stay at one element when
        # inference_only: resize() leaves them alone (gaussians.h:132-197, 236-262)
        for prefix in ("grad_", "first_moment_", "second_moment_"):
            for name in self.PARAMS:
                shape = tuple(getattr(self, name).shape)
                setattr(self, prefix + name, z(*shape))
        self.pruning_weight = z(1, 1)
        self.pruning_counter = z(1, 1, dt=torch.int32)
        self.was_visible = z(1, 1, dt=torch.bool)
        self.marked = z(1, 1)
        self._attach_grads()

    def _attach_grads(self) -> None:
        if self.inference_only:
            return
        for name in self.PARAMS:
            getattr(self, name).grad = getattr(self, "grad_" + name)

    def _shape(self, name: str, n: int):
        if name in ("mean", "scale"):
            return (n, 3)
        if name == "rotation":
            return (n, 4)
        if name == "opacity":
            return (n, 1)
        if name == "channels":
            return (n, L.NUM_CHANNELS)
        if name == "sh_coeffs_dc":
            return (n, 1, L.NUM_CHANNELS)
        return (n, self.num_sh_coeffs, L.NUM_CHANNELS)

    def resize(self, n: int) -> None:
        """Reference gaussians.h:198-263 (in-place resize_, storage kept)."""
        if self.count > 1 and n > self.count:
            raise RuntimeError("The new number of Gaussians must not be larger than the current count.")
        old = self.count
        with torch.no_grad():
            self.count = int(n)
            # resize_ keeps the prefix; elements it newly exposes are
            # uninitialised memory, so they are cleared (learning rates copy [0])
            prefixes = ("",) if self.inference_only else ("", "grad_", "first_moment_", "second_moment_")
            for name in self.PARAMS:
                shp = self._shape(name, n)
                for prefix in prefixes:
                    t = getattr(self, prefix + name)
                    t.resize_(shp)
                    if n > old:
                        t[old:].zero_()
            for name in self.LRS:
                t = getattr(self, name)
                t.resize_(n, 1)
                if n > old:
                    t[old:].copy_(t[:1].expand(n - old, 1))
            if not self.inference_only:
                for name in ("pruning_weight", "pruning_counter", "was_visible", "marked"):
                    t = getattr(self, name)
                    t.resize_(n, 1)
                    if n > old:
                        t[old:].zero_()
        self._attach_grads()

    def increment_sh_degree(self) -> None:
        with torch.no_grad():
            self.current_sh_degree.add_((self.current_sh_degree < self.max_sh_degree).to(torch.int32))

    def params(self) -> L.Params:
        return L.Params(
            self.count, self.num_sh_coeffs, 1 if self.inference_only else 0, self.current_sh_degree.data_ptr(),
            self.mean.data_ptr(), self.rotation.data_ptr(), self.scale.data_ptr(), self.opacity.data_ptr(),
            self.sh_coeffs_dc.data_ptr(), self.sh_coeffs_rest.data_ptr(),
        )


class MetaDataHolder:
    """Reference cuda/core/metadata.h:14-46."""

    def __init__(self, device):
        self.run_forward_pass = torch.ones(1, dtype=torch.bool, device=device)
        self.run_backward_pass = torch.ones(1, dtype=torch.bool, device=device)
        self.grads_enabled = torch.ones(1, dtype=torch.bool, device=device)
        self.total_num_calls = torch.zeros(1, dtype=torch.int32, device=device)
        self.total_num_steps = torch.zeros(1, dtype=torch.int32, device=device)

    def update(self) -> None:
        self.grads_enabled.fill_(torch.is_grad_enabled())
        with torch.no_grad():
            self.total_num_calls.add_(1)


class StatsDataHolder:
    """Reference cuda/core/stats.h:10-39."""

    def __init__(self, width: int, height: int, device):
        self.num_gaussians_hit = torch.zeros(height, width, dtype=torch.int32, device=device)
        self.num_gaussians_accumulated = torch.zeros(height, width, dtype=torch.int32, device=device)

    def reset(self) -> None:
        self.num_gaussians_hit.zero_()
        self.num_gaussians_accumulated.zero_()


class HitStoreHolder:
    """Per-ray composited hit lists kept by the forward for the backward
    (replaces the reference per-pixel linked lists, per_pixel_linked_list.h).
    hit_ids/hit_trans are [capacity, H*W] (slot-major), n_comp [H*W]."""

    def __init__(self, width: int, height: int, capacity: int, device):
        r = width * height
        self.capacity = int(capacity)
        self.hit_ids = torch.zeros(max(capacity, 1), r, dtype=torch.int32, device=device)
        self.hit_trans = torch.zeros(max(capacity, 1), r, dtype=F32, device=device)
        self.n_comp = torch.zeros(r, dtype=torch.int32, device=device)
        self.comp_rgb = torch.zeros(r, L.NUM_CHANNELS, dtype=F32, device=device)


class PhaseTimer:
    """HIP-event timing of the native calls, on the stream they are issued on
    (torch's current stream).  Enabled by assigning NativeRaytracer.timer."""

    def __init__(self):
        self.events = {}

    def start(self, name):
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return (name, ev)

    def stop(self, token):
        name, ev0 = token
        ev1 = torch.cuda.Event(enable_timing=True)
        ev1.record()
        self.events.setdefault(name, []).append((ev0, ev1))

    def reset(self):
        self.events = {}

    def summary(self):
        """{phase: (launches, mean ms, total ms)} -- synchronises."""
        torch.cuda.synchronize()
        out = {}
        for name, pairs in self.events.items():
            ms = [a.elapsed_time(b) for a, b in pairs]
            out[name] = (len(ms), sum(ms) / len(ms), sum(ms))
        return out


class _NoTimer:
    def start(self, name):
        return None

    def stop(self, token):
        pass


def hit_capacity_for(pixels: int, ppll_backward_size: int, max_capacity: int = 64) -> int:
    """Stored hits per ray from the reference's backward-list budget
    (config.py:46-47, entries): a ray whose composited list is longer is
    recomputed by re-traversal in the backward."""
    if pixels <= 0:
        return 0
    return int(max(0, min(max_capacity, ppll_backward_size // pixels)))


class NativeRaytracer:
    """Mirror of the reference torch custom class gray.Raytracer."""

    def __init__(self, width: int, height: int, num_gaussians: int, sh_max_degree: int,
                 ppll_forward_size: int, ppll_backward_size: int, inference_only: bool = False,
                 hit_capacity: int | None = None):
        dev = _device()
        self.device = dev
        self.image_width = int(width)
        self.image_height = int(height)
        self.width = int(width)
        self.height = int(height)
        self.inference_only = bool(inference_only)
        self.ctx = L.Context()
        self.camera_data = CameraDataHolder(dev)
        self.config_data = ConfigDataHolder(dev)
        self.framebuffer_data = FramebufferDataHolder(width, height, dev)
        self.gaussian_data = GaussianDataHolder(sh_max_degree, inference_only, dev)
        self.meta_data = MetaDataHolder(dev)
        self.stats_data = StatsDataHolder(width, height, dev)
        if hit_capacity is None:
            hit_capacity = 0 if inference_only else hit_capacity_for(width * height, ppll_backward_size)
        self.hits = HitStoreHolder(width, height, hit_capacity, dev)
        self.ppll_forward_size = int(ppll_forward_size)
        self.ppll_backward_size = int(ppll_backward_size)
        self.use_bvh = True
        # hits collected per ray before compositing (0: composite while
        # traversing); rays with more hits are re-traversed per batch
        self.list_capacity = DEFAULT_LIST_CAPACITY
        self.collect_budget = 0  # walk iterations before work is split into tasks (0: built-in)
        # the reference's fp16 storage of hit distances and alphas
        # (per_pixel_linked_list.h:28-29); False: fp32 throughout
        self.fp16_hits = False
        self.wave_list_min = 0  # lists longer than this are composited by a whole wave (0: built-in, < 0: never)
        # reproducible gradients: walks without continuation tasks and
        # per-Gaussian sums in pixel order instead of atomics (slower)
        self.deterministic = False
        # collect pass: 0 = tile bins for pinhole views (BVH walk for explicit
        # rays), 1 = the BVH walk for every view
        self.collect_mode = 0
        self.tile_stats = None  # optional int32 [tiles, 4] traversal statistics of the forward
        self.timer = _NoTimer()
        self._built_count = -1
        self._frame_prepared = False  # prepare_frame made the forward's resets
        self._forward_saved = False
        if num_gaussians > 0:
            self.resize(num_gaussians)

    # ---- helpers ----------------------------------------------------------
    def _stream(self) -> int:
        return L.stream_ptr()

    def _view(self, save: bool) -> L.View:
        fb, cam, g, st, h = self.framebuffer_data, self.camera_data, self.gaussian_data, self.stats_data, self.hits
        cfg = self.config_data
        v = L.View()
        v.origin = cam.origin.data_ptr()
        v.rotation_c2w = cam.rotation_c2w.data_ptr()
        v.vertical_fov = cam.vertical_fov_radians.data_ptr()
        v.znear = cam.znear.data_ptr()
        v.zfar = cam.zfar.data_ptr()
        # explicit rays from the framebuffer, switched on the device by
        # config.rays_from_python (reference shaders.cu:80-85)
        v.rays_origin = fb.ray_origin.data_ptr()
        v.rays_direction = fb.ray_direction.data_ptr()
        v.rays_enable = cfg.rays_from_python.data_ptr()
        v.width, v.height = self.width, self.height
        v.stride, v.alloc_height = self.image_width, self.image_height
        v.rgb = g.channels.data_ptr()
        v.out_channels = fb.output_channels.data_ptr()
        v.out_depth = fb.output_depth.data_ptr()
        v.out_ray_direction = fb.ray_direction.data_ptr()  # without saved state: only with needs_ray_output
        if save:
            v.out_transmittance = fb.transmittance.data_ptr()
            v.out_full_transmittance = fb.full_transmittance.data_ptr()
            v.out_remaining_est = fb.remaining_channels_estimate.data_ptr()
            v.out_ray_origin = fb.ray_origin.data_ptr()
            v.out_ray_direction = fb.ray_direction.data_ptr()
            v.stat_hits = st.num_gaussians_hit.data_ptr()
            v.stat_accumulated = st.num_gaussians_accumulated.data_ptr()
            v.hit_ids = h.hit_ids.data_ptr()
            v.hit_trans = h.hit_trans.data_ptr()
            v.n_comp = h.n_comp.data_ptr()
            v.comp_rgb = h.comp_rgb.data_ptr()
        v.grad_out = fb.grad_output_channels.data_ptr()
        v.grad_rgb = g.grad_channels.data_ptr()
        return v

    def _args(self, view: L.View, save: bool) -> L.RenderArgs:
        a = L.RenderArgs()
        a.knobs = self.config_data.knobs(self.meta_data.total_num_calls)
        a.num_views = 1
        self._view_keep = (view,)
        a.views = C.pointer(view)
        a.save = 1 if save else 0
        a.hit_capacity = self.hits.capacity
        a.use_bvh = 1 if self.use_bvh else 0
        a.tile_stats = self.tile_stats.data_ptr() if self.tile_stats is not None else None
        a.list_capacity = int(self.list_capacity)
        a.collect_budget = int(self.collect_budget)
        a.fp16_hits = 1 if self.fp16_hits else 0
        a.wave_list_min = int(self.wave_list_min)
        a.deterministic = 1 if self.deterministic else 0
        a.collect_mode = int(self.collect_mode)
        return a

    def _require_bvh(self) -> None:
        if self._built_count != self.gaussian_data.count:
            raise RuntimeError("rebuild_bvh() must be called after setting the Gaussians (and after resize)")

    def _sh_forward(self) -> None:
        g = self.gaussian_data
        origins = (C.c_void_p * 1)(self.camera_data.origin.data_ptr())
        outs = (C.c_void_p * 1)(g.channels.data_ptr())
        tk = self.timer.start("sh_forward")
        L.check(L.LIB.gray_sh_forward(self.ctx.handle, C.byref(g.params()), C.byref(self.config_data.knobs()), 1,
                                      origins, outs, self._stream()))
        self.timer.stop(tk)

    # ---- passes -----------------------------------------------------------
    def prepare_frame(self, origin, rotation_c2w, vertical_fov: float, znear: float, zfar: float) -> None:
        """One launch for the holder updates before a frame: the camera fills
        and set_pose of the Python layer (gray/raytracer.py:126-132, host
        values here), rays_from_python cleared, and the metadata update and
        was_visible reset that forward_pass would make
        (raytracer.h:88-93).  The next forward_pass skips those two."""
        cam, g = self.camera_data, self.gaussian_data
        a = L.FrameArgs()
        a.origin, a.rotation_c2w, a.rotation_w2c = (cam.origin.data_ptr(), cam.rotation_c2w.data_ptr(),
                                                    cam.rotation_w2c.data_ptr())
        a.vertical_fov, a.znear, a.zfar = (cam.vertical_fov_radians.data_ptr(), cam.znear.data_ptr(),
                                           cam.zfar.data_ptr())
        a.rays_from_python = self.config_data.rays_from_python.data_ptr()
        a.grads_enabled = self.meta_data.grads_enabled.data_ptr()
        a.total_num_calls = self.meta_data.total_num_calls.data_ptr()
        a.was_visible = g.was_visible.data_ptr() if g.was_visible.numel() > 0 else None
        a.count = g.was_visible.numel()
        o = np.asarray(origin, dtype=np.float32).reshape(3)
        rot = np.asarray(rotation_c2w, dtype=np.float32).reshape(9)
        a.origin_value[:] = o.tolist()
        a.c2w_value[:] = rot.tolist()
        a.fov_value, a.znear_value, a.zfar_value = float(vertical_fov), float(znear), float(zfar)
        a.grads_enabled_value = 1 if torch.is_grad_enabled() else 0
        L.check(L.LIB.gray_frame_setup(C.byref(a), self._stream()))
        self._frame_prepared = True

    def forward_pass(self) -> None:
        """Reference raytracer.h:88-102."""
        self._require_bvh()
        if self._frame_prepared:
            self._frame_prepared = False
        else:
            self.meta_data.update()
            self.gaussian_data.was_visible.fill_(False)
        save = torch.is_grad_enabled() and not self.inference_only
        if save and self._ellipsoids_on():
            raise RuntimeError("render_ellipsoids=True is only supported for no-grad display renders")
        if not save and self._ellipsoids_on():
            self._sh_forward()
            self._render_ellipsoids()
            self._forward_saved = False
            return
        view = self._view(save)
        tk = self.timer.start("trace_forward")
        self._render(view, save)
        self.timer.stop(tk)
        self._forward_saved = save

    def _render(self, view: L.View, save: bool) -> None:
        """One forward launch; the launch computes the SH colours itself
        (render args sh_params) beside its collect pass."""
        args = self._args(view, save)
        params = self.gaussian_data.params()
        args.sh_params = C.cast(C.pointer(params), C.c_void_p)
        L.check(L.LIB.gray_render_forward(self.ctx.handle, C.byref(args), self._stream()))

    def forward_pass_display(self) -> None:
        """Reference raytracer.h:104-114 (no metadata update, no backward state)."""
        self._require_bvh()
        if self._ellipsoids_on():
            self._sh_forward()
            self._render_ellipsoids()
            return
        self._render(self._view(False), False)

    def _ellipsoids_on(self) -> bool:
        """config.render_ellipsoids (host memory, no synchronisation): the
        reference's display switch (shaders.cu:90-93).  Its Python layer
        rejects it with gradients (gray/raytracer.py:142-144), and so does
        forward_pass here."""
        return bool(self.config_data.render_ellipsoids.item())

    def _render_ellipsoids(self) -> None:
        """Reference ellipsoid_viewer.cu: nearest proxy ellipsoid, flat shaded."""
        view = self._view(False)
        L.check(L.LIB.gray_render_ellipsoids(self.ctx.handle, C.byref(self._args(view, False)),
                                             self.config_data.ellipsoid_min_opacity.data_ptr(), self._stream()))

    def backward_pass(self) -> None:
        """Reference raytracer.h:116-124."""
        if self.inference_only:
            raise RuntimeError("backward_pass is unavailable for an inference-only raytracer")
        if not self._forward_saved:
            raise RuntimeError("backward_pass needs a forward_pass with gradients enabled")
        g = self.gaussian_data
        view = self._view(True)
        grads = L.Grads(g.grad_mean.data_ptr(), g.grad_rotation.data_ptr(), g.grad_scale.data_ptr(),
                        g.grad_opacity.data_ptr(), g.pruning_weight.data_ptr(), g.was_visible.data_ptr())
        params = g.params()
        tk = self.timer.start("trace_backward")
        L.check(L.LIB.gray_render_backward(self.ctx.handle, C.byref(self._args(view, True)), C.byref(params),
                                           C.byref(grads), self._stream()))
        self.timer.stop(tk)
        origins = (C.c_void_p * 1)(self.camera_data.origin.data_ptr())
        rgb = (C.c_void_p * 1)(g.channels.data_ptr())
        grgb = (C.c_void_p * 1)(g.grad_channels.data_ptr())
        tk = self.timer.start("sh_backward")
        L.check(L.LIB.gray_sh_backward(self.ctx.handle, C.byref(params), C.byref(self.config_data.knobs()), 1,
                                       origins, rgb, grgb, g.was_visible.data_ptr(), g.grad_sh_coeffs_dc.data_ptr(),
                                       g.grad_sh_coeffs_rest.data_ptr(), g.grad_mean.data_ptr(), 0, self._stream()))
        self.timer.stop(tk)

    def adam_struct(self, update_pruning_stats: bool) -> L.Adam:
        g, cfg = self.gaussian_data, self.config_data
        a = L.Adam()
        for i, name in enumerate(GaussianDataHolder.PARAMS):
            p = getattr(g, name)
            a.param[i] = p.data_ptr()
            a.grad[i] = getattr(g, "grad_" + name).data_ptr()
            a.m1[i] = getattr(g, "first_moment_" + name).data_ptr()
            a.m2[i] = getattr(g, "second_moment_" + name).data_ptr()
            a.lr[i] = getattr(g, GaussianDataHolder.LRS[i]).data_ptr()
            a.numel[i] = p.numel()
        a.beta1, a.beta2, a.epsilon = g.beta_1.data_ptr(), g.beta_2.data_ptr(), g.epsilon.data_ptr()
        if update_pruning_stats:
            a.was_visible = g.was_visible.data_ptr()
            a.pruning_counter = g.pruning_counter.data_ptr()
        a.count = g.count
        a.update_channels = cfg.update_channels.data_ptr()
        a.zero_grads = cfg.zero_grads.data_ptr()
        a.enable_sh = cfg.enable_sh.data_ptr()
        a.sh_update_laziness = g.sh_update_laziness.data_ptr()
        return a

    def step(self, update_pruning_stats: bool = False) -> None:
        """Reference raytracer.h:126-133 + adam.cu:70-84 in one launch.  With
        update_pruning_stats the counter update of gray/raytracer.py:361-366 is
        fused into the same launch."""
        if self.inference_only:
            raise RuntimeError("step is unavailable for an inference-only raytracer")
        with torch.no_grad():
            self.meta_data.total_num_steps.add_(1)
        a = self.adam_struct(update_pruning_stats)
        # the kernel reads the incremented counter on the device (raytracer.h:127-128)
        a.step_device = self.meta_data.total_num_steps.data_ptr()
        tk = self.timer.start("adam")
        L.check(L.LIB.gray_adam_step(self.ctx.handle, C.byref(a), 0, 0, self._stream()))
        self.timer.stop(tk)

    def update_bvh(self) -> None:
        """Reference raytracer.h:135 (BVHWrapper::update: snapshot + refit)."""
        if self._built_count != self.gaussian_data.count:
            self.rebuild_bvh()
            return
        g = self.gaussian_data
        tk = self.timer.start("bvh_refit")
        L.check(L.LIB.gray_bvh_update(self.ctx.handle, C.byref(g.params()), C.byref(self.config_data.knobs()), 0,
                                      self._stream()))
        self.timer.stop(tk)

    def rebuild_bvh(self) -> None:
        """Reference raytracer.h:137-142 (full LBVH build)."""
        g = self.gaussian_data
        L.check(L.LIB.gray_bvh_update(self.ctx.handle, C.byref(g.params()), C.byref(self.config_data.knobs()), 1,
                                      self._stream()))
        self._built_count = g.count

    def resize(self, new_num_gaussians: int) -> None:
        """Reference raytracer.h:144-149."""
        self.gaussian_data.resize(int(new_num_gaussians))

    def set_render_resolution(self, new_width: int, new_height: int) -> None:
        """Reference raytracer.h:151-158."""
        if new_width < 1 or new_height < 1:
            raise RuntimeError("render resolution must be positive")
        if new_width > self.image_width or new_height > self.image_height:
            raise RuntimeError("render resolution cannot exceed the allocated framebuffer size")
        self.width, self.height = int(new_width), int(new_height)

    # ---- getters (reference raytracer.h:171-182) ---------------------------
    def get_camera(self):
        return self.camera_data

    def get_config(self):
        return self.config_data

    def get_framebuffer(self):
        return self.framebuffer_data

    def get_gaussians(self):
        return self.gaussian_data

    def get_metadata(self):
        return self.meta_data

    def get_stats(self):
        return self.stats_data

    def get_ppll_forward_data(self):
        return self.hits

    def get_ppll_backward_data(self):
        return self.hits

    def get_num_channels(self) -> int:
        return L.NUM_CHANNELS

    def get_max_alpha(self) -> float:
        return L.MAX_ALPHA

    def device_status(self) -> None:
        """Raises if a kernel flagged a device-side error since the last call."""
        self.ctx.status()

    # ---- parity helpers (no reference counterpart) --------------------------
    def debug_snapshot(self):
        """Geometry snapshot of the last BVH update: geo [N,16], aabb [N,6]."""
        n = self.gaussian_data.count
        geo = torch.zeros(n, 16, dtype=F32, device=self.device)
        aabb = torch.zeros(n, 6, dtype=F32, device=self.device)
        L.check(L.LIB.gray_debug_snapshot(self.ctx.handle, geo.data_ptr(), aabb.data_ptr(), self._stream()))
        return geo, aabb

    def debug_bvh(self):
        """LBVH of the last rebuild: nodes [N-1,16] float32 (refs as int bits),
        sorted_ids [N], keys [N]."""
        n = self.gaussian_data.count
        nn = max(n - 1, 1)
        nodes = torch.zeros(nn, 16, dtype=F32, device=self.device)
        ids = torch.zeros(n, dtype=torch.int32, device=self.device)
        keys = torch.zeros(n, dtype=torch.int32, device=self.device)
        L.check(L.LIB.gray_debug_bvh(self.ctx.handle, nodes.data_ptr(), ids.data_ptr(), keys.data_ptr(),
                                     self._stream()))
        return nodes, ids, keys

    def debug_wide(self):
        """4-wide nodes of the collect pass at the binary node index: nodes
        [N-1,16] int32 (half-pair boxes, child refs) and flags [N-1] uint8 (1 =
        a 4-wide node: internal node at even depth)."""
        nn = max(self.gaussian_data.count - 1, 1)
        nodes = torch.zeros(nn, 16, dtype=torch.int32, device=self.device)
        flags = torch.zeros(nn, dtype=torch.uint8, device=self.device)
        L.check(L.LIB.gray_debug_wide(self.ctx.handle, nodes.data_ptr(), flags.data_ptr(), self._stream()))
        return nodes, flags

    def set_device_timing(self, enable: bool) -> None:
        """Record HIP events around BVH rebuilds and the forward's stages."""
        L.check(L.LIB.gray_bvh_set_timing(self.ctx.handle, 1 if enable else 0))

    def last_build_ms(self) -> float:
        """Device milliseconds of the last timed BVH rebuild (-1: none)."""
        return float(L.LIB.gray_bvh_last_build_ms(self.ctx.handle))

    def forward_stage_ms(self) -> dict:
        """Device milliseconds of the last timed forward's stages; waits for it."""
        ms = (C.c_float * 4)()
        L.check(L.LIB.gray_render_stage_ms(self.ctx.handle, ms, 4))
        return {"collect": ms[0], "collect_tasks": ms[1], "composite_lists": ms[2], "retraverse": ms[3]}

    def debug_list_counts(self):
        """Hits per ray of the last forward that collected lists, int32
        [tiles*64] in tile order (lane = 8*(y%8) + x%8; -1 = re-traversed)."""
        n = C.c_int64(0)
        L.check(L.LIB.gray_debug_list_counts(self.ctx.handle, None, 0, C.byref(n), self._stream()))
        out = torch.zeros(max(n.value, 1), dtype=torch.int32, device=self.device)
        L.check(L.LIB.gray_debug_list_counts(self.ctx.handle, out.data_ptr(), out.numel(), C.byref(n),
                                             self._stream()))
        return out[:n.value]

    def debug_hitlists(self, max_hits: int, use_bvh: bool = True):
        """Every hit of every ray of the current camera, sorted by (distance,
        id): ids [R, max_hits], dists [R, max_hits], counts [R], R = w*h."""
        self._require_bvh()
        r = self.width * self.height
        ids = torch.zeros(r, max_hits, dtype=torch.int32, device=self.device)
        dists = torch.zeros(r, max_hits, dtype=F32, device=self.device)
        counts = torch.zeros(r, dtype=torch.int32, device=self.device)
        view = self._view(False)
        args = self._args(view, False)
        args.use_bvh = 1 if use_bvh else 0
        L.check(L.LIB.gray_debug_hitlists(self.ctx.handle, C.byref(args), max_hits, ids.data_ptr(),
                                          dists.data_ptr(), counts.data_ptr(), self._stream()))
        return ids, dists, counts
