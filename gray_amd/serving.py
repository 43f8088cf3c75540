"""Serving loop: one scene resident in HBM, camera requests of mixed
resolution rendered with continuous batching.

`SceneServer` keeps the Gaussians on the device, with one geometry snapshot
and one BVH (the scene's library context), and owns `slots` execution
slots.  Each slot is a library context that renders the shared scene
(gray_ctx_share_scene) with its own hit-list and task buffers, reserved up
front, and its own stream.  The host loop:
  * moves arrived requests to a FIFO queue;
  * retires slots whose batch finished (HIP event query) and stamps their
    requests with the device completion time;
  * admits queued requests into a free slot as one batch: at most 8 views
    (one launch) and at most the slot's reserved lanes, the HBM budget given
    to the server divided over its slots; the batch (SH colours of every view
    + one forward launch of all views, each at its own resolution) is queued
    on the slot's stream and the loop goes on, so batches of different slots
    run on the device together while new requests keep arriving.
Plain mode is the comparison: one request per launch, in arrival order, on a
single slot.  `submit`/`poll` drive the same admission without a clock of
their own, for request sources that arrive over time (the remote viewer).

A request's display mode is the reference viewer's (view.py:153-190):
"Gaussians" (the image), "Depth" (the image and the depth output) or
"Ellipsoids" (the ellipsoid display, its own launch).  Requests of one batch
share the launch kind, and ellipsoid batches their minimum opacity.

The per-request work is the reference's forward_pass_display
(cuda/raytracer.h:104-114) for that camera.  Only the request's own view
influences its image, so a request renders the same alone and in any batch.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .camera import CameraInfo
from .functional import RenderSettings

F32 = torch.float32
RESOLUTIONS = ((256, 256), (512, 384), (1280, 720), (1600, 900))
MAX_VIEWS = 8
TASK_BYTES_PER_LANE = 256
CAM_FLOATS = 16  # origin 3, c2w 9, fov, znear, zfar, ellipsoid minimum opacity
MODES = ("Gaussians", "Depth", "Ellipsoids")


def lanes_of(width: int, height: int) -> int:
    return -(-width // 8) * -(-height // 8) * 64


@dataclass
class Request:
    rid: int
    camera: CameraInfo
    arrival: float = 0.0          # seconds after the start of serve()
    done: float = math.nan        # device completion, seconds after the start
    batch_size: int = 0
    image: Optional[torch.Tensor] = field(default=None, repr=False)
    mode: str = "Gaussians"
    znear: Optional[float] = None     # None: the server's setting
    min_opacity: float = 0.05         # ellipsoid display
    depth: Optional[torch.Tensor] = field(default=None, repr=False)  # [h, w, 1] in "Depth" mode
    error: Optional[str] = None       # set by poll() for a request it could not admit
    tag: object = field(default=None, repr=False)  # the caller's

    def __post_init__(self):
        if self.mode not in MODES:
            raise ValueError(f"unknown display mode {self.mode!r} (one of {', '.join(MODES)})")

    @property
    def width(self) -> int:
        return self.camera.image_width

    @property
    def height(self) -> int:
        return self.camera.image_height

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class _Slot:
    def __init__(self, server: "SceneServer", max_lanes: int):
        self.ctx = L.Context()
        # renders read the server's one snapshot and BVH; the slot owns only scratch
        L.check(L.LIB.gray_ctx_share_scene(self.ctx.handle, server.scene.handle))
        self.stream = torch.cuda.Stream()
        self.max_lanes = max_lanes
        self.batch: List[Request] = []
        self.end: Optional[torch.cuda.Event] = None
        self.keep = None  # host structures of the batch in flight
        with torch.cuda.stream(self.stream):
            self.cams = torch.zeros(MAX_VIEWS, CAM_FLOATS, dtype=F32, device=server.device)
            self.rgb = torch.zeros(MAX_VIEWS, server.n, 3, dtype=F32, device=server.device)
        self.cams_host = torch.zeros(MAX_VIEWS, CAM_FLOATS, dtype=F32, pin_memory=True)
        if server.list_capacity > 0:
            L.check(L.LIB.gray_render_reserve(self.ctx.handle, max_lanes, server.list_capacity))
        self.stream.synchronize()
        L.check(L.LIB.gray_device_status(self.ctx.handle, L.stream_ptr(self.stream)))


class SceneServer:
    def __init__(self, mean, rotation, scale, opacity, sh_dc, sh_rest, settings: Optional[RenderSettings] = None,
                 slots: int = 2, budget_bytes: int = 32 << 30, list_capacity: int = 256, device=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        f = lambda t: t.detach().to(self.device, F32).contiguous()  # noqa: E731
        self.tensors = [f(t) for t in (mean, rotation, scale, opacity, sh_dc, sh_rest)]
        self.n = self.tensors[0].shape[0]
        self.settings = s = settings or RenderSettings()
        self.list_capacity = int(list_capacity)
        t = lambda v, dt=F32: torch.tensor([v], dtype=dt, device=self.device)  # noqa: E731
        self._k = [t(s.exp_power), t(s.alpha_threshold), t(s.t_threshold), t(s.global_scale_factor),
                   t(s.eps_forward_normalization), t(s.eps_scale_grad),
                   torch.tensor(list(s.background), dtype=F32, device=self.device), t(False, torch.bool),
                   t(True, torch.bool)]
        self._k_depth = list(self._k)
        self._k_depth[7] = t(True, torch.bool)
        self._deg = torch.tensor([s.sh_degree], dtype=torch.int32, device=self.device)
        p = L.Params()
        p.count, p.num_sh_coeffs, p.sh_is_fp16 = self.n, self.tensors[5].shape[1], 0
        p.current_sh_degree = self._deg.data_ptr()
        p.mean, p.rotation, p.scale, p.opacity, p.sh_dc, p.sh_rest = (x.data_ptr() for x in self.tensors)
        self.params = p
        per_lane = self.list_capacity * 16 + TASK_BYTES_PER_LANE
        self.slot_lanes = int(budget_bytes // max(1, slots) // per_lane)
        self.budget_bytes = budget_bytes
        # the one resident snapshot + BVH of the scene
        self.scene = L.Context()
        self._rebuild()
        self.slots = [_Slot(self, self.slot_lanes) for _ in range(slots)]
        self._queue: List[Request] = []
        self._h0 = time.perf_counter()

    def knobs(self, depth: bool = False) -> L.Knobs:
        return L.Knobs(*(k.data_ptr() for k in (self._k_depth if depth else self._k)))

    def _rebuild(self) -> None:
        """Snapshot + BVH of the scene, on the current stream, finished before
        any slot renders it."""
        stream = torch.cuda.current_stream(self.device)
        L.check(L.LIB.gray_bvh_update(self.scene.handle, C.byref(self.params), C.byref(self.knobs()), 1,
                                      L.stream_ptr(stream)))
        stream.synchronize()
        self.scene.status(stream)

    def set_global_scale_factor(self, value: float) -> None:
        """The viewer's scaling modifier (view.py:171-176, 407-410): waits for
        the batches in flight (they keep their images, made with the old
        factor; poll() still returns them) and rebuilds the scene's BVH, which
        every slot renders."""
        if float(self._k[3].item()) == float(value):
            return
        for sl in self.slots:
            sl.stream.synchronize()
        self._k[3].fill_(float(value))
        self._k_depth[3].fill_(float(value))
        self._rebuild()

    # ---- one batch ---------------------------------------------------------
    def _launch(self, slot: _Slot, reqs: Sequence[Request]) -> None:
        V = len(reqs)
        s = self.settings
        for i, r in enumerate(reqs):
            c = r.camera
            row = np.zeros(CAM_FLOATS, np.float32)
            row[0:3] = c.origin
            row[3:12] = c.rotation_c2w_blender().reshape(-1)
            row[12], row[13], row[14] = c.fov_y, s.znear if r.znear is None else r.znear, s.zfar
            row[15] = r.min_opacity
            slot.cams_host[i] = torch.from_numpy(row)
        views = (L.View * V)()
        outs = []
        with torch.cuda.stream(slot.stream):
            slot.cams[:V].copy_(slot.cams_host[:V], non_blocking=True)
            ellipsoids = reqs[0].mode == "Ellipsoids"
            depth_on = any(r.mode == "Depth" for r in reqs)
            for i, r in enumerate(reqs):
                img = torch.empty(r.height, r.width, 3, dtype=F32, device=self.device)
                outs.append(img)
                r.depth = None
                if r.mode == "Depth":
                    r.depth = torch.zeros(r.height, r.width, 1, dtype=F32, device=self.device)
                v = views[i]
                base = slot.cams[i]
                v.origin = base.data_ptr()
                v.rotation_c2w = base.data_ptr() + 3 * 4
                v.vertical_fov = base.data_ptr() + 12 * 4
                v.znear = base.data_ptr() + 13 * 4
                v.zfar = base.data_ptr() + 14 * 4
                v.width, v.height, v.stride, v.alloc_height = r.width, r.height, r.width, r.height
                v.rgb = slot.rgb[i].data_ptr()
                v.out_channels = img.data_ptr()
                v.out_depth = r.depth.data_ptr() if r.depth is not None else None
            sp = L.stream_ptr(slot.stream)
            knobs = self.knobs(depth_on)
            a = L.RenderArgs()
            a.knobs = knobs
            a.num_views = V
            a.views = C.cast(views, C.POINTER(L.View))
            a.save, a.hit_capacity, a.use_bvh, a.list_capacity = 0, 0, 1, self.list_capacity
            a.fp16_hits = 1 if s.fp16_hits else 0
            if ellipsoids:
                origins = (C.c_void_p * V)(*[views[i].origin for i in range(V)])
                rgbs = (C.c_void_p * V)(*[views[i].rgb for i in range(V)])
                L.check(L.LIB.gray_sh_forward(slot.ctx.handle, C.byref(self.params), C.byref(knobs), V, origins, rgbs,
                                              sp))
                min_op = C.c_void_p(slot.cams[0].data_ptr() + 15 * 4)
                L.check(L.LIB.gray_render_ellipsoids(slot.ctx.handle, C.byref(a), min_op, sp))
            else:
                # the launch computes the colours itself, beside its collect pass
                a.sh_params = C.cast(C.pointer(self.params), C.c_void_p)
                L.check(L.LIB.gray_render_forward(slot.ctx.handle, C.byref(a), sp))
            slot.end = torch.cuda.Event(enable_timing=True)
            slot.end.record(slot.stream)
        for r, img in zip(reqs, outs):
            r.image = img
            r.batch_size = V
        slot.batch = list(reqs)
        slot.keep = views

    def render_alone(self, request: Request) -> torch.Tensor:
        """The request in a batch of its own (for comparison)."""
        slot = self.slots[0]
        self._launch(slot, [request])
        slot.stream.synchronize()
        L.check(L.LIB.gray_device_status(slot.ctx.handle, L.stream_ptr(slot.stream)))
        slot.batch = []
        return request.image

    # ---- the loop ----------------------------------------------------------
    def _take(self, queue: List[Request], slot: _Slot, batched: bool) -> List[Request]:
        if lanes_of(queue[0].width, queue[0].height) > slot.max_lanes:
            raise ValueError(f"request {queue[0].rid} ({queue[0].width}x{queue[0].height}) exceeds the per-slot "
                             f"budget of {slot.max_lanes} lanes")
        if not batched:
            return [queue.pop(0)]
        head = queue[0]
        kind = lambda r: (r.mode == "Ellipsoids", r.min_opacity if r.mode == "Ellipsoids" else 0.0)  # noqa: E731
        out, lanes = [], 0
        while queue and len(out) < MAX_VIEWS and kind(queue[0]) == kind(head):
            need = lanes_of(queue[0].width, queue[0].height)
            if lanes + need > slot.max_lanes:
                break
            lanes += need
            out.append(queue.pop(0))
        return out

    def submit(self, request: Request) -> None:
        """Queues a request for the next poll(); its arrival is stamped now
        (seconds since the server was made)."""
        request.arrival = time.perf_counter() - self._h0
        request.done = math.nan
        self._queue.append(request)

    def poll(self) -> List[Request]:
        """Retires the slots whose batch finished (requests get .done in the
        same clock as .arrival) and admits queued requests into free slots.
        Returns the finished requests, and those that could not be admitted
        (with .error set).  Does not block."""
        finished: List[Request] = []
        for sl in self.slots:
            if sl.batch and sl.end.query():
                L.check(L.LIB.gray_device_status(sl.ctx.handle, L.stream_ptr(sl.stream)))
                t = time.perf_counter() - self._h0
                for r in sl.batch:
                    r.done = t
                finished.extend(sl.batch)
                sl.batch = []
        for sl in self.slots:
            if not self._queue:
                break
            if sl.batch:
                continue
            try:
                reqs = self._take(self._queue, sl, True)
            except ValueError as e:
                r = self._queue.pop(0)
                r.error = str(e)
                finished.append(r)
                continue
            self._launch(sl, reqs)
        return finished

    @property
    def busy(self) -> bool:
        return bool(self._queue) or any(sl.batch for sl in self.slots)

    def serve(self, requests: Sequence[Request], batched: bool = True) -> dict:
        """Runs the arrival process in real time; returns latency statistics
        (seconds).  Requests get .done, .image and .batch_size."""
        pending = sorted(requests, key=lambda r: r.arrival)
        slots = self.slots if batched else self.slots[:1]
        for sl in slots:
            sl.stream.synchronize()
            sl.batch = []
        queue: List[Request] = []
        done: List[Request] = []
        launches = 0
        torch.cuda.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t0.record()
        h0 = time.perf_counter()
        while len(done) < len(requests):
            now = time.perf_counter() - h0
            while pending and pending[0].arrival <= now:
                queue.append(pending.pop(0))
            progressed = False
            for sl in slots:
                if sl.batch and sl.end.query():
                    t = t0.elapsed_time(sl.end) / 1e3
                    for r in sl.batch:
                        r.done = t
                    done.extend(sl.batch)
                    sl.batch = []
                    progressed = True
            for sl in slots:
                if queue and not sl.batch:
                    self._launch(sl, self._take(queue, sl, batched))
                    launches += 1
                    progressed = True
            if not progressed:
                time.sleep(0.0001)
        for sl in slots:
            L.check(L.LIB.gray_device_status(sl.ctx.handle, L.stream_ptr(sl.stream)))
        lat = np.array(sorted(r.latency for r in requests))
        span = max(r.done for r in requests) - min(r.arrival for r in requests)
        return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_ms": float(lat.mean() * 1e3), "frames_per_s": len(requests) / span, "launches": launches,
                "mean_batch": float(np.mean([r.batch_size for r in requests]))}


def poisson_requests(count: int, rate: float, seed: int, radius=(4.0, 8.0)) -> List[Request]:
    """Seeded requests: exponential inter-arrival times at `rate` per second,
    poses uniform on spheres of radius 4-8 looking at the origin, resolution
    uniform over RESOLUTIONS."""
    from .scenes import sphere_camera

    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        out.append(Request(i, sphere_camera(rng, radius[0], radius[1], w, h, uid=i), arrival=t))
    return out
