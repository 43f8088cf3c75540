"""Serving loop on the GPU: every request renders the same alone and inside
mixed-resolution batches, in continuous batching and in plain one-launch-per-
request mode, and the device result equals the single-camera drop-in path."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from gray_amd import scenes
from gray_amd.functional import RenderSettings
from gray_amd.serving import Request, SceneServer, lanes_of

from gpu_util import setup_native

pytestmark = pytest.mark.gpu

SIZES = ((40, 32), (64, 48), (24, 24), (72, 40))


def small_requests(count, seed, spacing=0.0):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        w, h = SIZES[i % len(SIZES)]
        out.append(Request(i, scenes.sphere_camera(rng, 2.5, 4.0, w, h, uid=i), arrival=i * spacing))
    return out


@pytest.fixture(scope="module")
def served():
    scene = scenes.tiny_scene(1024, seed=8)
    t = [getattr(scene, k).cuda() for k in ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")]
    budget = 3 * lanes_of(72, 40) * (64 * 16 + 256) * 2
    server = SceneServer(*t, settings=RenderSettings(t_threshold=0.03), slots=2, budget_bytes=budget,
                         list_capacity=64)
    return scene, server


def test_alone_equals_mixed_batches(served):
    scene, server = served
    reqs = small_requests(12, seed=1)
    alone = [server.render_alone(r).clone() for r in small_requests(12, seed=1)]
    stats = server.serve(reqs, batched=True)
    assert max(r.batch_size for r in reqs) > 1, "requests must have been batched"
    for r, a in zip(reqs, alone):
        assert (r.image - a).abs().max().item() <= 1e-6, r.rid
    plain = small_requests(12, seed=1)
    server.serve(plain, batched=False)
    for r, a in zip(plain, alone):
        assert r.batch_size == 1
        assert (r.image - a).abs().max().item() <= 1e-6, r.rid
    assert stats["p99_ms"] >= stats["p50_ms"] > 0


def test_served_image_equals_drop_in_render(served):
    scene, server = served
    r = small_requests(1, seed=2)[0]
    img = server.render_alone(r)
    nr = setup_native(scene, r.camera, t_threshold=0.03)
    nr.forward_pass()
    ref = nr.get_framebuffer().output_channels[:r.height, :r.width]
    assert (img - ref).abs().max().item() <= 1e-6


def test_arrivals_over_time(served):
    _, server = served
    reqs = small_requests(40, seed=3, spacing=0.0005)
    stats = server.serve(reqs, batched=True)
    assert all(not math.isnan(r.done) and r.done >= r.arrival for r in reqs)
    assert stats["frames_per_s"] > 0


def _drop_in_display(scene, cam, **fills):
    """The drop-in path's forward_pass_display with config knobs set."""
    nr = setup_native(scene, cam, t_threshold=0.03)
    cfg = nr.get_config()
    for k, v in fills.items():
        getattr(cfg, k).fill_(v)
    nr.rebuild_bvh()
    with torch.no_grad():
        nr.forward_pass_display()
    torch.cuda.synchronize()
    nr.device_status()
    fb = nr.get_framebuffer()
    H, W = cam.image_height, cam.image_width
    return fb.output_channels[:H, :W].clone(), fb.output_depth[:H, :W].clone()


def test_display_modes_match_drop_in(served):
    """Depth and ellipsoid requests (the viewer's display modes) against the
    drop-in path with config.render_depth / render_ellipsoids."""
    scene, server = served
    cam = small_requests(1, seed=5)[0].camera
    ref, ref_depth = _drop_in_display(scene, cam, render_depth=True)
    r = Request(0, cam, mode="Depth")
    img = server.render_alone(r)
    assert (ref_depth > 0).sum().item() > 100
    assert (img - ref).abs().max().item() <= 1e-6
    assert (r.depth - ref_depth).abs().max().item() <= 1e-5
    ref, _ = _drop_in_display(scene, cam, render_ellipsoids=True, ellipsoid_min_opacity=0.3)
    img = server.render_alone(Request(1, cam, mode="Ellipsoids", min_opacity=0.3))
    assert (img - ref).abs().max().item() <= 1e-6
    assert (ref.sum(-1) > 0).sum().item() > 100
    # mixed modes in one serve(): batches split by launch kind, images unchanged
    reqs = [Request(i, c.camera, mode=("Gaussians", "Ellipsoids", "Depth")[i % 3], min_opacity=0.3)
            for i, c in enumerate(small_requests(9, seed=6))]
    alone = [server.render_alone(Request(i, q.camera, mode=q.mode, min_opacity=0.3)).clone()
             for i, q in enumerate(reqs)]
    server.serve(reqs, batched=True)
    for q, a in zip(reqs, alone):
        assert (q.image - a).abs().max().item() <= 1e-6, (q.rid, q.mode)


def test_scale_factor_rebuilds_every_slot(served):
    scene, server = served
    cam = small_requests(1, seed=7)[0].camera
    base = server.render_alone(Request(0, cam)).clone()
    ref, _ = _drop_in_display(scene, cam, global_scale_factor=1.5)
    server.set_global_scale_factor(1.5)
    try:
        for sl in range(len(server.slots)):
            r = Request(sl, cam)
            server._launch(server.slots[sl], [r])
            server.slots[sl].stream.synchronize()
            server.slots[sl].batch = []
            assert (r.image - ref).abs().max().item() <= 1e-6
        assert (ref - base).abs().max().item() > 1e-3
    finally:
        server.set_global_scale_factor(1.0)
    assert (server.render_alone(Request(0, cam)) - base).abs().max().item() <= 1e-6


def test_submit_poll(served):
    _, server = served
    reqs = small_requests(10, seed=9)
    alone = [server.render_alone(r).clone() for r in small_requests(10, seed=9)]
    for r in reqs:
        server.submit(r)
    done = []
    for _ in range(200000):
        done += server.poll()
        if len(done) == len(reqs):
            break
    assert not server.busy and sorted(r.rid for r in done) == list(range(10))
    for r, a in zip(reqs, alone):
        assert r.error is None and r.done >= r.arrival
        assert (r.image - a).abs().max().item() <= 1e-6
    big = Request(99, scenes.sphere_camera(np.random.default_rng(0), 2.5, 4.0, 4000, 4000))
    server.submit(big)
    out = server.poll()
    assert out == [big] and "exceeds" in big.error


def test_slots_share_one_scene():
    """Every slot renders the server's one snapshot and BVH: a further slot
    adds only its render scratch to device memory, far less than the scene's
    geometry, and it sees that geometry."""
    from gray_amd import _lib as L
    from gray_amd.serving import _Slot

    scene = scenes.surface_scene(300_000, 20, -4.5, 0.4, seed=2)
    t = [getattr(scene, k).cuda() for k in ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")]
    server = SceneServer(*t, settings=RenderSettings(t_threshold=0.03), slots=1, budget_bytes=1 << 30,
                         list_capacity=0)
    torch.cuda.synchronize()
    free0, reserved0 = torch.cuda.mem_get_info()[0], torch.cuda.memory_reserved()
    extra = _Slot(server, server.slot_lanes)
    torch.cuda.synchronize()
    free1, reserved1 = torch.cuda.mem_get_info()[0], torch.cuda.memory_reserved()
    library_bytes = (free0 - free1) - (reserved1 - reserved0)  # outside torch's allocator
    scene_bytes = 300_000 * (64 + 48 + 24 + 64 + 64)  # snapshot, L2W, AABB, binary and 4-wide nodes
    assert library_bytes < scene_bytes // 4, (library_bytes, scene_bytes)
    n = C.c_int64()
    L.check(L.LIB.gray_bvh_info(extra.ctx.handle, C.byref(n), None, None))
    assert n.value == 300_000
    # the slot's geometry is the scene's, and a slot cannot rebuild it
    geo_scene = torch.zeros(300_000, 16, device="cuda")
    geo_slot = torch.zeros(300_000, 16, device="cuda")
    L.check(L.LIB.gray_debug_snapshot(server.scene.handle, geo_scene.data_ptr(), None, L.stream_ptr()))
    L.check(L.LIB.gray_debug_snapshot(extra.ctx.handle, geo_slot.data_ptr(), None, L.stream_ptr()))
    torch.cuda.synchronize()
    assert torch.equal(geo_scene, geo_slot)
    rc = L.LIB.gray_bvh_update(extra.ctx.handle, C.byref(server.params), C.byref(server.knobs()), 1,
                               L.stream_ptr())
    assert rc == L.GRAY_ERR_STATE
