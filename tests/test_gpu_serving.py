"""Serving loop on the GPU: every request renders the same alone and inside
mixed-resolution batches, in continuous batching and in plain one-launch-per-
request mode, and the device result equals the single-camera drop-in path."""
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
