"""Reference batch semantics on the GPU (forward_pass.cu:150-230): batches of
32 hits strictly beyond tmin, served here in sub-rounds of 16.  Exact distance
ties inside a batch are both composited; at a batch boundary the later one is
dropped.  Long lists exercise the re-traversal (recompute) backward."""
import numpy as np
import pytest
import torch

from gray_amd.native import NativeRaytracer
from oracle import oracle as O

from gpu_util import rel_err

pytestmark = pytest.mark.gpu

N_AXIS = 80


def axis_scene(dup_at=None):
    z = -np.linspace(0.1, 4.0, N_AXIS)
    if dup_at is not None:
        z[dup_at + 1] = z[dup_at]  # identical Gaussians -> identical distances
    mean = np.stack([np.zeros(N_AXIS), np.zeros(N_AXIS), z], 1)
    q = np.tile([1.0, 0.0, 0.0, 0.0], (N_AXIS, 1))
    s = np.full((N_AXIS, 3), np.log(0.05))
    logit = np.full((N_AXIS, 1), -2.5)
    colors = np.random.default_rng(0).uniform(size=(N_AXIS, 3))
    return mean, q, s, logit, colors


def run(mean, q, s, logit, colors, tthr, capacity, list_capacity=512):
    W = H = 8
    nr = NativeRaytracer(W, H, N_AXIS, 0, 1, 1, hit_capacity=capacity)
    # a negative capacity means: lists of that size, walks split into tasks every iteration
    nr.list_capacity = abs(list_capacity)
    nr.collect_budget = 1 if list_capacity < 0 else 0
    g = nr.get_gaussians()
    cfg = nr.get_config()
    cfg.t_threshold.fill_(tthr)
    cfg.enable_sh.fill_(False)
    with torch.no_grad():
        for name, v in (("mean", mean), ("rotation", q), ("scale", s), ("opacity", logit), ("channels", colors)):
            getattr(g, name).copy_(torch.tensor(v, dtype=torch.float32))
    fb = nr.get_framebuffer()
    # explicit rays (rays_from_python): all near the -z axis from the origin
    off = (np.arange(W * H, dtype=np.float32).reshape(H, W, 1) - 32) * 1e-4
    fb.ray_origin.copy_(torch.tensor(np.concatenate([off, -off, np.zeros_like(off)], 2)))
    d = np.zeros((H, W, 3), np.float32)
    d[..., 2] = -1.0
    fb.ray_direction.copy_(torch.tensor(d))
    cfg.rays_from_python.fill_(True)
    cam = nr.get_camera()
    cam.zfar.fill_(99999.9)
    nr.rebuild_bvh()
    torch.set_grad_enabled(True)
    nr.forward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    return nr


@pytest.mark.parametrize("list_capacity", [512, 0, 64, -512], ids=["lists", "traverse", "overflow", "tasks"])
@pytest.mark.parametrize("dup_at,tthr,expect_drop", [(None, 0.0, 0), (None, 0.05, 0), (15, 0.0, 0), (31, 0.0, 1),
                                                     (47, 0.0, 0), (63, 0.0, 1)])
def test_batches_and_ties(dup_at, tthr, expect_drop, list_capacity):
    mean, q, s, logit, colors = axis_scene(dup_at)
    nr = run(mean, q, s, logit, colors, tthr, capacity=64, list_capacity=list_capacity)
    fb = nr.get_framebuffer()
    ro = fb.ray_origin.reshape(-1, 3).cpu().numpy()
    rd = fb.ray_direction.reshape(-1, 3).cpu().numpy()
    geo, _ = nr.debug_snapshot()
    k = O.make_knobs(t_threshold=tthr)
    ref = O.render(geo.cpu().numpy(), colors.astype(np.float32), ro, rd, k, cap=128)
    img = fb.output_channels.reshape(-1, 3).cpu().numpy()
    ncomp = nr.hits.n_comp.cpu().numpy()
    assert np.array_equal(ref["nhits"], np.full(64, N_AXIS))
    assert np.array_equal(ncomp, ref["ncomp"])
    if tthr == 0.0:
        assert np.all(ncomp == N_AXIS - expect_drop)
    assert np.abs(img - ref["C"]).max() < 1e-5
    assert np.abs(fb.full_transmittance.reshape(-1).cpu().numpy() - ref["F"]).max() < 1e-6


def test_recompute_backward_long_lists():
    mean, q, s, logit, colors = axis_scene(31)
    grads = []
    for cap in (128, 10, 0):
        nr = run(mean, q, s, logit, colors, 0.0, capacity=cap)
        fb = nr.get_framebuffer()
        fb.grad_output_channels.copy_(torch.randn(fb.grad_output_channels.shape,
                                                  generator=torch.Generator().manual_seed(3)))
        g = nr.get_gaussians()
        nr.backward_pass()
        torch.cuda.synchronize()
        nr.device_status()
        grads.append({k: getattr(g, k).cpu().numpy().copy() for k in
                      ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels")})
    for other in grads[1:]:
        for key in grads[0]:
            assert rel_err(other[key], grads[0][key]) < 1e-3, key
