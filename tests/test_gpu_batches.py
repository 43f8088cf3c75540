"""Reference batch semantics on the GPU (forward_pass.cu:150-230): batches of
32 hits strictly beyond tmin, served here in sub-rounds of 16.  Exact distance
ties inside a batch are both composited; at a batch boundary the later one is
dropped.  Long lists exercise the re-traversal (recompute) backward.

The 80-hit lists are composited by a whole wave per ray by default and by one
lane per ray in the "lists-lane" case; the two kernels are also compared with
each other on a larger scene (composited hits, saved lists and
transmittances bit for bit)."""
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


def run(mean, q, s, logit, colors, tthr, capacity, list_capacity=512, wave_list_min=0):
    W = H = 8
    nr = NativeRaytracer(W, H, N_AXIS, 0, 1, 1, hit_capacity=capacity)
    # a negative capacity means: lists of that size, walks split into tasks every iteration
    nr.list_capacity = abs(list_capacity)
    nr.wave_list_min = wave_list_min
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


@pytest.mark.parametrize("list_capacity,wave", [(512, 0), (512, -1), (0, 0), (64, 0), (-512, 0)],
                         ids=["lists", "lists-lane", "traverse", "overflow", "tasks"])
@pytest.mark.parametrize("dup_at,tthr,expect_drop", [(None, 0.0, 0), (None, 0.05, 0), (15, 0.0, 0), (31, 0.0, 1),
                                                     (47, 0.0, 0), (63, 0.0, 1)])
def test_batches_and_ties(dup_at, tthr, expect_drop, list_capacity, wave):
    mean, q, s, logit, colors = axis_scene(dup_at)
    nr = run(mean, q, s, logit, colors, tthr, capacity=64, list_capacity=list_capacity, wave_list_min=wave)
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


@pytest.mark.parametrize("fp16,znear,zfar,tthr", [(False, 0.0, 99999.9, 0.03), (True, 0.0, 99999.9, 0.03),
                                                  (False, 2.8, 3.3, 0.0), (False, 0.0, 99999.9, 0.0)])
def test_wave_matches_lane(fp16, znear, zfar, tthr):
    """Every list of two hits or more composited by a whole wave against the
    per-lane kernel on the same snapshot: composited counts, saved lists,
    transmittances, full transmittance and depth bit for bit; colour and
    truncation estimate (skip-set sums in another order) within 1e-6."""
    from gray_amd import scenes
    from gpu_util import setup_native
    scene = scenes.tiny_scene(20000, seed=5)
    cam = scenes.tiny_camera(64, 48)
    out = {}
    for wave in (-1, 1):
        nr = setup_native(scene, cam, t_threshold=tthr, znear=znear, zfar=zfar, hit_capacity=64,
                          wave_list_min=wave)
        nr.fp16_hits = fp16
        nr.get_config().render_depth.fill_(True)
        torch.set_grad_enabled(True)
        nr.forward_pass()
        torch.cuda.synchronize()
        nr.device_status()
        fb, h = nr.get_framebuffer(), nr.hits
        out[wave] = dict(C=fb.output_channels.cpu().numpy(), T=fb.transmittance.cpu().numpy(),
                         F=fb.full_transmittance.cpu().numpy(), D=fb.output_depth.cpu().numpy(),
                         E=fb.remaining_channels_estimate.cpu().numpy(),
                         n=h.n_comp.cpu().numpy(), ids=h.hit_ids.cpu().numpy(), Th=h.hit_trans.cpu().numpy(),
                         lists=nr.debug_list_counts().cpu().numpy())
    a, b = out[-1], out[1]
    assert (b["lists"] >= 2).sum() > 1000, "the wave kernel must see many rays"
    # the per-lane kernel flags composited entries past the first 128 in the
    # list records instead of its register mask: both kinds must occur
    assert b["lists"].max() > 128, b["lists"].max()
    assert np.array_equal(a["n"], b["n"])
    assert a["n"].max() > 16, "lists must span several rounds"
    for slot in range(min(64, int(a["n"].max()))):
        live = a["n"] > slot
        assert np.array_equal(a["ids"][slot][live], b["ids"][slot][live]), slot
        assert np.array_equal(a["Th"][slot][live].view(np.uint32), b["Th"][slot][live].view(np.uint32)), slot
    assert np.array_equal(a["T"].view(np.uint32), b["T"].view(np.uint32))
    assert np.array_equal(a["D"].view(np.uint32), b["D"].view(np.uint32))
    assert np.abs(a["F"] - b["F"]).max() < 1e-6
    assert np.abs(a["C"] - b["C"]).max() < 1e-6
    assert np.abs(a["E"] - b["E"]).max() < 1e-5
