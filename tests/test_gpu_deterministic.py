"""Deterministic gradient reduction (gray_render_args.deterministic) on the GPU.

With the flag set the forward walks each ray's tree in one lane and the
backward writes one record per composited hit, sorts the records by Gaussian
with a stable radix sort and sums each Gaussian's terms in pixel and hit
order.  Checked here:

  * two forward + backward runs on the same inputs give bit-identical images
    and gradients (stored lists, re-traversed rays and a mix of both; one
    view and a batch of four views);
  * the gradients equal the oracle's per element (the bound of the other
    parity tests: |a - b| <= 1e-3 |b| + 1e-6 max|b|) and the default
    (atomic) mode's;
  * the image equals the default mode's within 1e-6 (the same composited
    hits; only the truncation-estimate sums may be added in another order).
"""
import math

import numpy as np
import pytest
import torch

from gray_amd import scenes
from gray_amd.functional import MultiViewRenderer, RenderSettings
from oracle import oracle as O

from gpu_util import (assert_grads_close, comparable_gaussians, oracle_knobs, per_element_err, rays_of,
                      setup_native)

pytestmark = pytest.mark.gpu

GRADS = ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels", "grad_sh_coeffs_dc",
         "grad_sh_coeffs_rest", "pruning_weight", "was_visible")


def run(scene, cam, deterministic, hit_capacity=None, list_capacity=None, t_threshold=0.03, seed=1):
    """One forward + backward of `scene` seen from `cam`; returns the image
    and every gradient tensor (CPU copies) and the raytracer."""
    nr = setup_native(scene, cam, t_threshold=t_threshold, hit_capacity=hit_capacity, list_capacity=list_capacity)
    nr.deterministic = deterministic
    torch.set_grad_enabled(True)
    nr.forward_pass()
    fb = nr.get_framebuffer()
    h, w = nr.height, nr.width
    gen = torch.Generator().manual_seed(seed)
    fb.grad_output_channels.zero_()
    fb.grad_output_channels[:h, :w].copy_(torch.randn(h, w, 3, generator=gen).cuda() / (h * w))
    g = nr.get_gaussians()
    with torch.no_grad():
        for name in GRADS:
            getattr(g, name).zero_()
    nr.backward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    out = {name: getattr(g, name).cpu().clone() for name in GRADS}
    out["image"] = fb.output_channels[:h, :w].cpu().clone()
    return out, nr


# stored lists (default capacity), every ray re-traversed (0), and a mix (3)
@pytest.mark.parametrize("cap", [None, 0, 3], ids=["stored", "recompute", "mixed"])
def test_repeat_runs_bit_identical(cap):
    scene = scenes.surface_scene(200_000, 30, -4.5, 0.4, seed=2)
    cam = scenes.ring_camera(5.0, 0.7, 320, 240, 60.0)
    a, _ = run(scene, cam, True, hit_capacity=cap, t_threshold=0.01)
    b, nr = run(scene, cam, True, hit_capacity=cap, t_threshold=0.01)
    for key in a:
        assert torch.equal(a[key], b[key]), key
    assert a["grad_opacity"].abs().max() > 0 and a["image"].abs().max() > 0.05
    n = nr.hits.n_comp[: 320 * 240].cpu()
    if cap == 3:
        assert (n > 3).any() and ((n > 0) & (n <= 3)).any(), "both backward paths must run"
    # against the default (atomic) reduction
    c, _ = run(scene, cam, False, hit_capacity=cap, t_threshold=0.01)
    assert (a["image"] - c["image"]).abs().max().item() <= 1e-6
    assert torch.equal(a["was_visible"], c["was_visible"])
    for key in ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels", "pruning_weight"):
        err = per_element_err(a[key].numpy(), c[key].numpy())
        assert err <= 1e-3, (key, err)


@pytest.mark.parametrize("cap", [None, 3], ids=["stored", "mixed"])
def test_deterministic_backward_matches_oracle(cap):
    scene = scenes.tiny_scene(1024, seed=0)
    cam = scenes.tiny_camera(64, 64)
    out, nr = run(scene, cam, True, hit_capacity=cap)
    g = nr.get_gaussians()
    geo, _ = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    _, l2w = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, 0.01, 2.0)
    ro, rd = rays_of(nr)
    rgb = g.channels.cpu().numpy()
    k = oracle_knobs(nr)
    fwd = O.render(geo, rgb, ro, rd, k)
    assert np.abs(out["image"].reshape(-1, 3).numpy() - fwd["C"]).max() < 1e-5
    gout = nr.get_framebuffer().grad_output_channels[:64, :64].reshape(-1, 3).cpu().numpy()
    ob = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, fwd, gout, k, 64 * 64)
    ok = comparable_gaussians(nr.hits.n_comp[: 64 * 64].cpu().numpy(), fwd, geo, ro, rd)
    assert_grads_close([
        ("channels", out["grad_channels"], ob["rgb"]), ("opacity", out["grad_opacity"], ob["opacity"]),
        ("scale", out["grad_scale"], ob["scale"]), ("rotation", out["grad_rotation"], ob["rotation"]),
        ("pruning", out["pruning_weight"], ob["pruning"]),
    ], ok)


def test_batched_views_bit_identical():
    scene = scenes.surface_scene(100_000, 20, -4.5, 0.4, seed=5)
    cams = [scenes.ring_camera(4.0, 2 * math.pi * i / 4 + 0.2, 160, 120, 60.0, uid=i) for i in range(4)]
    runs = []
    for det in (True, True, False):
        params = [getattr(scene, k).cuda().clone().requires_grad_() for k in
                  ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")]
        r = MultiViewRenderer(160, 120, RenderSettings(t_threshold=0.01, deterministic=det))
        img = r.render(cams, *params, rebuild=True)
        w = torch.randn(img.shape, generator=torch.Generator().manual_seed(3)).cuda()
        (img * w).sum().backward()
        torch.cuda.synchronize()
        r.device_status()
        runs.append([img.detach().cpu()] + [p.grad.cpu() for p in params])
    for x, y in zip(runs[0], runs[1]):
        assert torch.equal(x, y)
    assert (runs[0][0] - runs[2][0]).abs().max().item() <= 1e-6
    for x, y in zip(runs[0][1:], runs[2][1:]):
        assert per_element_err(x.numpy(), y.numpy()) <= 1e-3
