"""Forward and backward knobs of the reference config on the GPU against the
oracle: kernel exponent p (exp_power 1, 2, 3), alpha threshold, background,
depth output (render_depth, depth normalised by 1 - T), and the znear/zfar
interval, in the collected-lists forward (per-lane compositing, and "wave":
every list of two hits or more composited by a whole wave) and the
re-traversal forward."""
import numpy as np
import pytest
import torch

from gray_amd import scenes
from oracle import oracle as O

from gpu_util import assert_grads_close, comparable_gaussians, rays_of, setup_native

pytestmark = pytest.mark.gpu

CASES = [
    dict(exp_power=1.0, alpha_threshold=0.01, bg=(0.0, 0.0, 0.0)),
    dict(exp_power=3.0, alpha_threshold=0.05, bg=(0.2, 0.5, 0.9)),
    dict(exp_power=2.0, alpha_threshold=0.2, bg=(1.0, 1.0, 1.0)),
    dict(exp_power=2.0, alpha_threshold=0.01, bg=(0.1, 0.1, 0.1), znear=2.6, zfar=3.4),
]


@pytest.mark.parametrize("lists,wave", [(None, -1), (0, 0), (None, 1)], ids=["lists", "traverse", "wave"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_knobs_forward_backward(case, lists, wave):
    kw = dict(CASES[case])
    znear, zfar = kw.pop("znear", 0.0), kw.pop("zfar", 99999.9)
    scene = scenes.tiny_scene(1024, seed=20 + case)
    cam = scenes.tiny_camera(40, 32)
    nr = setup_native(scene, cam, t_threshold=0.02, znear=znear, zfar=zfar, list_capacity=lists,
                      wave_list_min=wave, **kw)
    nr.get_config().render_depth.fill_(True)
    torch.set_grad_enabled(True)
    nr.forward_pass()
    fb = nr.get_framebuffer()
    fb.grad_output_channels.copy_(torch.randn(fb.grad_output_channels.shape,
                                              generator=torch.Generator().manual_seed(case)).cuda())
    g = nr.get_gaussians()
    for name in ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels"):
        getattr(g, name).zero_()
    nr.backward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    geo, _ = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    ro, rd = rays_of(nr)
    rgb = g.channels.cpu().numpy()
    k = O.make_knobs(t_threshold=0.02, znear=znear, zfar=zfar, render_depth=True, **kw)
    ref = O.render(geo, rgb, ro, rd, k)
    img = fb.output_channels[:32, :40].reshape(-1, 3).cpu().numpy()
    depth = fb.output_depth[:32, :40].reshape(-1).cpu().numpy()
    assert np.abs(img - ref["C"]).max() < 1e-5
    assert np.abs(depth - ref["D"]).max() < 1e-4 * max(1.0, np.abs(ref["D"]).max())
    if znear > 0.0:
        k_open = O.make_knobs(t_threshold=0.02, render_depth=True, **kw)
        assert ref["ncomp"].sum() < O.render(geo, rgb, ro, rd, k_open)["ncomp"].sum(), "the interval must cut hits"
        assert ref["D"].max() <= zfar
    _, l2w = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, kw["alpha_threshold"],
                        kw["exp_power"])
    gout = fb.grad_output_channels[:32, :40].reshape(-1, 3).cpu().numpy()
    ob = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, ref, gout, k, 40 * 32)
    ok = comparable_gaussians(nr.hits.n_comp[:40 * 32].cpu().numpy(), ref, geo, ro, rd, kw["alpha_threshold"])
    assert_grads_close([(name, getattr(g, name).cpu(), ob[key]) for name, key in (
        ("grad_opacity", "opacity"), ("grad_scale", "scale"), ("grad_rotation", "rotation"),
        ("grad_channels", "rgb"))], ok)


@pytest.mark.parametrize("lists,hit_cap,wave", [(None, None, -1), (0, None, 0), (3, None, 0), (None, 3, -1),
                                                (None, None, 1), (None, 3, 1)],
                         ids=["lists", "traverse", "mixed", "recompute-backward", "wave", "wave-recompute-backward"])
def test_fp16_hits_forward_backward(lists, hit_cap, wave):
    """The reference's fp16 hit storage: distances and compositing alphas
    rounded to fp16, full transmittance from fp32 alphas."""
    scene = scenes.tiny_scene(1024, seed=31)
    cam = scenes.tiny_camera(48, 40)
    bg = (0.2, 0.3, 0.1)
    imgs = {}
    for fp16 in (False, True):
        nr = setup_native(scene, cam, t_threshold=0.02, bg=bg, list_capacity=lists, hit_capacity=hit_cap,
                          wave_list_min=wave)
        nr.fp16_hits = fp16
        nr.get_config().render_depth.fill_(True)
        torch.set_grad_enabled(True)
        nr.forward_pass()
        fb = nr.get_framebuffer()
        imgs[fp16] = fb.output_channels[:40, :48].reshape(-1, 3).cpu().numpy()
    fb.grad_output_channels.copy_(torch.randn(fb.grad_output_channels.shape,
                                              generator=torch.Generator().manual_seed(3)).cuda())
    g = nr.get_gaussians()
    for name in ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels"):
        getattr(g, name).zero_()
    nr.backward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    assert np.abs(imgs[True] - imgs[False]).max() > 1e-6, "the mode must change the image"
    geo, _ = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    ro, rd = rays_of(nr)
    rgb = g.channels.cpu().numpy()
    k = O.make_knobs(t_threshold=0.02, bg=bg, render_depth=True, fp16_hits=True)
    ref = O.render(geo, rgb, ro, rd, k)
    assert np.abs(imgs[True] - ref["C"]).max() < 1e-5
    depth = fb.output_depth[:40, :48].reshape(-1).cpu().numpy()
    assert np.abs(depth - ref["D"]).max() < 1e-4 * max(1.0, np.abs(ref["D"]).max())
    T = fb.transmittance[:40, :48].reshape(-1).cpu().numpy()
    F = fb.full_transmittance[:40, :48].reshape(-1).cpu().numpy()
    assert np.abs(T - ref["T"]).max() < 1e-5 and np.abs(F - ref["F"]).max() < 1e-5
    _, l2w = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, 0.01, 2.0)
    gout = fb.grad_output_channels[:40, :48].reshape(-1, 3).cpu().numpy()
    ob = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, ref, gout, k, 48 * 40)
    ok = comparable_gaussians(nr.hits.n_comp[:48 * 40].cpu().numpy(), ref, geo, ro, rd)
    assert_grads_close([(name, getattr(g, name).cpu(), ob[key]) for name, key in (
        ("grad_opacity", "opacity"), ("grad_scale", "scale"), ("grad_rotation", "rotation"),
        ("grad_channels", "rgb"))], ok)


@pytest.mark.parametrize("use_bvh", [True, False])
def test_ellipsoid_mode_matches_oracle(use_bvh):
    """config.render_ellipsoids (reference ellipsoid_viewer.cu) through
    forward_pass and forward_pass_display without gradients."""
    scene = scenes.tiny_scene(1024, seed=41)
    cam = scenes.tiny_camera(48, 40)
    bg = (0.1, 0.2, 0.3)
    nr = setup_native(scene, cam, bg=bg, znear=1.0, zfar=4.0)
    nr.use_bvh = use_bvh
    torch.set_grad_enabled(True)
    nr.forward_pass()  # writes the rays used below
    normal = nr.get_framebuffer().output_channels[:40, :48].clone()
    nr.backward_pass()
    ro, rd = rays_of(nr)
    cfg = nr.get_config()
    cfg.render_ellipsoids.fill_(True)
    cfg.ellipsoid_min_opacity.fill_(0.3)
    cfg.render_depth.fill_(True)
    geo, _ = nr.debug_snapshot()
    rgb = nr.get_gaussians().channels.cpu().numpy()
    C, D = O.ellipsoids(geo.cpu().numpy(), rgb, ro, rd, 0.3, 1.0, 4.0, bg)
    assert (D > 0).sum() > 200 and (D == 0).sum() > 50
    for display in (False, True):
        fb = nr.get_framebuffer()
        fb.output_channels.zero_()
        fb.output_depth.zero_()
        with torch.no_grad():
            (nr.forward_pass_display if display else nr.forward_pass)()
        torch.cuda.synchronize()
        nr.device_status()
        img = fb.output_channels[:40, :48].reshape(-1, 3).cpu().numpy()
        depth = fb.output_depth[:40, :48].reshape(-1).cpu().numpy()
        assert np.abs(img - C).max() < 1e-5
        assert np.abs(depth - D).max() < 1e-5 * 4.0
    cfg.render_ellipsoids.fill_(False)
    with torch.no_grad():
        nr.forward_pass()
    torch.cuda.synchronize()
    assert torch.allclose(nr.get_framebuffer().output_channels[:40, :48], normal, atol=1e-6)


def test_jittered_primary_rays_match_oracle():
    """config.jitter_primary_rays: the rays follow the reference jitter
    stream seeded by (pixel, metadata total_num_calls)."""
    scene = scenes.tiny_scene(1024, seed=51)
    cam = scenes.tiny_camera(48, 40)
    nr = setup_native(scene, cam)
    nr.get_config().jitter_primary_rays.fill_(True)
    torch.set_grad_enabled(True)
    dirs = []
    for _ in range(2):
        nr.forward_pass()
        torch.cuda.synchronize()
        calls = int(nr.get_metadata().total_num_calls.item())
        ro, rd = rays_of(nr)
        oro, ord_ = O.pinhole_rays_jittered(48, 40, cam.fov_y, cam.origin.astype(np.float32),
                                            cam.rotation_c2w_blender(), calls)
        assert np.abs(ro - oro).max() == 0.0
        assert np.abs(rd - ord_).max() < 2e-6
        dirs.append(rd)
        geo, _ = nr.debug_snapshot()
        ref = O.render(geo.cpu().numpy(), nr.get_gaussians().channels.cpu().numpy(), ro, rd,
                       O.make_knobs(t_threshold=0.03))
        img = nr.get_framebuffer().output_channels[:40, :48].reshape(-1, 3).cpu().numpy()
        assert np.abs(img - ref["C"]).max() < 1e-5
        nr.backward_pass()
    assert np.abs(dirs[0] - dirs[1]).max() > 1e-4, "a new call draws new offsets"


def test_framebuffer_write_rules_without_gradients():
    """Reference Framebuffer::write (framebuffer.h:33-52): without gradients
    only the colour, depth when render_depth or needs_ray_output is set, and
    the ray direction when needs_ray_output is set."""
    scene = scenes.tiny_scene(1024, seed=61)
    cam = scenes.tiny_camera(48, 40)
    nr = setup_native(scene, cam)
    torch.set_grad_enabled(True)
    nr.forward_pass()
    fb = nr.get_framebuffer()
    rd = fb.ray_direction.clone()
    nr.backward_pass()
    cfg = nr.get_config()
    for render_depth, ray_out in ((False, False), (True, False), (False, True)):
        cfg.render_depth.fill_(render_depth)
        cfg.needs_ray_output.fill_(ray_out)
        fb.output_depth.fill_(-7.0)
        fb.ray_direction.fill_(-7.0)
        fb.transmittance.fill_(-7.0)
        with torch.no_grad():
            nr.forward_pass()
        torch.cuda.synchronize()
        depth_written = bool((fb.output_depth[:40, :48] != -7.0).all())
        assert depth_written == (render_depth or ray_out)
        if ray_out:
            assert torch.equal(fb.ray_direction[:40, :48], rd[:40, :48])
        else:
            assert bool((fb.ray_direction == -7.0).all())
        assert bool((fb.transmittance == -7.0).all())
