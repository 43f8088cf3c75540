"""The Python API entry points the reference's callers use, on the GPU,
through the drop-in Raytracer:

  * set_render_resolution (gray/raytracer.py:199-203, raytracer.h:151-158),
    which train.py:237-241 calls every iteration: an active window smaller
    than the allocation, rays from the launch dimensions, pixel ids by the
    allocated stride, and the pruning weight normalised by the launch
    dimensions (backward_pass.cu:129);
  * save_safetensors / from_safetensors (gray/raytracer.py:368-438), with the
    fp16 SH of an inference-only load (gaussians.h:132);
  * from_point_cloud (gray/raytracer.py:205-294), with and without binning.

Images are checked against the oracle on the GPU's own snapshot and rays
(1e-5), gradients element by element (1e-3 relative, see per_element_err)."""
import os

import numpy as np
import pytest
import torch

from gray_amd import Raytracer, RaytracerConfig, scenes
from gray_amd.raytracer import SH_C0, BasicPointCloud
from oracle import oracle as O

from gpu_util import oracle_knobs, per_element_err, rel_err

pytestmark = pytest.mark.gpu

GRAD_NAMES = (("grad_channels", "rgb"), ("grad_opacity", "opacity"), ("grad_scale", "scale"),
              ("grad_rotation", "rotation"))


def load_scene(rt, scene, degree=3):
    g = rt.cuda_module.get_gaussians()
    with torch.no_grad():
        g.mean.copy_(scene.mean)
        g.rotation.copy_(scene.rotation)
        g.scale.copy_(scene.scale)
        g.opacity.copy_(scene.opacity)
        g.sh_coeffs_dc.copy_(scene.sh_dc)
        g.sh_coeffs_rest.copy_(scene.sh_rest)
        g.current_sh_degree.fill_(degree)
    rt.cuda_module.rebuild_bvh()


def active_rays(rt):
    fb = rt.cuda_module.get_framebuffer()
    h, w = rt.render_height, rt.render_width
    return (fb.ray_origin[:h, :w].reshape(-1, 3).cpu().numpy(), fb.ray_direction[:h, :w].reshape(-1, 3).cpu().numpy())


def zero_grads(g):
    for name in ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels", "grad_sh_coeffs_dc",
                 "grad_sh_coeffs_rest", "pruning_weight"):
        getattr(g, name).zero_()


def test_reduced_resolution_matches_oracle():
    """Half-resolution render inside the full allocation: image, rays,
    gradients and the pruning weight against the oracle at the launch
    dimensions; the framebuffer outside the window is left alone; the image
    equals that of a raytracer allocated at the reduced size."""
    scene = scenes.tiny_scene(1024, seed=4)
    W, H = 96, 80
    w, h = W // 2, H // 2
    cfg = RaytracerConfig()
    rt = Raytracer(cfg, scene.count, W, H)
    load_scene(rt, scene)
    fb = rt.cuda_module.get_framebuffer()
    with torch.no_grad():
        fb.output_channels.fill_(-7.0)
    rt.set_render_resolution(w, h)
    cam = scenes.tiny_camera(w, h)
    render = rt(cam)
    assert tuple(render.shape) == (3, h, w)
    torch.cuda.synchronize()
    rt.cuda_module.device_status()
    # outside the active window nothing is written
    assert bool((fb.output_channels[h:, :] == -7.0).all()) and bool((fb.output_channels[:, w:] == -7.0).all())
    # rays from the launch dimensions (aspect w/h), at pixel ids y*W + x
    ro, rd = active_rays(rt)
    oro, ord_ = O.pinhole_rays(w, h, cam.fov_y, cam.origin.astype(np.float32), cam.rotation_c2w_blender())
    assert np.abs(ro - oro).max() == 0.0
    assert np.abs(rd - ord_).max() < 2e-6
    nr = rt.cuda_module
    geo, _ = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    g = nr.get_gaussians()
    rgb = g.channels.cpu().numpy()
    k = oracle_knobs(nr)
    fwd = O.render(geo, rgb, ro, rd, k)
    img = render.detach().permute(1, 2, 0).reshape(-1, 3).cpu().numpy()
    assert np.abs(img - fwd["C"]).max() < 1e-5
    assert (fwd["ncomp"] > 0).sum() > 200
    # the same camera in a raytracer allocated at the reduced size
    rt2 = Raytracer(cfg, scene.count, w, h)
    load_scene(rt2, scene)
    render2 = rt2(cam)
    # the same composited hits; the truncation-estimate sums may round differently
    # where task rounds appended a ray's hits in another order
    assert (render.detach() - render2.detach()).abs().max().item() <= 1e-6

    target = torch.rand(3, h, w, generator=torch.Generator().manual_seed(9)).cuda()
    zero_grads(g)
    loss = torch.nn.functional.l1_loss(render, target)
    rt.backward(loss)
    torch.cuda.synchronize()
    nr.device_status()
    gout = fb.grad_output_channels[:h, :w].reshape(-1, 3).cpu().numpy()
    assert np.abs(gout).max() > 0 and bool((fb.grad_output_channels[h:, :] == 0).all())
    _, l2w = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, cfg.alpha_threshold, cfg.exp_power)
    ob = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, fwd, gout, k, float(w * h))
    vis = g.was_visible.reshape(-1).cpu().numpy().astype(np.uint8)
    _, _, gmean = O.sh_backward(scene.mean, scene.sh_rest, cam.origin.astype(np.float32), rgb, ob["rgb"], vis, 3,
                                grad_mean=ob["mean"])
    ok = comparable(nr, fwd, geo, ro, rd, w, h)
    assert ok.mean() > 0.99
    for name, key in GRAD_NAMES + (("pruning_weight", "pruning"),):
        a, b = getattr(g, name).cpu().numpy(), ob[key]
        # pruning weight: every composited weight / (launch width * height)
        assert per_element_err(a[ok], b[ok]) < 1e-3, name
        assert rel_err(a, b) < 1e-3, name
    assert per_element_err(g.grad_mean.cpu().numpy()[ok], gmean[ok]) < 1e-3
    ob_full = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, fwd, gout, k,
                         float(W * H))
    assert rel_err(g.pruning_weight.cpu(), ob_full["pruning"]) > 1.0  # normalised by the window, not the allocation


def comparable(nr, fwd, geo, ro, rd, w, h):
    """Gaussians whose gradients compare element by element: those hit by no
    ray whose composited count differs between the GPU and the oracle.  The
    stop test T < t_threshold compares alphas from two exp implementations,
    so a ray can composite one hit more or less (a handful per frame); the
    Gaussians such a ray hits are held to the frame-level bound only."""
    n = geo.shape[0]
    # n_comp is stored at pixel ids y*stride + x of the allocation
    stride = nr.image_width
    ncomp = nr.hits.n_comp.reshape(nr.image_height, stride)[:h, :w].reshape(-1).cpu().numpy()
    diff = np.nonzero(ncomp != fwd["ncomp"])[0]
    ok = np.ones(n, bool)
    if diff.size:
        assert diff.size <= 4, f"{diff.size} rays composite differently"
        ids, _, counts = O.hitlists(geo, ro[diff], rd[diff], float(nr.get_config().alpha_threshold), n)
        for r in range(diff.size):
            ok[ids[r, :counts[r]]] = False
    return ok


def test_safetensors_round_trip(tmp_path):
    """save_safetensors -> from_safetensors renders the same image (1e-6); keys and
    shapes are the reference's; an inference-only load keeps SH in fp16 and
    renders what the oracle renders from the fp16-rounded coefficients."""
    scene = scenes.tiny_scene(1024, seed=5)
    W, H = 64, 48
    cfg = RaytracerConfig()
    rt = Raytracer(cfg, scene.count, W, H)
    load_scene(rt, scene)
    cam = scenes.tiny_camera(W, H)
    with torch.no_grad():
        ref = rt(cam).clone()
    path = rt.save_safetensors(str(tmp_path), 7)
    assert os.path.basename(path) == "gaussians_00007.safetensors"
    import safetensors.torch

    sd = safetensors.torch.load_file(path)
    n = scene.count
    shapes = {"mean": (n, 3), "rotation": (n, 4), "scale": (n, 3), "opacity": (n, 1), "sh_coeffs_dc": (n, 1, 3),
              "sh_coeffs_rest": (n, 15, 3), "current_sh_degree": (1,)}
    assert set(sd) == set(shapes)
    for key, shp in shapes.items():
        assert tuple(sd[key].shape) == shp, key
    assert sd["current_sh_degree"].dtype == torch.int32 and int(sd["current_sh_degree"][0]) == 3
    assert torch.equal(sd["mean"], scene.mean) and torch.equal(sd["sh_coeffs_rest"], scene.sh_rest)

    rt2 = Raytracer.from_safetensors(cfg, path, W, H)
    with torch.no_grad():
        img2 = rt2(cam)
    assert (img2 - ref).abs().max().item() <= 1e-6

    rti = Raytracer.from_safetensors(cfg, path, W, H, inference_only=True)
    rti.cuda_module.get_config().needs_ray_output.fill_(True)  # ray directions without backward state
    gi = rti.cuda_module.get_gaussians()
    assert gi.sh_coeffs_dc.dtype == torch.float16 and gi.sh_coeffs_rest.dtype == torch.float16
    # no gradient or optimiser state for an inference-only raytracer (gaussians.h:236-262)
    assert gi.grad_sh_coeffs_rest.numel() == 45 and gi.first_moment_mean.numel() == 3
    assert gi.pruning_weight.numel() == 1
    with torch.no_grad():
        imgi = rti(cam)
    torch.cuda.synchronize()
    rti.cuda_module.device_status()
    dc16 = scene.sh_dc.half().float().numpy()
    rest16 = scene.sh_rest.half().float().numpy()
    orgb = O.sh_forward(scene.mean, dc16, rest16, cam.origin.astype(np.float32), 3)
    rgb_i = gi.channels.cpu().numpy()
    assert np.abs(rgb_i - orgb).max() < 2e-6
    geo, _ = rti.cuda_module.debug_snapshot()
    ro, rd = gpu_pinhole_rays(rti, cam)
    full = O.render(geo.cpu().numpy(), rgb_i, ro, rd, oracle_knobs(rti.cuda_module))
    assert np.abs(imgi.permute(1, 2, 0).reshape(-1, 3).cpu().numpy() - full["C"]).max() < 1e-5
    assert 0 < (imgi - ref).abs().max().item() < 1e-2  # fp16 coefficients move the image, a little

    # without SH the colours are saved instead
    cfg0 = RaytracerConfig(sh=False)
    rt0 = Raytracer(cfg0, scene.count, W, H)
    g0 = rt0.cuda_module.get_gaussians()
    with torch.no_grad():
        for name in ("mean", "rotation", "scale", "opacity"):
            getattr(g0, name).copy_(getattr(scene, name))
        g0.channels.copy_(scene.channels())
    rt0.cuda_module.rebuild_bvh()
    with torch.no_grad():
        ref0 = rt0(cam).clone()
    p0 = rt0.save_safetensors(str(tmp_path), 8)
    sd0 = safetensors.torch.load_file(p0)
    assert set(sd0) == {"mean", "rotation", "scale", "opacity", "channels"}
    with torch.no_grad():
        img0 = Raytracer.from_safetensors(cfg0, p0, W, H)(cam)
    assert (img0 - ref0).abs().max().item() <= 1e-6


def gpu_pinhole_rays(rt, cam):
    """The GPU's own rays of a no-gradient render with needs_ray_output set:
    directions from the framebuffer, origins the camera origin (pinhole)."""
    fb = rt.cuda_module.get_framebuffer()
    h, w = rt.render_height, rt.render_width
    rd = fb.ray_direction[:h, :w].reshape(-1, 3).cpu().numpy()
    ro = np.tile(cam.origin.astype(np.float32), (rd.shape[0], 1))
    oro, ord_ = O.pinhole_rays(w, h, cam.fov_y, cam.origin.astype(np.float32), cam.rotation_c2w_blender())
    assert np.abs(rd - ord_).max() < 2e-6
    return ro, rd


def binned_cloud(points, colors, dists, bin_size, radius):
    """numpy restatement of the binning of gray/raytracer.py:214-245: points
    rounded to the bin grid, lexicographically sorted unique cells, per-cell
    means of position, colour and distance."""
    rounded = np.round(points / np.float32(bin_size * radius))
    cells, inverse, counts = np.unique(rounded, axis=0, return_inverse=True, return_counts=True)
    inverse = inverse.reshape(-1)
    m = cells.shape[0]
    ps, cs, ds = np.zeros((m, 3)), np.zeros((m, 3)), np.zeros(m)
    np.add.at(ps, inverse, points.astype(np.float64))
    np.add.at(cs, inverse, colors.astype(np.float64))
    np.add.at(ds, inverse, dists.astype(np.float64))
    return ps / counts[:, None], cs / counts[:, None], ds / counts


@pytest.mark.parametrize("binning", [False, True])
def test_from_point_cloud(binning):
    rng = np.random.default_rng(11)
    cfg = RaytracerConfig(init_binning=binning)
    cell = cfg.init_bin_size
    # 600 cell centres on a sphere of radius 0.6, four jittered copies of each
    # (jitter well inside a cell, so the binning keeps exactly 600)
    base = rng.normal(size=(600, 3))
    base = np.round(0.6 * base / np.linalg.norm(base, axis=1, keepdims=True) / cell) * cell
    points = (np.repeat(base, 4, axis=0) + rng.uniform(-0.1, 0.1, size=(2400, 3)) * cell).astype(np.float32)
    colors = rng.uniform(0.1, 0.9, size=(2400, 3)).astype(np.float32)
    dists = rng.uniform(20.0, 40.0, size=2400).astype(np.float32)
    pc = BasicPointCloud(points=points, colors=colors, distances_to_cam=dists, radius=1.0)
    W = H = 64
    rt = Raytracer.from_point_cloud(cfg, pc, W, H)
    g = rt.cuda_module.get_gaussians()
    if binning:
        ep, ec, ed = binned_cloud(points, colors, dists, cfg.init_bin_size, 1.0)
        assert g.count == 600
    else:
        ep, ec, ed = points, colors, dists
        assert g.count == 2400
    assert np.abs(g.mean.cpu().numpy() - ep).max() < 1e-6
    assert np.abs(g.channels.cpu().numpy() - ec).max() < 1e-6
    scale = g.scale.cpu().numpy()
    assert np.abs(scale - np.log(ed * cfg.init_scale)[:, None]).max() < 1e-5
    assert np.array_equal(g.rotation.cpu().numpy(), np.tile([1.0, 0.0, 0.0, 0.0], (g.count, 1)))
    assert np.abs(g.opacity.cpu().numpy() - np.log(cfg.init_opacity / (1 - cfg.init_opacity))).max() < 1e-6
    dc = g.sh_coeffs_dc.cpu().numpy()[:, 0, :]
    assert np.abs(dc - (ec - 0.5) / SH_C0).max() < 1e-5
    assert int(g.current_sh_degree[0]) == 0 and float(g.sh_coeffs_rest.abs().max()) == 0.0
    # the initialised scene renders as the oracle renders it (SH degree 0)
    cam = scenes.tiny_camera(W, H)
    rt.cuda_module.get_config().needs_ray_output.fill_(True)
    with torch.no_grad():
        img = rt(cam)
    torch.cuda.synchronize()
    rt.cuda_module.device_status()
    geo, _ = rt.cuda_module.debug_snapshot()
    orgb = O.sh_forward(g.mean.cpu().numpy(), g.sh_coeffs_dc.cpu().numpy(), g.sh_coeffs_rest.cpu().numpy(),
                        cam.origin.astype(np.float32), 0)
    rgb = g.channels.cpu().numpy()
    assert np.abs(rgb - orgb).max() < 2e-6
    ro, rd = gpu_pinhole_rays(rt, cam)
    full = O.render(geo.cpu().numpy(), rgb, ro, rd, oracle_knobs(rt.cuda_module))
    assert (full["ncomp"] > 0).sum() > 100, "the cloud must be visible"
    assert np.abs(img.permute(1, 2, 0).reshape(-1, 3).cpu().numpy() - full["C"]).max() < 1e-5
