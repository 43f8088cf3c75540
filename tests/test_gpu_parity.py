"""GPU (HIP) path against the CPU oracle on the same seeded inputs.

Bars: hit sets/order/distances bit-exact (same snapshot and rays); image
within 1e-4 max-abs (1e-5 with the GPU's own snapshot and rays); gradients
within 1e-3 relative element by element (|a - b| <= 1e-3 |b| + 1e-6 max|b|)
and to the largest component."""
import numpy as np
import pytest
import torch

from gray_amd import scenes
from oracle import oracle as O

from gpu_util import assert_grads_close, comparable_gaussians, mock_camera, rays_of, rel_err, setup_native

pytestmark = pytest.mark.gpu


# hit sources of the forward: collected lists (default), composite while
# traversing (0), lists so short that most rays of a tile fall back to
# re-traversal while their neighbours composite from lists (3), and lists
# whose walks stop every 2 iterations so that every ray's remaining subtrees
# are finished by the task rounds
# lists: tile bins; walk / tasks: the BVH walk (tasks: budget 2, so most work
# goes through the pooled task rounds); mixed: lists of 3 (most rays overflow
# to re-traversal); traverse: no lists
@pytest.fixture(scope="module", params=[(None, 0, 0), (0, 0, 0), (3, 0, 0), (None, 0, 1), (None, 2, 1)],
                ids=["lists", "traverse", "mixed", "walk", "tasks"])
def tiny(request):
    torch.manual_seed(0)
    scene = scenes.tiny_scene(1024, seed=0)
    cam = scenes.tiny_camera(64, 64)
    cap, budget, mode = request.param
    nr = setup_native(scene, cam, list_capacity=cap, collect_budget=budget, collect_mode=mode)
    torch.set_grad_enabled(True)
    nr.forward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    return scene, cam, nr


def oracle_knobs(nr):
    c = nr.get_config()
    cam = nr.get_camera()
    return O.make_knobs(exp_power=float(c.exp_power), alpha_threshold=float(c.alpha_threshold),
                        t_threshold=float(c.t_threshold), eps_fwd=float(c.eps_forward_normalization),
                        eps_scale_grad=float(c.eps_scale_grad), bg=tuple(c.background_channels.tolist()),
                        znear=float(cam.znear), zfar=float(cam.zfar), render_depth=bool(c.render_depth))


def test_snapshot_matches_oracle(tiny):
    scene, cam, nr = tiny
    geo, aabb = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    ogeo, _ = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, 0.01, 2.0)
    vis = ogeo[:, 12] > 0
    assert np.array_equal(vis, geo[:, 12] > 0)
    assert np.abs(geo[:, 9:12] - ogeo[:, 9:12]).max() == 0.0
    for row in range(3):
        a, b = geo[vis, 3 * row:3 * row + 3], ogeo[vis, 3 * row:3 * row + 3]
        scale = np.abs(b).max(axis=1, keepdims=True)
        assert np.all(np.abs(a - b) <= 1e-5 * scale), row
    for col in (12, 13):
        a, b = geo[vis, col], ogeo[vis, col]
        assert np.all(np.abs(a - b) <= 4e-6 * np.abs(b)), col
    # AABB contains every ellipsoid point (sample the surface)
    box = aabb.cpu().numpy()
    l2w = np.linalg.inv(ogeo[:, :9].reshape(-1, 3, 3).astype(np.float64))
    u = np.random.default_rng(0).normal(size=(64, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    pts = ogeo[:, None, 9:12] + np.einsum("nij,kj->nki", l2w, u)
    assert np.all(pts[vis] >= box[vis, None, :3] - 1e-6) and np.all(pts[vis] <= box[vis, None, 3:] + 1e-6)


def test_rays_match_oracle(tiny):
    scene, cam, nr = tiny
    ro, rd = rays_of(nr)
    oro, ord_ = O.pinhole_rays(64, 64, cam.fov_y, cam.origin.astype(np.float32), cam.rotation_c2w_blender())
    assert np.abs(ro - oro).max() == 0.0
    assert np.abs(rd - ord_).max() < 2e-6


def test_sh_colours_match_oracle(tiny):
    scene, cam, nr = tiny
    rgb = nr.get_gaussians().channels.cpu().numpy()
    orgb = O.sh_forward(scene.mean, scene.sh_dc, scene.sh_rest, cam.origin.astype(np.float32), 3)
    assert np.abs(rgb - orgb).max() < 2e-6


@pytest.mark.parametrize("use_bvh", [True, False])
def test_hit_lists_bit_exact(tiny, use_bvh):
    scene, cam, nr = tiny
    geo, _ = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    ro, rd = rays_of(nr)
    ids, dists, counts = nr.debug_hitlists(256, use_bvh=use_bvh)
    ids, dists, counts = ids.cpu().numpy(), dists.cpu().numpy(), counts.cpu().numpy()
    oids, odists, ocounts = O.hitlists(geo, ro, rd, 0.01, 256)
    assert counts.max() < 256
    assert np.array_equal(counts, ocounts)
    assert counts.sum() > 1000, "scene must produce hits"
    for r in range(len(counts)):
        c = counts[r]
        assert np.array_equal(ids[r, :c].astype(np.uint32), oids[r, :c])
        assert np.array_equal(dists[r, :c].view(np.uint32), odists[r, :c].view(np.uint32))


def test_forward_matches_oracle(tiny):
    scene, cam, nr = tiny
    fb = nr.get_framebuffer()
    img = fb.output_channels[:64, :64].reshape(-1, 3).cpu().numpy()
    geo, _ = nr.debug_snapshot()
    ro, rd = rays_of(nr)
    rgb = nr.get_gaussians().channels.cpu().numpy()
    k = oracle_knobs(nr)
    same = O.render(geo.cpu().numpy(), rgb, ro, rd, k)
    assert np.abs(img - same["C"]).max() < 1e-5
    T = fb.transmittance[:64, :64].reshape(-1).cpu().numpy()
    F = fb.full_transmittance[:64, :64].reshape(-1).cpu().numpy()
    assert np.abs(T - same["T"]).max() < 1e-5 and np.abs(F - same["F"]).max() < 1e-5
    ncomp = nr.hits.n_comp[: 64 * 64].cpu().numpy()
    assert (ncomp != same["ncomp"]).sum() <= 2
    # stored composited lists replay the oracle's front-to-back order
    cap = nr.hits.capacity
    hid = nr.hits.hit_ids[:, : 64 * 64].cpu().numpy().T
    for r in np.nonzero(ncomp == same["ncomp"])[0]:
        c = min(ncomp[r], cap)
        assert np.array_equal(hid[r, :c].astype(np.uint32), same["ids"][r, :c])
    # full oracle: its own snapshot, rays and SH
    ogeo, _ = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, 0.01, 2.0)
    oro, ord_ = O.pinhole_rays(64, 64, cam.fov_y, cam.origin.astype(np.float32), cam.rotation_c2w_blender())
    orgb = O.sh_forward(scene.mean, scene.sh_dc, scene.sh_rest, cam.origin.astype(np.float32), 3)
    full = O.render(ogeo, orgb, oro, ord_, k)
    assert np.abs(img - full["C"]).max() < 1e-4


def test_list_counts_match_oracle(tiny):
    scene, cam, nr = tiny
    if nr.list_capacity == 0:
        assert nr.debug_list_counts().numel() == 0
        return
    counts = nr.debug_list_counts().cpu().numpy().reshape(8, 8, 8, 8)  # tile y, tile x, lane y, lane x
    counts = counts.transpose(0, 2, 1, 3).reshape(-1)
    geo, _ = nr.debug_snapshot()
    ro, rd = rays_of(nr)
    same = O.render(geo.cpu().numpy(), nr.get_gaussians().channels.cpu().numpy(), ro, rd, oracle_knobs(nr))
    complete = counts >= 0
    assert np.array_equal(counts[complete], same["nhits"][complete])
    assert np.all(same["nhits"][~complete] > nr.list_capacity)
    if nr.list_capacity < 8:
        assert 0 < complete.sum() < complete.size


def test_backward_matches_oracle(tiny):
    scene, cam, nr = tiny
    fb = nr.get_framebuffer()
    g = nr.get_gaussians()
    img = fb.output_channels[:64, :64].clone().requires_grad_()
    target = torch.rand(64, 64, 3, generator=torch.Generator().manual_seed(1)).cuda()
    loss = torch.nn.functional.l1_loss(img, target)
    loss.backward()
    with torch.no_grad():
        fb.grad_output_channels[:64, :64].copy_(img.grad)
        for name in ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels",
                     "grad_sh_coeffs_dc", "grad_sh_coeffs_rest", "pruning_weight"):
            getattr(g, name).zero_()
    nr.backward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    geo, _ = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    _, l2w = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, 0.01, 2.0)
    ro, rd = rays_of(nr)
    rgb = g.channels.cpu().numpy()
    k = oracle_knobs(nr)
    fwd = O.render(geo, rgb, ro, rd, k)
    gout = img.grad.reshape(-1, 3).cpu().numpy()
    ob = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, fwd, gout, k, 64 * 64)
    vis = g.was_visible.reshape(-1).cpu().numpy().astype(np.uint8)
    gdc, grest, gmean = O.sh_backward(scene.mean, scene.sh_rest, cam.origin.astype(np.float32), rgb, ob["rgb"],
                                      vis, 3, grad_mean=ob["mean"])
    ok = comparable_gaussians(nr.hits.n_comp[: 64 * 64].cpu().numpy(), fwd, geo, ro, rd)
    assert_grads_close([
        ("channels", g.grad_channels.cpu(), ob["rgb"]), ("opacity", g.grad_opacity.cpu(), ob["opacity"]),
        ("scale", g.grad_scale.cpu(), ob["scale"]), ("rotation", g.grad_rotation.cpu(), ob["rotation"]),
        ("mean", g.grad_mean.cpu(), gmean), ("sh_dc", g.grad_sh_coeffs_dc.cpu(), gdc),
        ("sh_rest", g.grad_sh_coeffs_rest.cpu(), grest), ("pruning", g.pruning_weight.cpu(), ob["pruning"]),
    ], ok)
    assert np.array_equal(vis > 0, (np.abs(ob["opacity"].reshape(-1)) > 0) | (ob["pruning"].reshape(-1) > 0))


def pixel_grid_scene(cam, blocks, logit=0.0):
    """Tiny Gaussians on pixel rays: for each (depths, columns) block, one per
    depth on the ray of every pixel in those columns.  Each is hit by its own
    pixel's ray only (radius under a seventh of the pixel spacing at these
    depths, centre a third of it off the ray), so an 8x8 tile of a block composites 64 x len(depths) distinct
    Gaussians."""
    w, h = cam.image_width, cam.image_height
    ro, rd = O.pinhole_rays(w, h, cam.fov_y, cam.origin.astype(np.float32), cam.rotation_c2w_blender())
    means = []
    for depths, columns in blocks:
        pix = np.array([y * w + x for y in range(h) for x in columns])
        means += [ro[pix] + t * rd[pix] for t in depths]
    mean = np.concatenate(means)
    n = mean.shape[0]
    gen = torch.Generator().manual_seed(11)
    # off the ray by about a third of the radius, so the hit is not at the
    # centre, where the rotation and scale gradients vanish to rounding
    off = torch.randn(n, 3, generator=gen).numpy()
    mean = (mean + 0.002 * off / np.linalg.norm(off, axis=1, keepdims=True)).astype(np.float32)
    rotation = torch.nn.functional.normalize(torch.randn(n, 4, generator=gen), dim=1)
    scale = torch.full((n, 3), -5.8) + 0.2 * torch.rand(n, 3, generator=gen)
    opacity = logit + torch.randn(n, 1, generator=gen) * 0.5
    sh_dc = 0.5 + 0.2 * torch.randn(n, 1, 3, generator=gen)
    sh_rest = 0.05 * torch.randn(n, 15, 3, generator=gen)
    return scenes.GaussianScene(torch.from_numpy(mean), rotation, scale, opacity, sh_dc, sh_rest)


@pytest.mark.parametrize("size,blocks,logit,cap", [
    (64, [((2.6, 3.0, 3.4), range(64))], 0.0, 16),
    (64, [((2.6, 3.0, 3.4), range(32)), ((3.0,), range(32, 64))], 0.0, 16),
    (16, [(tuple(np.linspace(2.0, 4.0, 70)), range(16))], -3.5, 80),
    (64, [((2.6, 2.9, 3.2, 3.5, 3.8), range(64))], -1.0, 16),
    (16, [(tuple(np.linspace(2.0, 4.0, 40)), range(16))], -3.5, 48),
], ids=["every-tile", "half", "deep", "second-table-full", "replay"])
def test_backward_table_overflow(size, blocks, logit, cap):
    """Tiles whose composited hits span more distinct Gaussians than the
    stored-list backward's per-tile table holds (128 slots): three Gaussians on
    every pixel's ray make 192 per 8x8 tile.  The hits whose Gaussian found no
    slot (at most 12 per lane) are redone after the flush and summed per
    Gaussian in the cleared table, then added to the device.  In the "half"
    frame the right half has one Gaussian per pixel (64 per tile), so tables
    that overflow and tables that do not share the launch.  With five per ray
    ("second-table-full", 320 per tile) the Gaussians left over do not all fit
    the second table either, and those hits go to the device one by one.  With
    40 per ray ("replay") a lane has more failed hits than it keeps, and
    replays its list to them.  In the "deep" frame 70 faint Gaussians lie on
    every ray, so slots run out past the 64 steps a lane can mark: those tiles
    drop their table and the recompute kernel redoes their stored lists.
    Gradients equal the oracle's element by element, on two backward calls in
    a row (the first clears the marks it leaves)."""
    cam = scenes.tiny_camera(size, size)
    scene = pixel_grid_scene(cam, blocks, logit)
    nr = setup_native(scene, cam, t_threshold=0.0001, hit_capacity=cap, sh=False, channels=scene.channels())
    torch.set_grad_enabled(True)
    nr.forward_pass()
    fb = nr.get_framebuffer()
    g = nr.get_gaussians()
    ncomp = nr.hits.n_comp[: size * size].cpu().numpy().reshape(size, size)
    for depths, columns in blocks:
        # every Gaussian of a short column is composited; a deep column passes
        # step 64; of 40, a few of the faintest may fall under the alpha threshold
        need = len(depths) - 5 if 10 < len(depths) < 64 else min(len(depths), 65)
        assert (ncomp[:, list(columns)] >= need).all(), "grid pixels composite their Gaussians"
    gout = torch.randn(fb.grad_output_channels.shape, generator=torch.Generator().manual_seed(4)).cuda()
    fb.grad_output_channels.copy_(gout)
    names = ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels", "pruning_weight")
    runs = []
    for _ in range(2):
        with torch.no_grad():
            for name in names:
                getattr(g, name).zero_()
        nr.backward_pass()
        torch.cuda.synchronize()
        nr.device_status()
        runs.append({k: getattr(g, k).cpu().numpy().copy() for k in names})
    geo, _ = nr.debug_snapshot()
    geo = geo.cpu().numpy()
    _, l2w = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, 0.01, 2.0)
    ro, rd = rays_of(nr)
    rgb = g.channels.cpu().numpy()
    k = oracle_knobs(nr)
    fwd = O.render(geo, rgb, ro, rd, k)
    ob = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, fwd,
                    gout[:size, :size].reshape(-1, 3).cpu().numpy(), k, size * size)
    ok = comparable_gaussians(ncomp.reshape(-1), fwd, geo, ro, rd)
    for run in runs:
        assert_grads_close([
            ("channels", run["grad_channels"], ob["rgb"]), ("opacity", run["grad_opacity"], ob["opacity"]),
            ("scale", run["grad_scale"], ob["scale"]), ("rotation", run["grad_rotation"], ob["rotation"]),
            ("mean", run["grad_mean"], ob["mean"]), ("pruning", run["pruning_weight"], ob["pruning"]),
        ], ok)


def test_recompute_backward_equals_stored():
    scene = scenes.tiny_scene(1024, seed=3)
    cam = scenes.tiny_camera(48, 40)
    grads = []
    for cap in (64, 0, 3):
        nr = setup_native(scene, cam, t_threshold=0.001, hit_capacity=cap)
        torch.set_grad_enabled(True)
        nr.forward_pass()
        fb = nr.get_framebuffer()
        fb.grad_output_channels.copy_(torch.randn(fb.grad_output_channels.shape,
                                                  generator=torch.Generator().manual_seed(5)).cuda())
        g = nr.get_gaussians()
        for name in ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels"):
            getattr(g, name).zero_()
        nr.backward_pass()
        torch.cuda.synchronize()
        nr.device_status()
        grads.append({k: getattr(g, k).cpu().numpy().copy() for k in
                      ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels")})
    for other in grads[1:]:
        for key in grads[0]:
            assert rel_err(other[key], grads[0][key]) < 2e-4, key


@pytest.mark.parametrize("athr", [0.01, 0.3])
def test_reference_single_gaussian(athr):
    from gray_amd import Raytracer, RaytracerConfig
    from oracle import torch_ref
    from test_oracle import target_image

    cfg = RaytracerConfig(sh=False)
    rt = Raytracer(cfg, 1, 32, 32)
    config = rt.cuda_module.get_config()
    config.zero_grads.fill_(False)
    config.alpha_threshold.fill_(athr)
    config.background_channels.copy_(torch.tensor([0.1, 0.1, 0.1]))
    mu = np.array([[-0.05, -0.05, 0.0]])
    ls = np.log([[0.1, 0.3, 0.3]])
    col = np.array([[1.0, 0.0, 0.0]])
    q = np.array([[-0.5, 0.0, 0.0, 1.0]])
    lo = np.array([[0.0]])
    g = rt.cuda_module.get_gaussians()
    with torch.no_grad():
        for name, v in (("mean", mu), ("scale", ls), ("channels", col), ("rotation", q), ("opacity", lo)):
            getattr(g, name).copy_(torch.tensor(v, dtype=torch.float32))
    rt.cuda_module.rebuild_bvh()
    render = rt(mock_camera(32))
    target = torch.tensor(target_image()).permute(2, 0, 1).cuda()
    loss = torch.nn.functional.l1_loss(render, target)
    rt.backward(loss)
    fb = rt.cuda_module.get_framebuffer()
    ro = fb.ray_origin.reshape(-1, 3).cpu().double()
    rd = fb.ray_direction.reshape(-1, 3).cpu().double()
    t = lambda a: torch.tensor(np.asarray(a, np.float64), requires_grad=True)  # noqa: E731
    tm, ts, tc, tq, to = t(mu), t(ls), t(col), t(q), t(lo)
    ref, _ = torch_ref.render(tm, tq, ts, to, tc, ro, rd, 2.0, athr, 0.03, (0.1, 0.1, 0.1))
    ref_img = ref.reshape(32, 32, 3).permute(2, 0, 1)
    diff = (render.detach().cpu().double() - ref_img.detach()).abs()
    assert diff.mean() < 1e-5 and diff.max() < 1e-4
    torch.nn.functional.l1_loss(ref_img, target.cpu().double()).backward()
    for name, ref_t in (("mean", tm), ("rotation", tq), ("scale", ts), ("opacity", to), ("channels", tc)):
        assert (getattr(g, name).grad.cpu().double() - ref_t.grad).abs().max() < 1e-4, name


@pytest.mark.parametrize("tthr", [0.0, 0.5])
def test_reference_multiple_gaussians(tthr):
    from gray_amd import Raytracer, RaytracerConfig
    from oracle import torch_ref
    from test_oracle import target_image

    cfg = RaytracerConfig(t_threshold=tthr, sh=False)
    rt = Raytracer(cfg, 3, 32, 32)
    rt.cuda_module.get_config().zero_grads.fill_(False)
    mu = np.array([[0.05, 0.05, -0.25], [-0.05, -0.05, 0.25], [0.0, 0.0, 0.0]])
    ls = np.log([[1.0, 0.2, 0.1], [0.1, 0.3, 0.3], [0.4, 0.4, 0.4]])
    col = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    q = np.array([[1.0, 0.0, 0.0, 1.0]] * 3)
    opv = np.array([[0.5], [0.7], [0.7]])
    lo = np.log(opv / (1 - opv))
    g = rt.cuda_module.get_gaussians()
    with torch.no_grad():
        for name, v in (("mean", mu), ("scale", ls), ("channels", col), ("rotation", q), ("opacity", lo)):
            getattr(g, name).copy_(torch.tensor(v, dtype=torch.float32))
    rt.cuda_module.rebuild_bvh()
    render = rt(mock_camera(32))
    target = torch.tensor(target_image()).permute(2, 0, 1).cuda()
    rt.backward(torch.nn.functional.l1_loss(render, target))
    fb = rt.cuda_module.get_framebuffer()
    ro = fb.ray_origin.reshape(-1, 3).cpu().double()
    rd = fb.ray_direction.reshape(-1, 3).cpu().double()
    t = lambda a: torch.tensor(np.asarray(a, np.float64), requires_grad=True)  # noqa: E731
    tm, ts, tc, tq, to = t(mu), t(ls), t(col), t(q), t(lo)
    ref, _ = torch_ref.render(tm, tq, ts, to, tc, ro, rd, 2.0, 0.01, tthr, (0, 0, 0))
    ref_img = ref.reshape(32, 32, 3).permute(2, 0, 1)
    diff = (render.detach().cpu().double() - ref_img.detach()).abs()
    assert diff.mean() < 1e-5 and diff.max() < 1e-4
    torch.nn.functional.l1_loss(ref_img, target.cpu().double()).backward()
    for name, ref_t in (("mean", tm), ("rotation", tq), ("scale", ts), ("opacity", to), ("channels", tc)):
        assert (getattr(g, name).grad.cpu().double() - ref_t.grad).abs().max() < 1e-4, name


def test_adam_matches_oracle():
    scene = scenes.tiny_scene(700, seed=9)
    cam = scenes.tiny_camera(16, 16)
    nr = setup_native(scene, cam)
    g = nr.get_gaussians()
    gen = torch.Generator().manual_seed(2)
    host = {}
    for name in g.PARAMS:
        grad = torch.randn(getattr(g, name).shape, generator=gen)
        # Gaussians no ray reached have +0 gradients (the kernel skips storing
        # them again), and -0 must stay -0 as the reference's g *= 0 leaves it
        grad[100:300] = 0.0
        grad[300:340] = -0.0
        getattr(g, "grad_" + name).copy_(grad)
        host[name] = [getattr(g, name).cpu().numpy().copy(), grad.numpy().copy(),
                      np.zeros(grad.shape, np.float32), np.zeros(grad.shape, np.float32)]
    lrs = dict(zip(g.PARAMS, (1.6e-4, 4e-3, 5e-3, 2e-2, 2.5e-3, 2.5e-3, 6e-4)))
    for name, lr_name in zip(g.PARAMS, g.LRS):
        getattr(g, lr_name).fill_(lrs[name])
    g.was_visible.fill_(True)
    nr.step(update_pruning_stats=True)
    torch.cuda.synchronize()
    for name in g.PARAMS:
        p, gr, m, v = host[name]
        if name == "channels":
            # SH on: channels are not optimised, their gradient is set to +0 (adam.cu:35-40)
            gc = getattr(g, "grad_" + name).cpu().numpy()
            assert np.all(gc == 0) and not np.signbit(gc).any()
            continue
        O.adam(p, gr, m, v, lrs[name], 0.9, 0.999, 1e-15, 1)
        assert np.abs(getattr(g, name).cpu().numpy() - p).max() < 1e-6, name
        assert np.abs(getattr(g, "first_moment_" + name).cpu().numpy() - m).max() < 1e-6
        assert np.abs(getattr(g, "second_moment_" + name).cpu().numpy() - v).max() < 1e-6
        ga = getattr(g, "grad_" + name).cpu().numpy()
        assert np.all(ga == 0)
        assert np.array_equal(np.signbit(ga), np.signbit(gr)), name  # g *= 0 keeps the sign (adam.cu:19-20)
    assert np.all(g.pruning_counter.cpu().numpy() == 1)


def test_inference_only_fp16_sh_matches_oracle():
    """inference_only keeps SH coefficients in fp16 (reference gaussians.h:132);
    the image equals the oracle's on the same (fp16-rounded) coefficients."""
    from gray_amd.native import NativeRaytracer

    scene = scenes.tiny_scene(1024, seed=12)
    cam = scenes.tiny_camera(48, 40)
    nr = NativeRaytracer(48, 40, scene.count, 3, 300_000_000, 120_000_000, inference_only=True)
    g = nr.get_gaussians()
    assert g.sh_coeffs_rest.dtype == torch.float16
    with torch.no_grad():
        g.mean.copy_(scene.mean)
        g.rotation.copy_(scene.rotation)
        g.scale.copy_(scene.scale)
        g.opacity.copy_(scene.opacity)
        g.sh_coeffs_dc.copy_(scene.sh_dc)
        g.sh_coeffs_rest.copy_(scene.sh_rest)
        g.current_sh_degree.fill_(3)
    from gpu_util import set_camera

    set_camera(nr, cam)
    nr.rebuild_bvh()
    with torch.no_grad():
        nr.forward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    dc16 = scene.sh_dc.half().float().numpy()
    rest16 = scene.sh_rest.half().float().numpy()
    rgb = O.sh_forward(scene.mean.numpy(), dc16, rest16, cam.origin.astype(np.float32), 3)
    assert np.abs(g.channels.cpu().numpy() - rgb).max() < 1e-5
    geo, _ = nr.debug_snapshot()
    # no backward state in inference: the rays come from the closed form
    ro, rd = O.pinhole_rays(48, 40, cam.fov_y, cam.origin.astype(np.float32), cam.rotation_c2w_blender())
    ref = O.render(geo.cpu().numpy(), rgb, ro, rd, oracle_knobs(nr))
    img = nr.get_framebuffer().output_channels[:40, :48].reshape(-1, 3).cpu().numpy()
    assert np.abs(img - ref["C"]).max() < 1e-5
