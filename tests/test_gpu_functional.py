"""Multi-camera autograd rendering (gray_amd.functional) on the GPU: a batch of
views equals the same views rendered one at a time, its gradients equal the
per-view oracle gradients summed, and they match float64 central finite
differences (h = 1e-4) of an independent restatement (oracle/torch_ref.py)."""
import math

import numpy as np
import pytest
import torch

from gray_amd import scenes
from gray_amd.functional import MultiViewRenderer, RenderSettings
from oracle import oracle as O
from oracle import torch_ref

from gpu_util import assert_grads_close, comparable_gaussians, rel_err, setup_native

pytestmark = pytest.mark.gpu

PARAMS = ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")


def scene_tensors(scene, requires_grad=True):
    return [getattr(scene, k).cuda().clone().requires_grad_(requires_grad) for k in PARAMS]


def ring(count, w, h, radius=3.0):
    return [scenes.ring_camera(radius, 2 * math.pi * i / count + 0.3, w, h, 60.0, uid=i) for i in range(count)]


@pytest.mark.parametrize("views", [3, 10])
def test_batch_equals_single_views(views):
    scene = scenes.tiny_scene(1024, seed=4)
    cams = ring(views, 40, 32)
    r = MultiViewRenderer(40, 32, RenderSettings(t_threshold=0.03))
    with torch.no_grad():
        imgs = r.render(cams, *scene_tensors(scene, False))
    torch.cuda.synchronize()
    r.device_status()
    for i, cam in enumerate(cams):
        nr = setup_native(scene, cam, t_threshold=0.03)
        nr.forward_pass()
        single = nr.get_framebuffer().output_channels[:32, :40]
        assert (imgs[i] - single).abs().max().item() < 1e-6, i


def test_gradients_equal_summed_oracle_views():
    scene = scenes.tiny_scene(1024, seed=5)
    cams = ring(3, 32, 24)
    r = MultiViewRenderer(32, 24, RenderSettings(t_threshold=0.03))
    params = scene_tensors(scene)
    imgs = r.render(cams, *params)
    w = torch.randn(imgs.shape, generator=torch.Generator().manual_seed(2)).cuda()
    (imgs * w).sum().backward()
    torch.cuda.synchronize()
    r.device_status()
    k = O.make_knobs(t_threshold=0.03)
    _, l2w = O.snapshot(scene.mean.numpy(), scene.rotation.numpy(), scene.scale.numpy(), scene.opacity.numpy(),
                        0.01, 2.0)
    geo = r.debug_snapshot()[0].cpu().numpy()  # the GPU's snapshot: the same hit sets as the render
    tot = {key: 0 for key in ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")}
    ok = np.ones(scene.count, bool)
    for i, cam in enumerate(cams):
        ro = r.ro[i].reshape(-1, 3).cpu().numpy()
        rd = r.rd[i].reshape(-1, 3).cpu().numpy()
        origin = cam.origin.astype(np.float32)
        rgb = O.sh_forward(scene.mean.numpy(), scene.sh_dc.numpy(), scene.sh_rest.numpy(), origin, 3)
        fwd = O.render(geo, rgb, ro, rd, k)
        assert np.abs(imgs[i].detach().reshape(-1, 3).cpu().numpy() - fwd["C"]).max() < 1e-5
        ok &= comparable_gaussians(r.n_comp[i].cpu().numpy(), fwd, geo, ro, rd)
        b = O.backward(geo, l2w, rgb, scene.rotation.numpy(), scene.scale.numpy(), scene.opacity.numpy(), ro, rd,
                       fwd, w[i].reshape(-1, 3).cpu().numpy(), k, 32 * 24)
        vis = ((np.abs(b["opacity"]).reshape(-1) > 0) | (b["pruning"].reshape(-1) > 0)).astype(np.uint8)
        gdc, grest, gmean = O.sh_backward(scene.mean.numpy(), scene.sh_rest.numpy(), origin, rgb, b["rgb"], vis, 3,
                                          grad_mean=b["mean"])
        tot["mean"] = tot["mean"] + gmean
        tot["rotation"] = tot["rotation"] + b["rotation"]
        tot["scale"] = tot["scale"] + b["scale"]
        tot["opacity"] = tot["opacity"] + b["opacity"]
        tot["sh_dc"] = tot["sh_dc"] + gdc
        tot["sh_rest"] = tot["sh_rest"] + grest
    assert_grads_close([(name, p.grad.cpu().reshape(p.shape[0], -1), tot[name].reshape(p.shape[0], -1))
                        for name, p in zip(PARAMS, params)], ok)


def test_gradients_match_fp64_finite_differences():
    """t_threshold 0 (every hit composited): the forward is then smooth in every
    parameter away from the hit-criterion boundary, and the reference backward
    is its exact derivative.  Parameters whose +-h perturbation changes any
    ray's hit set sit on that boundary and are skipped."""
    scene = scenes.tiny_scene(300, seed=6)
    W = H = 24
    cams = ring(4, W, H, radius=2.5)
    s = RenderSettings(t_threshold=0.0)
    r = MultiViewRenderer(W, H, s)
    params = scene_tensors(scene)
    imgs = r.render(cams, *params)
    wgt = torch.randn(imgs.shape, generator=torch.Generator().manual_seed(3)).cuda()
    (imgs * wgt).sum().backward()
    torch.cuda.synchronize()
    r.device_status()
    ro = torch.tensor(r.ro.reshape(4, -1, 3).cpu().numpy(), dtype=torch.float64)
    rd = torch.tensor(r.rd.reshape(4, -1, 3).cpu().numpy(), dtype=torch.float64)
    wv = torch.tensor(wgt.reshape(4, -1, 3).cpu().numpy(), dtype=torch.float64)
    p64 = [getattr(scene, k).to(torch.float64) for k in PARAMS]
    origins = [torch.tensor(c.origin, dtype=torch.float64) for c in cams]

    def masks(pp, j):
        sl = [t[j:j + 1] for t in pp[:4]]
        return [torch_ref.hit_mask(*sl, ro[v], rd[v])[:, 0] for v in range(4)]

    def loss(pp, rays):
        tot = torch.zeros((), dtype=torch.float64)
        for v in range(4):
            idx = rays[v]
            if idx.numel() == 0:
                continue
            col = torch_ref.sh_colors(pp[0], pp[4], pp[5], origins[v])
            img, _ = torch_ref.render(pp[0], pp[1], pp[2], pp[3], col, ro[v][idx], rd[v][idx], 2.0, 0.01, 0.0)
            tot = tot + (img * wv[v][idx]).sum()
        return tot

    rng = np.random.default_rng(7)
    seen = torch.zeros(300, dtype=torch.bool)
    for v in range(4):
        seen |= torch_ref.hit_mask(*p64[:4], ro[v], rd[v]).any(0)
    cand = np.nonzero(seen.numpy())[0]
    h = 1e-4
    checked, worst = 0, 0.0
    for _ in range(400):
        if checked == 64:
            break
        j = int(rng.choice(cand))
        which = int(rng.integers(0, 6))
        comp = int(rng.integers(0, p64[which][j].numel()))
        base = masks(p64, j)
        pert = []
        for sign in (1.0, -1.0):
            pp = [t.clone() for t in p64]
            pp[which][j].view(-1)[comp] += sign * h
            pert.append(pp)
        if any(not torch.equal(a, b) for pp in pert for a, b in zip(masks(pp, j), base)):
            continue  # on the hit-criterion boundary
        rays = [torch.nonzero(m).reshape(-1) for m in base]
        fd = float(loss(pert[0], rays) - loss(pert[1], rays)) / (2 * h)
        g = float(params[which].grad[j].reshape(-1)[comp])
        ref = float(params[which].grad.abs().max())
        err = abs(g - fd) / (abs(fd) + 1e-3 * ref + 1e-9)
        worst = max(worst, err)
        assert err < 1e-3, (PARAMS[which], j, comp, g, fd)
        checked += 1
    assert checked == 64, checked
    print(f"64 parameters, worst relative error {worst:.2e}")


def test_config3_scale_under_memory_cap():
    """200,000 Gaussians (surface generator), 8 cameras at 512x512 in one
    launch, hit lists under a 4 GB cap: composited lists longer than the
    derived store are recomputed in the backward.  Gradients equal those of an
    uncapped store, and 64 random parameters match float64 central
    differences (h = 1e-4; t_threshold 0 as above) taken with the GPU's hit
    structure held fixed, half of them drawn from the largest fifth of the
    Gaussians so that large footprints are covered; none is skipped."""
    scene = scenes.surface_scene(200_000, 20, -4.5, 0.4, seed=1)
    cams = scenes.ring_cameras(8, 6.0, 512, 512)
    capped = MultiViewRenderer(512, 512, RenderSettings(t_threshold=0.0, memory_cap_bytes=4 << 30))
    params = scene_tensors(scene)
    imgs = capped.render(cams, *params)
    wgt = torch.randn(imgs.shape, generator=torch.Generator().manual_seed(9)).cuda()
    (imgs * wgt).sum().backward()
    torch.cuda.synchronize()
    capped.device_status()
    stored, lists = capped.last_capacities
    ncomp = capped.n_comp.max().item()
    assert ncomp > stored, "the cap must force recomputed lists for this check"
    grads = [p.grad.clone() for p in params]
    roomy = MultiViewRenderer(512, 512, RenderSettings(t_threshold=0.0, memory_cap_bytes=64 << 30,
                                                       hit_capacity=64, list_capacity=512))
    params2 = scene_tensors(scene)
    (roomy.render(cams, *params2) * wgt).sum().backward()
    torch.cuda.synchronize()
    for name, a, b in zip(PARAMS, grads, params2):
        assert rel_err(a.cpu(), b.grad.cpu()) < 1e-4, name
    # float64 central differences on 64 random parameters, each within the
    # GPU's piece of the piecewise-smooth forward: every ray keeps the hit
    # structure the GPU composited (from the oracle on the GPU's snapshot and
    # rays, which matches the GPU hit for hit); float64 alone would decide a
    # few hits and near-equal distances differently on rays of large Gaussians
    geo, aabb = capped.debug_snapshot()
    geo = geo.cpu().numpy()
    aabb = aabb.cpu().numpy()
    ro32 = capped.ro.reshape(8, -1, 3).cpu().numpy()
    rd32 = capped.rd.reshape(8, -1, 3).cpu().numpy()
    wv = wgt.reshape(8, -1, 3).cpu().numpy().astype(np.float64)
    p64 = [getattr(scene, k).to(torch.float64) for k in PARAMS]
    origins = [torch.tensor(c.origin, dtype=torch.float64) for c in cams]
    k = O.make_knobs(t_threshold=0.0)

    def rays_of_gaussian(j):
        """Per view: the rays whose hit list holds j, their hit lists and composited lists."""
        out = []
        for v in range(8):
            o, d = ro32[v], rd32[v]
            with np.errstate(divide="ignore", invalid="ignore"):
                t0 = (aabb[j, :3] - o) / d
                t1 = (aabb[j, 3:] - o) / d
            tn = np.nanmax(np.minimum(t0, t1), axis=1)
            tf = np.nanmin(np.maximum(t0, t1), axis=1)
            near = np.nonzero((tn <= tf) & (tf >= 0))[0]
            if near.size == 0:
                continue
            ids, _, counts = O.hitlists(geo, o[near], d[near], 0.01, 4096)
            has = np.array([j in ids[i, :counts[i]] for i in range(near.size)], bool)
            if not has.any():
                continue
            rays = near[has]
            fwd = O.render(geo, capped.rgb[v].cpu().numpy(), o[rays], d[rays], k)
            hits = [ids[i, :counts[i]] for i in np.nonzero(has)[0]]
            comp = [fwd["ids"][i, :fwd["ncomp"][i]] for i in range(rays.size)]
            out.append((v, rays, hits, comp))
        return out

    def loss(pp, per_view, sub):
        pos = {int(g): i for i, g in enumerate(sub)}
        q = [t[torch.from_numpy(sub)] for t in pp]
        tot = torch.zeros((), dtype=torch.float64)
        for v, rays, hits, comp in per_view:
            fixed = []
            for hl, cl in zip(hits, comp):
                cset = set(int(g) for g in cl)
                fixed.append(([pos[int(g)] for g in cl], [pos[int(g)] for g in hl if int(g) not in cset]))
            col = torch_ref.sh_colors(q[0], q[4], q[5], origins[v])
            img, _ = torch_ref.render(q[0], q[1], q[2], q[3], col, torch.from_numpy(ro32[v][rays]).double(),
                                      torch.from_numpy(rd32[v][rays]).double(), 2.0, 0.01, 0.0, fixed=fixed)
            tot = tot + (img * torch.from_numpy(wv[v][rays])).sum()
        return tot

    rng = np.random.default_rng(11)
    vis = np.nonzero(params[3].grad.abs().reshape(-1).cpu().numpy() > 0)[0]
    # stratified: every other draw from the largest fifth of the visible
    # Gaussians (by largest axis), so Gaussians seen by hundreds of rays are
    # checked as well as small ones
    size = scene.scale.max(dim=1).values.numpy()[vis]
    big = vis[np.argsort(size)[-max(1, len(vis) // 5):]]
    h = 1e-4
    checked, worst, many_rays = 0, 0.0, 0
    for trial in range(64):
        j = int(rng.choice(big if trial % 2 else vis))
        which = int(rng.integers(0, 6))
        comp = int(rng.integers(0, p64[which][j].numel()))
        per_view = rays_of_gaussian(j)
        n_rays = sum(len(pv[1]) for pv in per_view)
        assert n_rays > 0, j
        sub = np.unique(np.concatenate([np.concatenate(pv[2]) for pv in per_view] + [np.array([j])]))
        pert = []
        for sign in (1.0, -1.0):
            pp = [t.clone() for t in p64]
            pp[which][j].view(-1)[comp] += sign * h
            pert.append(pp)
        fd = float(loss(pert[0], per_view, sub) - loss(pert[1], per_view, sub)) / (2 * h)
        g = float(params[which].grad[j].reshape(-1)[comp])
        ref = float(params[which].grad.abs().max())
        err = abs(g - fd) / (abs(fd) + 1e-3 * ref + 1e-9)
        worst = max(worst, err)
        assert err < 1e-3, (PARAMS[which], j, comp, g, fd, n_rays)
        checked += 1
        many_rays += int(n_rays > 64)
    assert checked == 64, checked
    assert many_rays >= 8, many_rays
    print(f"config 3: capacities stored {stored} / lists {lists}, max composited {ncomp}; "
          f"64 parameters ({many_rays} seen by more than 64 rays), worst relative error {worst:.2e}")
