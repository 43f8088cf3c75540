"""The training frame at full size (1M Gaussians on 20 surfaces, 980x545
camera on a ring of radius 6) against the oracle's brute force over all 1M
Gaussians on a seeded sample of pixels: hit lists bit-exact, image and
transmittances within 1e-5 (same snapshot and rays), and the gradients of a
loss that only those pixels see within 1e-3 relative, element by element."""
import numpy as np
import pytest
import torch

from gray_amd import scenes
from oracle import oracle as O

from gpu_util import assert_grads_close, comparable_gaussians, rays_of, setup_native

pytestmark = pytest.mark.gpu

W, H, N = 980, 545, 1_000_000


@pytest.fixture(scope="module")
def frame():
    scene = scenes.surface_scene(N, 20, -4.5, 0.4, seed=0)
    cam = scenes.ring_camera(6.0, 0.0, W, H, 60.0)
    nr = setup_native(scene, cam, t_threshold=0.01)
    torch.set_grad_enabled(True)
    nr.forward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    pix = np.random.default_rng(7).choice(W * H, 512, replace=False)
    geo, _ = nr.debug_snapshot()
    ro, rd = rays_of(nr)
    return scene, cam, nr, pix, geo.cpu().numpy(), ro[pix], rd[pix]


def test_fullsize_hit_lists_bit_exact(frame):
    _, _, nr, pix, geo, ro, rd = frame
    cap = 256  # the debug kernel keeps the first `cap` hits it meets: lists compare when all fit
    ids, dists, counts = nr.debug_hitlists(cap)
    sel = torch.from_numpy(pix).cuda()
    ids, dists, counts = ids[sel].cpu().numpy(), dists[sel].cpu().numpy(), counts[sel].cpu().numpy()
    oids, odists, ocounts = O.hitlists(geo, ro, rd, 0.01, cap)
    assert np.array_equal(counts, ocounts)
    assert counts.sum() > 2000, "the sample must see the surfaces"
    full = counts <= cap
    assert full.mean() > 0.98
    for r in np.nonzero(full)[0]:
        c = counts[r]
        assert np.array_equal(ids[r, :c].astype(np.uint32), oids[r, :c])
        assert np.array_equal(dists[r, :c].view(np.uint32), odists[r, :c].view(np.uint32))


def test_fullsize_forward_matches_oracle(frame):
    _, _, nr, pix, geo, ro, rd = frame
    fb = nr.get_framebuffer()
    rgb = nr.get_gaussians().channels.cpu().numpy()
    ref = O.render(geo, rgb, ro, rd, O.make_knobs(t_threshold=0.01))
    img = fb.output_channels[:H, :W].reshape(-1, 3)[torch.from_numpy(pix).cuda()].cpu().numpy()
    T = fb.transmittance[:H, :W].reshape(-1)[torch.from_numpy(pix).cuda()].cpu().numpy()
    F = fb.full_transmittance[:H, :W].reshape(-1)[torch.from_numpy(pix).cuda()].cpu().numpy()
    assert np.abs(img - ref["C"]).max() < 1e-5
    assert np.abs(T - ref["T"]).max() < 1e-5 and np.abs(F - ref["F"]).max() < 1e-5
    assert (ref["ncomp"] > 0).sum() > 100


def test_fullsize_backward_matches_oracle(frame):
    scene, _, nr, pix, geo, ro, rd = frame
    fb = nr.get_framebuffer()
    g = nr.get_gaussians()
    gen = torch.Generator().manual_seed(3)
    gsel = torch.randn(len(pix), 3, generator=gen) * 1e-3
    with torch.no_grad():
        fb.grad_output_channels.zero_()
        flat = fb.grad_output_channels[:H, :W].reshape(-1, 3)
        flat[torch.from_numpy(pix).cuda()] = gsel.cuda()
        fb.grad_output_channels[:H, :W].copy_(flat.reshape(H, W, 3))
        for name in ("grad_mean", "grad_rotation", "grad_scale", "grad_opacity", "grad_channels"):
            getattr(g, name).zero_()
    nr.backward_pass()
    torch.cuda.synchronize()
    nr.device_status()
    rgb = g.channels.cpu().numpy()
    k = O.make_knobs(t_threshold=0.01)
    ref = O.render(geo, rgb, ro, rd, k)
    _, l2w = O.snapshot(scene.mean, scene.rotation, scene.scale, scene.opacity, 0.01, 2.0)
    ob = O.backward(geo, l2w, rgb, scene.rotation, scene.scale, scene.opacity, ro, rd, ref, gsel.numpy(), k,
                    float(W * H))
    touched = np.abs(ob["opacity"]).reshape(-1) > 0
    assert touched.sum() > 500
    # the pruning weight sums every composited hit of the frame whatever the
    # output gradient, so only the gradients are comparable on a sample
    ncomp = nr.hits.n_comp[: W * H][torch.from_numpy(pix).cuda()].cpu().numpy()
    ok = comparable_gaussians(ncomp, ref, geo, ro, rd)
    assert_grads_close([(name, getattr(g, name).cpu(), ob[key]) for name, key in (
        ("grad_opacity", "opacity"), ("grad_scale", "scale"), ("grad_rotation", "rotation"),
        ("grad_channels", "rgb"))], ok)
