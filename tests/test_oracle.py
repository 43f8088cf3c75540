"""Pins the CPU oracle (oracle/gray_oracle.c) before it is trusted as the GPU
checker: closed forms, invariants, the reference tests' known-answer scenes
(tests/test_single_gaussian.py, tests/test_multiple_gaussians.py of the
reference, restated), autograd of an independent float64 formulation, and
finite differences.  CPU only."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import torch_ref

MAX_ALPHA = 0.9999


# ---------------------------------------------------------------------------
# spherical harmonics: reference complex recursion vs closed form
# ---------------------------------------------------------------------------
closed_form_basis = torch_ref.sh_basis


def test_closed_form_is_orthonormal():
    # Gauss-Legendre x uniform-phi quadrature is exact for these polynomials
    ct, wt = np.polynomial.legendre.leggauss(16)
    phi = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    Y = []
    W = []
    for c, w in zip(ct, wt):
        s = math.sqrt(1 - c * c)
        for p in phi:
            x, y, z = s * math.cos(p), s * math.sin(p), c
            Y.append(np.array([1 / (2 * math.sqrt(math.pi))] + closed_form_basis(x, y, z)))
            W.append(w * 2 * np.pi / len(phi))
    Y = np.array(Y)
    W = np.array(W)
    gram = (Y * W[:, None]).T @ Y
    assert np.abs(gram - np.eye(16)).max() < 1e-12


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_sh_recursion_matches_closed_form(oracle_lib, degree):
    rng = np.random.default_rng(degree)
    kc = (degree + 1) ** 2 - 1
    for _ in range(200):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        b, _, _, _ = O.sh_basis(degree, *v)
        ref = np.array(closed_form_basis(*v))[:kc]
        assert np.abs(b[:kc] - ref).max() < 2e-6


def test_sh_recursion_partials_tangential(oracle_lib):
    # partials of the recursion agree with the closed form along the sphere
    rng = np.random.default_rng(7)
    for _ in range(100):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        _, bx, by, bz = O.sh_basis(3, *v)
        t = np.cross(v, rng.normal(size=3))
        t /= np.linalg.norm(t)
        h = 1e-5
        fd = (np.array(closed_form_basis(*(v + h * t))) - np.array(closed_form_basis(*(v - h * t)))) / (2 * h)
        an = bx * t[0] + by * t[1] + bz * t[2]
        assert np.abs(an - fd).max() < 5e-5


def test_sh_forward_backward_against_autograd(oracle_lib):
    rng = np.random.default_rng(3)
    n = 64
    mean = rng.normal(size=(n, 3)).astype(np.float32)
    dc = rng.normal(0.5, 0.2, size=(n, 1, 3)).astype(np.float32)
    rest = rng.normal(0, 0.3, size=(n, 15, 3)).astype(np.float32)
    origin = np.array([0.3, -2.0, 4.0], np.float32)
    rgb = O.sh_forward(mean, dc, rest, origin, 3)
    # float64 autograd of the closed form
    m = torch.tensor(mean, dtype=torch.float64, requires_grad=True)
    d = torch.tensor(dc, dtype=torch.float64, requires_grad=True)
    r = torch.tensor(rest, dtype=torch.float64, requires_grad=True)
    delta = m - torch.tensor(origin, dtype=torch.float64)
    dirs = delta / delta.norm(dim=1, keepdim=True)
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    B = torch.stack([torch.as_tensor(v) for v in closed_form_basis(x, y, z)], 1)
    out = (0.28209479177387814 * d[:, 0] + (B[:, :, None] * r).sum(1) + 0.5).clamp(min=0)
    assert np.abs(out.detach().numpy() - rgb).max() < 1e-5
    g = torch.tensor(rng.normal(size=(n, 3)))
    (out * g).sum().backward()
    gdc, grest, gmean = O.sh_backward(mean, rest, origin, rgb, g.numpy().astype(np.float32), None, 3)
    assert np.abs(gdc - d.grad.numpy()).max() < 1e-5
    assert np.abs(grest - r.grad.numpy()).max() < 1e-5
    assert np.abs(gmean - m.grad.numpy()).max() < 1e-4 * max(1.0, np.abs(m.grad.numpy()).max())


# ---------------------------------------------------------------------------
# Adam: reference bias correction uses step+1 (adam.cu:16-17)
# ---------------------------------------------------------------------------
def test_adam_matches_torch_with_step_offset(oracle_lib):
    rng = np.random.default_rng(0)
    p0 = rng.normal(size=257).astype(np.float32)
    p = p0.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp = torch.tensor(p0.astype(np.float64), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-15)
    opt.state[tp] = {"step": torch.tensor(1.0), "exp_avg": torch.zeros_like(tp), "exp_avg_sq": torch.zeros_like(tp)}
    for step in range(1, 6):
        g = rng.normal(size=p.shape).astype(np.float32)
        gg = g.copy()
        O.adam(p, gg, m, v, 1e-2, 0.9, 0.999, 1e-15, step, zero_grads=True)
        assert np.all(gg == 0)
        tp.grad = torch.tensor(g.astype(np.float64))
        opt.step()
        assert np.abs(p - tp.detach().numpy()).max() < 1e-5


# ---------------------------------------------------------------------------
# camera and snapshot closed forms
# ---------------------------------------------------------------------------
def test_pinhole_rays_closed_form(oracle_lib):
    W, H, fov = 64, 48, 1.1
    c2w = np.diag([-1.0, 1.0, 1.0]).astype(np.float32)
    ro, rd = O.pinhole_rays(W, H, fov, np.array([0, 0, 1], np.float32), c2w)
    assert np.allclose(np.linalg.norm(rd, axis=1), 1.0, atol=1e-6)
    v = math.tan(fov / 2)
    a = W / H
    for (i, j) in [(0, 0), (W - 1, 0), (W // 2, H // 2), (5, 17)]:
        x = a * v * (2 * (i + 0.5) / W - 1)
        y = v * (1 - 2 * (j + 0.5) / H)
        d = c2w[:, 0] * x + c2w[:, 1] * y - c2w[:, 2]
        d /= np.linalg.norm(d)
        assert np.abs(rd[j * W + i] - d).max() < 1e-6


def test_snapshot_clip_invariant(oracle_lib):
    # at the clip boundary |x| = 1 the kernel times opacity equals alpha_threshold
    rng = np.random.default_rng(1)
    n = 50
    for p in (1.0, 2.0, 3.0, 2.5):
        logit = rng.normal(size=(n, 1)).astype(np.float32) + 1.0
        geo, _ = O.snapshot(np.zeros((n, 3)), np.tile([1, 0, 0, 0], (n, 1)), np.zeros((n, 3)), logit,
                            alpha_threshold=0.01, exp_power=p)
        op, clip = geo[:, 12], geo[:, 13]
        vis = op > 0
        g = np.exp(-(clip[vis].astype(np.float64) ** 2) ** p / (2 * p))
        assert np.abs(op[vis] * g - 0.01).max() < 1e-6


def test_snapshot_world_to_local(oracle_lib):
    rng = np.random.default_rng(2)
    n = 40
    mean = rng.normal(size=(n, 3))
    q = rng.normal(size=(n, 4))
    s = rng.normal(-1, 0.5, size=(n, 3))
    o = rng.normal(size=(n, 1)) + 2
    geo, l2w = O.snapshot(mean, q, s, o, 0.01, 2.0)
    R = torch_ref.build_rotation(torch.tensor(q / np.linalg.norm(q, axis=1, keepdims=True)))
    clip = geo[:, 13].astype(np.float64)
    A = R.numpy() * (np.exp(s) * clip[:, None])[:, None, :]
    assert np.abs(l2w[:, :9].reshape(n, 3, 3) - A).max() < 1e-5 * np.abs(A).max()
    M = geo[:, :9].reshape(n, 3, 3).astype(np.float64)
    eye = np.einsum("nij,njk->nik", M, A)
    assert np.abs(eye - np.eye(3)).max() < 1e-5


# ---------------------------------------------------------------------------
# known-answer scenes of the reference tests (restated)
# ---------------------------------------------------------------------------
def mock_camera_rays(oracle_lib, n=32):
    c2w = np.diag([-1.0, 1.0, 1.0]).astype(np.float32)  # -R with column 0 negated, R = -I
    return O.pinhole_rays(n, n, 1.5, np.array([0.0, 0.0, 1.0], np.float32), c2w)


def target_image():
    from PIL import Image
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "target_1.png")
    return np.asarray(Image.open(path).convert("RGB"), dtype=np.float32) / 255.0  # [H,W,3]


def run_case(oracle_lib, mean, log_scale, colors, rot, logit, bg, athr, tthr, p=2.0):
    ro, rd = mock_camera_rays(oracle_lib)
    geo, l2w = O.snapshot(mean, rot, log_scale, logit, athr, p)
    k = O.make_knobs(exp_power=p, alpha_threshold=athr, t_threshold=tthr, bg=bg)
    fwd = O.render(geo, colors, ro, rd, k, mode=0)
    # float64 autograd restatement
    t = lambda a: torch.tensor(np.asarray(a, np.float64), requires_grad=True)  # noqa: E731
    tm, ts, tc, tq, to = t(mean), t(log_scale), t(colors), t(rot), t(logit)
    img, extra = torch_ref.render(tm, tq, ts, to, tc, torch.tensor(ro, dtype=torch.float64),
                                  torch.tensor(rd, dtype=torch.float64), p, athr, tthr, bg)
    target = target_image().reshape(-1, 3)
    gout = np.sign(img.detach().numpy() - target) / img.numel()
    (img * torch.tensor(gout)).sum().backward()
    bwd = O.backward(geo, l2w, colors, rot, log_scale, logit, ro, rd, fwd, gout.astype(np.float32), k, 32 * 32)
    return fwd, img.detach().numpy(), bwd, {"mean": tm.grad, "rotation": tq.grad, "scale": ts.grad,
                                            "opacity": to.grad, "rgb": tc.grad}


def assert_grads_close(bwd, ref, tol=1e-4):
    for key in ("mean", "rotation", "scale", "opacity", "rgb"):
        a = bwd[key].reshape(-1)
        b = ref[key].numpy().reshape(-1)
        assert np.abs(a - b).max() < tol, (key, np.abs(a - b).max())


@pytest.mark.parametrize("athr", [0.01, 0.3])
def test_single_gaussian_known_answer(oracle_lib, athr):
    mean = np.array([[-0.05, -0.05, 0.0]])
    log_scale = np.log([[0.1, 0.3, 0.3]])
    colors = np.array([[1.0, 0.0, 0.0]])
    rot = np.array([[-0.5, 0.0, 0.0, 1.0]])
    logit = np.array([[0.0]])  # logit(0.5)
    bg = (0.1, 0.1, 0.1)
    fwd, ref_img, bwd, ref_g = run_case(oracle_lib, mean, log_scale, colors, rot, logit, bg, athr, 0.03)
    # closed form C = alpha c + (1 - alpha) bg
    alpha = (fwd["C"][:, 0] - 0.1) / 0.9
    assert np.abs(fwd["C"][:, 1] - 0.1 * (1 - alpha)).max() < 1e-6
    assert np.abs(fwd["C"] - ref_img).max() < 1e-5
    assert np.abs(fwd["C"] - ref_img).mean() < 1e-6
    assert (fwd["ncomp"] > 0).sum() > 20
    assert_grads_close(bwd, ref_g)


@pytest.mark.parametrize("tthr", [0.0, 0.5])
def test_multiple_gaussians_known_answer(oracle_lib, tthr):
    mean = np.array([[0.05, 0.05, -0.25], [-0.05, -0.05, 0.25], [0.0, 0.0, 0.0]])
    log_scale = np.log([[1.0, 0.2, 0.1], [0.1, 0.3, 0.3], [0.4, 0.4, 0.4]])
    colors = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    rot = np.array([[1.0, 0.0, 0.0, 1.0]] * 3)
    logit = np.log(np.array([[0.5], [0.7], [0.7]]) / (1 - np.array([[0.5], [0.7], [0.7]])))
    fwd, ref_img, bwd, ref_g = run_case(oracle_lib, mean, log_scale, colors, rot, logit, (0, 0, 0), 0.01, tthr)
    assert np.abs(fwd["C"] - ref_img).max() < 1e-5
    if tthr == 0.5:
        assert (fwd["est"].sum(1) > 0).any(), "truncation estimate must be exercised"
    assert_grads_close(bwd, ref_g)


def random_small(seed, n=48, scale_mu=-1.6):
    rng = np.random.default_rng(seed)
    mean = rng.normal(0, 0.35, size=(n, 3))
    q = rng.normal(size=(n, 4))
    s = rng.normal(scale_mu, 0.4, size=(n, 3))
    logit = rng.normal(size=(n, 1))
    colors = rng.uniform(size=(n, 3))
    return mean, q, s, logit, colors


@pytest.mark.parametrize("seed,tthr,p", [(0, 0.03, 2.0), (1, 0.0, 2.0), (2, 0.3, 1.0), (3, 0.03, 3.0)])
def test_random_scene_against_autograd(oracle_lib, seed, tthr, p):
    mean, q, s, logit, colors = random_small(seed)
    ro, rd = O.pinhole_rays(20, 20, 1.0, np.array([0, 0, 2.5], np.float32), np.diag([1.0, 1.0, 1.0]))
    geo, l2w = O.snapshot(mean, q, s, logit, 0.01, p)
    k = O.make_knobs(exp_power=p, alpha_threshold=0.01, t_threshold=tthr, bg=(0.2, 0.3, 0.1))
    fwd = O.render(geo, colors, ro, rd, k, mode=0)
    lit = O.render(geo, colors, ro, rd, k, mode=1)
    # the restated truncation bookkeeping equals the literal reference arithmetic
    assert np.abs(fwd["C"] - lit["C"]).max() < 2e-5
    assert np.abs(fwd["F"] - lit["F"]).max() < 2e-6
    t = lambda a: torch.tensor(np.asarray(a, np.float64), requires_grad=True)  # noqa: E731
    tm, ts, tc, tq, to = t(mean), t(s), t(colors), t(q), t(logit)
    img, extra = torch_ref.render(tm, tq, ts, to, tc, torch.tensor(ro, dtype=torch.float64),
                                  torch.tensor(rd, dtype=torch.float64), p, 0.01, tthr, (0.2, 0.3, 0.1))
    assert np.abs(fwd["C"] - img.detach().numpy()).max() < 5e-5
    assert np.array_equal(fwd["ncomp"], np.array(extra["ncomp"]))
    gout = np.random.default_rng(seed + 100).normal(size=img.shape) / img.numel()
    (img * torch.tensor(gout)).sum().backward()
    bwd = O.backward(geo, l2w, colors, q, s, logit, ro, rd, fwd, gout.astype(np.float32), k, 400.0)
    for key, ref in (("mean", tm.grad), ("rotation", tq.grad), ("scale", ts.grad), ("opacity", to.grad),
                     ("rgb", tc.grad)):
        a, b = bwd[key].reshape(-1), ref.numpy().reshape(-1)
        scale = max(np.abs(b).max(), 1e-8)
        assert np.abs(a - b).max() < 1e-3 * scale + 1e-7, (key, np.abs(a - b).max(), scale)


@pytest.mark.parametrize("tthr", [0.0, 0.3])
def test_fixed_structure_render_equals_free_render(tthr):
    """torch_ref.render with the hit structure given (composited order and
    skipped hits) equals the render that decides it, value and gradient, so
    finite differences with the structure held fixed differentiate the same
    function within a piece."""
    mean, q, s, logit, colors = random_small(5)
    ro, rd = O.pinhole_rays(12, 12, 1.0, np.array([0, 0, 2.5], np.float32), np.eye(3))
    ro, rd = torch.tensor(ro, dtype=torch.float64), torch.tensor(rd, dtype=torch.float64)
    t = lambda a: torch.tensor(np.asarray(a, np.float64), requires_grad=True)  # noqa: E731
    outs = []
    free = None
    for fixed_mode in (False, True):
        tm, ts, tc, tq, to = t(mean), t(s), t(colors), t(q), t(logit)
        fixed = None
        if fixed_mode:
            hits = torch_ref.hit_mask(tm, tq, ts, to, ro, rd)
            fixed = []
            for r in range(ro.shape[0]):
                comp = free[1]["order"][r]
                skipped = [int(g) for g in torch.nonzero(hits[r]).reshape(-1) if int(g) not in comp]
                fixed.append((comp, skipped))
        img, extra = torch_ref.render(tm, tq, ts, to, tc, ro, rd, 2.0, 0.01, tthr, (0.2, 0.3, 0.1), fixed=fixed)
        if not fixed_mode:
            free = (img, extra)
        gout = torch.tensor(np.random.default_rng(1).normal(size=img.shape))
        (img * gout).sum().backward()
        outs.append((img.detach(), [x.grad.clone() for x in (tm, ts, tc, tq, to)]))
    assert sum(len(o) for o in free[1]["order"]) > 50
    # equal up to the order of the (float64) truncation-estimate sums
    assert torch.allclose(outs[0][0], outs[1][0], rtol=1e-12, atol=1e-15)
    for a, b in zip(outs[0][1], outs[1][1]):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-15)


def test_invariants(oracle_lib):
    mean, q, s, logit, colors = random_small(11, n=96)
    ro, rd = O.pinhole_rays(16, 16, 1.0, np.array([0, 0, 2.5], np.float32), np.eye(3))
    geo, _ = O.snapshot(mean, q, s, logit, 0.01, 2.0)
    k = O.make_knobs(t_threshold=0.0)
    fwd = O.render(geo, colors, ro, rd, k)
    ids, dists, counts = O.hitlists(geo, ro, rd)
    for r in range(ro.shape[0]):
        n = fwd["ncomp"][r]
        Tb = fwd["Tb"][r, :n]
        assert np.all(np.diff(Tb) <= 0), "T must not increase along the composited list"
        assert n == counts[r], "with t_threshold 0 every hit is composited"
        assert np.all(np.diff(dists[r, :counts[r]]) >= 0)
        assert list(fwd["ids"][r, :n]) == list(ids[r, :n])
    # alpha range: every composited alpha in [MAX_ALPHA * athr * (1-1e-6), MAX_ALPHA]
    # (checked through T ratios)
    ratio = fwd["Tb"][:, 1:] / np.maximum(fwd["Tb"][:, :-1], 1e-30)
    mask = np.arange(ratio.shape[1])[None, :] < (fwd["ncomp"][:, None] - 1)
    a = 1 - ratio[mask]
    assert a.min() >= MAX_ALPHA * 0.01 * (1 - 1e-3) and a.max() <= MAX_ALPHA * (1 + 1e-6)
    # empty scene renders the background with T = F = 1
    empty = O.render(np.zeros((0, 16), np.float32), np.zeros((0, 3), np.float32), ro, rd,
                     O.make_knobs(bg=(0.25, 0.5, 0.75)))
    assert np.allclose(empty["C"], [0.25, 0.5, 0.75]) and np.all(empty["T"] == 1) and np.all(empty["F"] == 1)


def test_truncation_and_batches(oracle_lib):
    # 80 coincident-axis Gaussians along the ray: several batches of 32, then the estimate
    n = 80
    mean = np.zeros((n, 3))
    mean[:, 2] = -np.linspace(0.1, 4.0, n)
    q = np.tile([1.0, 0, 0, 0], (n, 1))
    s = np.full((n, 3), np.log(0.05))
    logit = np.full((n, 1), -2.5)  # opacity ~0.076
    colors = np.random.default_rng(0).uniform(size=(n, 3))
    ro = np.zeros((1, 3), np.float32)
    rd = np.array([[0, 0, -1.0]], np.float32)
    geo, _ = O.snapshot(mean, q, s, logit, 0.01, 2.0)
    for tthr in (0.0, 0.05):
        fwd = O.render(geo, colors, ro, rd, O.make_knobs(t_threshold=tthr), mode=0)
        lit = O.render(geo, colors, ro, rd, O.make_knobs(t_threshold=tthr), mode=1)
        assert fwd["nhits"][0] == n
        if tthr == 0.0:
            assert fwd["ncomp"][0] == n
        else:
            assert 32 < fwd["ncomp"][0] < n and fwd["T"][0] < tthr
        assert np.abs(fwd["C"] - lit["C"]).max() < 1e-5
        # closed form along the axis: alpha_i = MAX_ALPHA * op, composite in order
        op = 1 / (1 + np.exp(2.5))
        a = MAX_ALPHA * op
        T = 1.0
        c = np.zeros(3)
        k = 0
        for i in range(n):
            c += colors[i] * T * a
            T *= 1 - a
            k += 1
            if T < tthr:
                break
        rest = colors[k:]
        F = T * (1 - a) ** len(rest)
        est = rest.mean(0) if len(rest) else np.zeros(3)
        c += (T - F) * est
        assert np.abs(fwd["C"][0] - c).max() < 1e-5


def test_finite_differences_of_reference_formulation():
    # pins the float64 restatement's autograd against central differences
    mean, q, s, logit, colors = random_small(5, n=12, scale_mu=-1.2)
    ro = torch.tensor([[0.05, -0.02, 2.5], [0.1, 0.1, 2.5], [-0.2, 0.05, 2.5]], dtype=torch.float64)
    rd = torch.tensor([[0.0, 0.0, -1.0], [-0.03, -0.04, -1.0], [0.08, -0.02, -1.0]], dtype=torch.float64)
    rd = rd / rd.norm(dim=1, keepdim=True)
    params = [torch.tensor(a, dtype=torch.float64) for a in (mean, q, s, logit, colors)]

    def f(ps):
        img, _ = torch_ref.render(ps[0], ps[1], ps[2], ps[3], ps[4], ro, rd, 2.0, 0.01, 0.03, (0.1, 0.2, 0.3))
        w = torch.linspace(0.5, 1.5, img.numel(), dtype=torch.float64).reshape(img.shape)
        return (img * w).sum()

    ps = [p.clone().requires_grad_() for p in params]
    f(ps).backward()
    h = 1e-6
    rng = np.random.default_rng(0)
    checked = 0
    for _ in range(40):
        pi = int(rng.integers(0, 5))
        idx = tuple(int(rng.integers(0, d)) for d in params[pi].shape)
        if abs(float(ps[pi].grad[idx])) < 1e-9:
            continue
        plus = [p.clone() for p in params]
        minus = [p.clone() for p in params]
        plus[pi][idx] += h
        minus[pi][idx] -= h
        fd = (f(plus) - f(minus)).item() / (2 * h)
        an = float(ps[pi].grad[idx])
        assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (pi, idx, fd, an)
        checked += 1
    assert checked >= 10


# ---------------------------------------------------------------------------
# fp16 hit storage (the reference keeps distances and alphas as fp16,
# per_pixel_linked_list.h:28-29)
# ---------------------------------------------------------------------------
def test_fp16_round_matches_numpy(oracle_lib):
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100000) * 10.0 ** rng.uniform(-9, 5, 100000)).astype(np.float32)
    edges = np.array([65504, 65519.99, 65520, 65535, 1e5, -65520, 6.1035156e-05, 6.1e-5, 5.96e-8, 2.98e-8,
                      2.99e-8, 1e-9, 0.0, -0.0, np.inf, -np.inf, 1 + 2**-11, 1 + 3 * 2**-11, 2049, 2051, 4097],
                     np.float32)
    x = np.concatenate([x, edges])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(O.fp16_round(x).view(np.uint32), ref.view(np.uint32))


def test_fp16_hits_single_gaussian_closed_form(oracle_lib):
    # one hit: compositing uses the fp16 alpha (T = 1 - a16), the full
    # transmittance the fp32 one (F = 1 - a32, unchanged by the mode), the skip
    # set is empty, so C = a16 c + bg F; depth is the fp16 distance
    mean = np.array([[-0.05, -0.05, 0.0]])
    log_scale = np.log([[0.1, 0.3, 0.3]])
    colors = np.array([[0.7, 0.2, 0.4]], np.float32)
    rot = np.array([[-0.5, 0.0, 0.0, 1.0]])
    logit = np.array([[0.3]])
    bg = (0.1, 0.2, 0.3)
    ro, rd = mock_camera_rays(oracle_lib)
    geo, _ = O.snapshot(mean, rot, log_scale, logit, 0.01, 2.0)
    f32 = O.render(geo, colors, ro, rd, O.make_knobs(t_threshold=0.03, bg=bg, render_depth=True))
    f16 = O.render(geo, colors, ro, rd, O.make_knobs(t_threshold=0.03, bg=bg, render_depth=True, fp16_hits=True))
    hit = f32["ncomp"] > 0
    assert hit.sum() > 20 and np.array_equal(f16["ncomp"], f32["ncomp"])
    assert np.array_equal(f16["F"], f32["F"])
    a32 = 1.0 - f32["F"][hit].astype(np.float64)
    a16 = 1.0 - f16["T"][hit].astype(np.float64)
    assert np.abs(a16 - O.fp16_round(a32)).max() < 2e-7
    assert (f16["T"][hit] != f32["T"][hit]).any(), "the rounding must change the composited alpha"
    expect = a16[:, None] * colors[0] + np.array(bg) * f32["F"][hit][:, None]
    assert np.abs(f16["C"][hit] - expect).max() < 1e-6
    assert np.abs(f16["D"][hit] - O.fp16_round(f32["D"][hit])).max() <= 1e-6 * np.abs(f32["D"][hit]).max()
    assert np.array_equal(f16["C"][~hit], f32["C"][~hit])


def test_fp16_hits_literal_and_skip_set_agree(oracle_lib):
    # the skip-set formulation and the reference's literal one (all-hit sums
    # minus composited, forward_pass.cu:53-57,85-88) agree in fp16 mode too
    rng = np.random.default_rng(5)
    n = 300
    mean = rng.normal(0, 0.4, (n, 3))
    log_scale = rng.normal(-2.5, 0.4, (n, 3))
    rot = rng.normal(size=(n, 4))
    logit = rng.normal(size=(n, 1))
    colors = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    ro, rd = mock_camera_rays(oracle_lib)
    geo, _ = O.snapshot(mean, rot, log_scale, logit, 0.01, 2.0)
    k = O.make_knobs(t_threshold=0.2, bg=(0.3, 0.1, 0.2), fp16_hits=True)
    a = O.render(geo, colors, ro, rd, k, mode=0)
    b = O.render(geo, colors, ro, rd, k, mode=1)
    assert (a["nhits"] > a["ncomp"]).sum() > 50
    assert np.abs(a["C"] - b["C"]).max() < 1e-5
    assert np.abs(a["F"] - b["F"]).max() < 1e-5
    f32 = O.render(geo, colors, ro, rd, O.make_knobs(t_threshold=0.2, bg=(0.3, 0.1, 0.2)), mode=0)
    assert np.abs(a["C"] - f32["C"]).max() > 1e-5


# ---------------------------------------------------------------------------
# ellipsoid display mode (ellipsoid_viewer.cu)
# ---------------------------------------------------------------------------
def test_ellipsoid_mode_sphere_closed_form(oracle_lib):
    # a sphere of radius rho = exp(s) * clip at the origin, rays along -z
    # from (b, 0, 5): entry t = 5 - sqrt(rho^2 - b^2), world normal
    # (b, 0, sqrt(rho^2 - b^2)) / rho, shade max(0.5, cos) (c + 0.15)
    s, logit = -0.5, 1.0
    op = 1.0 / (1.0 + np.exp(-logit))
    clip = (4.0 * np.log(op / 0.01)) ** 0.25
    rho = np.exp(s) * clip
    geo, _ = O.snapshot(np.zeros((1, 3)), np.array([[1.0, 0.0, 0.0, 0.0]]), np.full((1, 3), s),
                        np.array([[logit]]), 0.01, 2.0)
    rgb = np.array([[0.2, 0.5, 0.7]], np.float32)
    bs = np.array([0.0, 0.3 * rho, 0.9 * rho, 1.1 * rho])
    ro = np.stack([bs, np.zeros(4), np.full(4, 5.0)], 1).astype(np.float32)
    rd = np.tile(np.array([[0.0, 0.0, -1.0]], np.float32), (4, 1))
    bg = (0.05, 0.1, 0.15)
    C, D = O.ellipsoids(geo, rgb, ro, rd, 0.0, 0.0, 99999.9, bg)
    h = np.sqrt(np.maximum(rho ** 2 - bs ** 2, 0.0))
    for i in range(3):
        assert abs(D[i] - (5.0 - h[i])) < 2e-6 * 5
        assert np.abs(C[i] - max(0.5, h[i] / rho) * (rgb[0] + 0.15)).max() < 2e-6
    assert D[3] == 0.0 and np.array_equal(C[3], np.array(bg, np.float32))
    # znear past the entry point: the exit point, facing away (shade 0.5)
    C, D = O.ellipsoids(geo, rgb, ro[:1], rd[:1], 0.0, 5.0, 99999.9, bg)
    assert abs(D[0] - (5.0 + rho)) < 1e-5 and np.abs(C[0] - 0.5 * (rgb[0] + 0.15)).max() < 2e-6
    # below the opacity threshold: background
    C, D = O.ellipsoids(geo, rgb, ro[:1], rd[:1], op + 0.01, 0.0, 99999.9, bg)
    assert D[0] == 0.0 and np.array_equal(C[0], np.array(bg, np.float32))
    # two spheres on the ray: the nearer one
    geo2, _ = O.snapshot(np.array([[0.0, 0.0, -2.0], [0.0, 0.0, 0.0]]), np.array([[1.0, 0, 0, 0]] * 2),
                         np.full((2, 3), s), np.array([[logit], [logit]]), 0.01, 2.0)
    C, D = O.ellipsoids(geo2, np.array([[1.0, 0, 0], [0, 1.0, 0]], np.float32), ro[:1], rd[:1])
    assert abs(D[0] - (5.0 - rho)) < 1e-5 and C[0, 1] > C[0, 0]


# ---------------------------------------------------------------------------
# sub-pixel ray jitter (camera.h:17-36, random.h)
# ---------------------------------------------------------------------------
def test_jitter_generators_known_answers(oracle_lib):
    # TEA's published test vector: zero key, zero plaintext, 32 cycles -> 0x41ea3a0a (first word)
    assert O.tea(0, 0, 32, [0, 0, 0, 0]) == 0x41EA3A0A
    # Numerical Recipes LCG from 0
    seq, x = [], 0
    for _ in range(3):
        x = O.lcg(x)
        seq.append(x)
    assert seq == [1013904223, 1196435762, 3519870697]


def test_jittered_rays_stay_within_the_pixel(oracle_lib):
    W, H, fov = 40, 30, 1.0
    origin = np.array([0.0, 0.0, 3.0], np.float32)
    c2w = np.diag([1.0, 1.0, 1.0]).astype(np.float32)
    ro, rd = O.pinhole_rays(W, H, fov, origin, c2w)
    ro1, rd1 = O.pinhole_rays_jittered(W, H, fov, origin, c2w, 7)
    _, rd2 = O.pinhole_rays_jittered(W, H, fov, origin, c2w, 8)
    assert np.array_equal(ro, ro1)
    assert not np.allclose(rd1, rd2)
    # back to pixel coordinates: the offsets lie in [-0.5, 0.5) and average 0
    vs = np.tan(fov / 2)
    def pix(d):
        x = d[:, 0] / -d[:, 2]
        y = d[:, 1] / -d[:, 2]
        return (x / (vs * W / H) + 1) * W / 2 - 0.5, (1 - y / vs) * H / 2 - 0.5
    px0, py0 = pix(rd)
    px1, py1 = pix(rd1)
    jx, jy = px1 - px0, py1 - py0
    assert np.all(np.abs(jx) <= 0.5 + 1e-4) and np.all(np.abs(jy) <= 0.5 + 1e-4)
    assert abs(jx.mean()) < 0.05 and abs(jy.mean()) < 0.05 and jx.std() > 0.2
