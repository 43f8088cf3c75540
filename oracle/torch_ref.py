"""Differentiable PyTorch restatement of the ray tracer for small scenes.

TEST INFRASTRUCTURE ONLY (tests/ and smoke()).  Independent of the C oracle
in formulation: it inverts L2W = R diag(s) numerically (torch.linalg.inv) and
normalises the local direction exactly as the reference test brute force does
(tests/test_multiple_gaussians.py:82-124, tests/utils.py:3-31), in float64,
with autograd for the gradients.  Compositing follows cuda/forward_pass.cu
(front to back by distance, stop once T < t_threshold; hits beyond the stop
enter the detached truncation estimate and the full transmittance).
"""
from __future__ import annotations

import math

import torch

MAX_ALPHA = 0.9999


def build_rotation(q: torch.Tensor) -> torch.Tensor:
    """tests/utils.py:3-20 (quaternion r, x, y, z -> rotation matrix)."""
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = torch.stack([
        1 - 2 * (y * y + z * z), 2 * (x * y - r * z), 2 * (x * z + r * y),
        2 * (x * y + r * z), 1 - 2 * (x * x + z * z), 2 * (y * z - r * x),
        2 * (x * z - r * y), 2 * (y * z + r * x), 1 - 2 * (x * x + y * y),
    ], dim=-1)
    return R.reshape(-1, 3, 3)


def sh_basis(x, y, z):
    """Real SH of degrees 1-3 in closed form (the 15 non-DC terms, in the
    order of the reference's coefficients, sh.cu:11-78)."""
    pi = math.pi
    c1 = math.sqrt(3 / (4 * pi))
    c2a, c2b, c2c = math.sqrt(15 / (4 * pi)), math.sqrt(5 / (16 * pi)), math.sqrt(15 / (16 * pi))
    c3a, c3b, c3c = math.sqrt(35 / (32 * pi)), math.sqrt(105 / (4 * pi)), math.sqrt(21 / (32 * pi))
    c3d, c3e = math.sqrt(7 / (16 * pi)), math.sqrt(105 / (16 * pi))
    xx, yy, zz = x * x, y * y, z * z
    return [
        -c1 * y, c1 * z, -c1 * x,
        c2a * x * y, -c2a * y * z, c2b * (2 * zz - xx - yy), -c2a * x * z, c2c * (xx - yy),
        -c3a * y * (3 * xx - yy), c3b * x * y * z, -c3c * y * (4 * zz - xx - yy),
        c3d * z * (2 * zz - 3 * xx - 3 * yy), -c3c * x * (4 * zz - xx - yy), c3e * z * (xx - yy),
        -c3a * x * (xx - 3 * yy),
    ]


def sh_colors(mean, sh_dc, sh_rest, origin):
    """Degree-3 colours max(C0 dc + sum Y_k c_k + 0.5, 0) seen from origin
    (sh.cu:224-247); differentiable in mean, sh_dc [N,1,3], sh_rest [N,15,3]."""
    delta = mean - origin
    d = delta / delta.norm(dim=1, keepdim=True)
    B = torch.stack([torch.as_tensor(v) for v in sh_basis(d[:, 0], d[:, 1], d[:, 2])], 1)
    return (0.28209479177387814 * sh_dc[:, 0] + (B[:, :, None] * sh_rest).sum(1) + 0.5).clamp(min=0)


def clip_factor(op: torch.Tensor, athr: float, p: float) -> torch.Tensor:
    lr = torch.log(op / athr)
    return torch.where(op > athr, (2 * p * lr.clamp(min=0)) ** (1.0 / (2 * p)), torch.zeros_like(op))


def hit_mask(mean, rotation, scale, opacity, ro, rd, exp_power=2.0, alpha_threshold=0.01):
    """[R, N] bool: which Gaussians each ray hits (the criterion of render)."""
    with torch.no_grad():
        op = torch.sigmoid(opacity.reshape(-1))
        clip = clip_factor(op, alpha_threshold, exp_power)
        q = rotation / rotation.norm(dim=-1, keepdim=True)
        W2L = torch.linalg.inv(build_rotation(q) * torch.exp(scale)[:, None, :])
        lo = torch.einsum("nij,rnj->rni", W2L, ro[:, None, :] - mean[None, :, :])
        ld = torch.einsum("nij,rj->rni", W2L, rd)
        ldn = ld / ld.norm(dim=-1)[..., None]
        hit_local = lo + (-(lo * ldn).sum(-1))[..., None] * ldn
        sq_clip = (hit_local ** 2).sum(-1) / (clip[None, :] ** 2).clamp(min=1e-300)
        behind = (lo * ld).sum(-1) > 0
        return (clip > 0)[None, :] & (~behind) & (sq_clip <= 1.0) & (op > alpha_threshold)[None, :]


def render(mean, rotation, scale, opacity, colors, ro, rd, exp_power=2.0, alpha_threshold=0.01, t_threshold=0.01,
           bg=(0.0, 0.0, 0.0), znear=0.0, zfar=99999.9, fixed=None):
    """Returns (image [R,3], dict of per-ray extras).  All inputs float64.

    fixed: optional per-ray (composited, skipped) lists of Gaussian indices.
    The hit structure (which Gaussians a ray composites, in which order, and
    which enter the truncation estimate) is then taken as given, typically
    from the GPU, and only the smooth weights are evaluated: finite
    differences of this function stay within one piece of the piecewise-
    smooth forward, whose pieces are set by float32 hit decisions and
    distance order that float64 does not reproduce at ties."""
    dt = mean.dtype
    op = torch.sigmoid(opacity.reshape(-1))
    clip = clip_factor(op.detach(), alpha_threshold, exp_power)
    visible = (clip > 0).detach()
    q = rotation / rotation.norm(dim=-1, keepdim=True)
    R = build_rotation(q)
    s = torch.exp(scale)
    # the clip factor only scales the proxy for the hit decision; alpha uses
    # x*clip, i.e. the unclipped local frame (kernel.cu:40-42, shaders.cu:50)
    A = R * s[:, None, :]
    W2L = torch.linalg.inv(A)  # [N,3,3]
    rel = ro[:, None, :] - mean[None, :, :]  # [R,N,3]
    lo = torch.einsum("nij,rnj->rni", W2L, rel)  # local origin (unclipped frame)
    ld = torch.einsum("nij,rj->rni", W2L, rd)
    norm = ld.norm(dim=-1)
    ldn = ld / norm[..., None]
    t_local = -(lo * ldn).sum(-1)
    hit_local = lo + t_local[..., None] * ldn  # = x * clip in the clipped frame
    dist = t_local / norm
    sq_clip = (hit_local.detach() ** 2).sum(-1) / (clip[None, :] ** 2).clamp(min=1e-300)
    behind = ((lo * ld).sum(-1) > 0).detach()
    hit = visible[None, :] & (~behind) & (sq_clip <= 1.0) & (op.detach() > alpha_threshold)[None, :]
    qn = (hit_local ** 2).sum(-1)
    g = torch.exp(-(qn ** exp_power) / (2 * exp_power))
    alpha = MAX_ALPHA * g * op[None, :]

    Rn = ro.shape[0]
    out = torch.zeros(Rn, 3, dtype=dt)
    extras = {"T": [], "F": [], "ncomp": [], "order": []}
    bgt = torch.tensor(bg, dtype=dt)
    for r in range(Rn):
        if fixed is not None:
            comp, skipped = fixed[r]
            T = torch.ones((), dtype=dt)
            pix = torch.zeros(3, dtype=dt)
            for gid in comp:
                a = alpha[r, gid]
                pix = pix + colors[gid] * (T * a)
                T = T * (1 - a)
            skip_c = torch.zeros(3, dtype=dt)
            skip_a = torch.zeros((), dtype=dt)
            skip_p = torch.ones((), dtype=dt)
            for gid in skipped:
                ad = alpha[r, gid].detach()
                skip_c = skip_c + colors[gid].detach() * ad
                skip_a = skip_a + ad
                skip_p = skip_p * (1 - ad)
            F = T * skip_p
            est = skip_c / skip_a if float(skip_a) > 0 else torch.zeros(3, dtype=dt)
            out[r] = pix + (T - F) * est + bgt * F
            extras["T"].append(float(T))
            extras["F"].append(float(F))
            extras["ncomp"].append(len(comp))
            extras["order"].append(list(comp))
            continue
        idx = torch.nonzero(hit[r]).reshape(-1)
        if idx.numel() == 0:
            out[r] = bgt
            extras["T"].append(1.0)
            extras["F"].append(1.0)
            extras["ncomp"].append(0)
            extras["order"].append([])
            continue
        d = dist[r, idx].detach()
        order = sorted(range(idx.numel()), key=lambda i: (float(d[i]), int(idx[i])))
        T = torch.ones((), dtype=dt)
        pix = torch.zeros(3, dtype=dt)
        skip_c = torch.zeros(3, dtype=dt)
        skip_a = torch.zeros((), dtype=dt)
        skip_p = torch.ones((), dtype=dt)
        stopped = False
        nc = 0
        comp_ids = []
        for i in order:
            gid = int(idx[i])
            a = alpha[r, gid]
            di = float(d[i])
            if not stopped and di > znear and di < zfar:
                pix = pix + colors[gid] * (T * a)
                T = T * (1 - a)
                nc += 1
                comp_ids.append(gid)
                if float(T) < t_threshold:
                    stopped = True
            else:
                ad = a.detach()
                skip_c = skip_c + colors[gid].detach() * ad
                skip_a = skip_a + ad
                skip_p = skip_p * (1 - ad)
        F = T * skip_p
        est = skip_c / skip_a if float(skip_a) > 0 else torch.zeros(3, dtype=dt)
        pix = pix + (T - F) * est + bgt * F
        out[r] = pix
        extras["T"].append(float(T))
        extras["F"].append(float(F))
        extras["ncomp"].append(nc)
        extras["order"].append(comp_ids)
    return out, extras
