"""Float64 restatement of the training loss of train.py:255-260.

TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU baseline leg).

    loss = (1 - lambda_ssim) * F.l1_loss(render, target) + lambda_ssim * (1 - fused_ssim(render[None], target[None]))

fused_ssim (git a7c48d6, pinned in the reference's uv.lock:705-707) is a
third-party extension absent here; it computes the SSIM of 3D Gaussian
Splatting's loss_utils.ssim: a per-channel 2-D convolution with the outer
product of an 11-tap Gaussian (sigma 1.5, normalised to sum 1), zero padding
of 5 ("same"), C1 = 0.01^2, C2 = 0.03^2, and the mean of the SSIM map ("valid"
crops 5 pixels from each border before the mean).  The reference has no test
of it; the restatement is pinned in tests/test_loss.py against
scipy.ndimage.gaussian_filter (an independent windowing and padding route),
against SSIM(x, x) = 1, symmetry, and against fp64 central differences.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

C1 = 0.01 ** 2
C2 = 0.03 ** 2
RADIUS = 5


def window(dtype=torch.float64) -> torch.Tensor:
    t = torch.arange(-RADIUS, RADIUS + 1, dtype=torch.float64)
    g = torch.exp(-t * t / (2 * 1.5 ** 2))
    return (g / g.sum()).to(dtype)


def ssim_map(x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
    """x, y [B, C, H, W] -> the per-pixel SSIM map (zero padding)."""
    c = x.shape[1]
    g = window(x.dtype).to(x.device)
    w = (g[:, None] * g[None, :]).expand(c, 1, 2 * RADIUS + 1, 2 * RADIUS + 1).contiguous()
    conv = lambda t: F.conv2d(t, w, padding=RADIUS, groups=c)  # noqa: E731
    mu1, mu2 = conv(x), conv(y)
    s11 = conv(x * x) - mu1 * mu1
    s22 = conv(y * y) - mu2 * mu2
    s12 = conv(x * y) - mu1 * mu2
    return ((2 * mu1 * mu2 + C1) * (2 * s12 + C2)) / ((mu1 * mu1 + mu2 * mu2 + C1) * (s11 + s22 + C2))


def ssim(x: torch.Tensor, y: torch.Tensor, padding: str = "same") -> torch.Tensor:
    """Mean SSIM as fused_ssim(x, y, padding) returns it; x, y [B, C, H, W]."""
    m = ssim_map(x, y)
    if padding == "valid":
        m = m[:, :, RADIUS:-RADIUS, RADIUS:-RADIUS]
    elif padding != "same":
        raise ValueError(padding)
    return m.mean()


def l1_ssim(render: torch.Tensor, target: torch.Tensor, lambda_ssim: float = 0.2,
            padding: str = "same") -> torch.Tensor:
    """train.py:255-260 for one [C, H, W] render (or a [B, C, H, W] batch)."""
    x = render if render.dim() == 4 else render[None]
    y = target if target.dim() == 4 else target[None]
    l1 = (x - y).abs().mean()
    if lambda_ssim <= 0.0:
        return l1
    return (1.0 - lambda_ssim) * l1 + lambda_ssim * (1.0 - ssim(x, y, padding))


def l1_ssim_grad(render: torch.Tensor, target: torch.Tensor, lambda_ssim: float = 0.2, padding: str = "same"):
    """(loss, dloss/drender) in float64 by autograd."""
    x = render.detach().to(torch.float64).requires_grad_()
    loss = l1_ssim(x, target.detach().to(torch.float64), lambda_ssim, padding)
    loss.backward()
    return float(loss.detach()), x.grad

