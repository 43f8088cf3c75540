"""The loss oracle (oracle/loss_ref.py) pinned on the CPU, and the host-side
checks of the fused loss (no device work).

fused_ssim has no test or fixture in the reference; its SSIM is pinned
against scipy.ndimage.gaussian_filter (separable Gaussian of radius 5,
sigma 1.5, zero padding: an independent route through the windowing and the
border), identities, and fp64 central differences of the gradient."""
import ctypes as C

import numpy as np
import pytest
import scipy.ndimage as ndi
import torch

from oracle import loss_ref as R


def scipy_ssim_map(x: np.ndarray, y: np.ndarray, sigma: float = 1.5) -> np.ndarray:
    """[C,H,W] float64 -> SSIM map with scipy's Gaussian filter (mode constant 0)."""
    f = lambda a: np.stack([ndi.gaussian_filter(c, sigma, mode="constant", cval=0.0,  # noqa: E731
                                                truncate=R.RADIUS / sigma) for c in a])
    mu1, mu2 = f(x), f(y)
    s11, s22, s12 = f(x * x) - mu1 ** 2, f(y * y) - mu2 ** 2, f(x * y) - mu1 * mu2
    return ((2 * mu1 * mu2 + R.C1) * (2 * s12 + R.C2)) / ((mu1 ** 2 + mu2 ** 2 + R.C1) * (s11 + s22 + R.C2))


def images(c=3, h=21, w=17, seed=0):
    g = torch.Generator().manual_seed(seed)
    x = torch.rand(c, h, w, generator=g, dtype=torch.float64)
    y = (x + 0.2 * torch.randn(c, h, w, generator=g, dtype=torch.float64)).clamp(0, 1)
    return x, y


def test_window_matches_scipy():
    delta = np.zeros(11)
    delta[5] = 1.0
    ref = ndi.gaussian_filter1d(delta, 1.5, mode="constant", truncate=5 / 1.5)
    assert np.abs(R.window().numpy() - ref).max() < 1e-15
    assert abs(float(R.window().sum()) - 1.0) < 1e-15


@pytest.mark.parametrize("shape", [(3, 21, 17), (1, 4, 30), (2, 40, 9)])
def test_ssim_map_matches_scipy(shape):
    x, y = images(*shape, seed=sum(shape))
    m = R.ssim_map(x[None], y[None])[0].numpy()
    ref = scipy_ssim_map(x.numpy(), y.numpy())
    assert np.abs(m - ref).max() < 1e-12
    # the check sees a wrong window width and a wrong border rule
    assert np.abs(scipy_ssim_map(x.numpy(), y.numpy(), sigma=1.0) - m).max() > 1e-3
    mirror = np.stack([ndi.gaussian_filter(c, 1.5, mode="reflect", truncate=5 / 1.5) for c in x.numpy()])
    zero = np.stack([ndi.gaussian_filter(c, 1.5, mode="constant", truncate=5 / 1.5) for c in x.numpy()])
    assert np.abs(mirror - zero).max() > 1e-2


def test_ssim_identities():
    x, y = images()
    assert abs(float(R.ssim(x[None], x[None])) - 1.0) < 1e-14
    assert abs(float(R.ssim(x[None], y[None]) - R.ssim(y[None], x[None]))) < 1e-14
    assert float(R.ssim(x[None], y[None])) < 0.9
    full = R.ssim_map(x[None], y[None])
    assert abs(float(R.ssim(x[None], y[None], "valid")) - float(full[:, :, 5:-5, 5:-5].mean())) < 1e-15
    # l1 term alone is torch's l1_loss; lambda 1 is 1 - SSIM
    assert abs(float(R.l1_ssim(x, y, 0.0)) - float(torch.nn.functional.l1_loss(x, y))) < 1e-15
    assert abs(float(R.l1_ssim(x, y, 1.0)) - (1.0 - float(R.ssim(x[None], y[None])))) < 1e-15


@pytest.mark.parametrize("padding", ["same", "valid"])
def test_gradient_matches_finite_differences(padding):
    x, y = images(3, 14, 13, seed=4)
    lam = 0.2
    _, g = R.l1_ssim_grad(x, y, lam, padding)
    rng = np.random.default_rng(1)
    h = 1e-6
    for _ in range(24):
        c, i, j = int(rng.integers(3)), int(rng.integers(14)), int(rng.integers(13))
        if abs(float(x[c, i, j] - y[c, i, j])) < 10 * h:
            continue
        xp, xm = x.clone(), x.clone()
        xp[c, i, j] += h
        xm[c, i, j] -= h
        fd = (float(R.l1_ssim(xp, y, lam, padding)) - float(R.l1_ssim(xm, y, lam, padding))) / (2 * h)
        assert abs(fd - float(g[c, i, j])) < 1e-7 + 1e-5 * abs(fd), (c, i, j)


def test_host_checks_without_device():
    from gray_amd import _lib as L
    from gray_amd import loss

    n = C.c_int64()
    assert L.LIB.gray_loss_workspace_bytes(0, 4, 4, C.byref(n)) == L.GRAY_ERR_INVALID
    assert L.LIB.gray_loss_workspace_bytes(3, 545, 980, C.byref(n)) == L.GRAY_OK
    tiles = 62 * 35
    assert n.value == (3 * 3 * 545 * 980 * 4 + 255) // 256 * 256 + 2 * 4 * tiles  # one group of 3 planes
    assert L.LIB.gray_loss_forward(None, None) == L.GRAY_ERR_INVALID
    a = L.LossArgs()
    a.planes, a.height, a.width, a.valid_padding = 3, 8, 8, 1
    a.render = a.target = a.workspace = a.loss = 256  # never dereferenced: the shape check fails first
    assert L.LIB.gray_loss_forward(C.byref(a), None) == L.GRAY_ERR_INVALID
    assert b"valid padding" in L.LIB.gray_last_error()
    with pytest.raises(TypeError):
        loss.l1_ssim_loss(torch.zeros(3, 4, 4), torch.zeros(3, 4, 4))
