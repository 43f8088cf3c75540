"""Host-side logic of the drop-in API (CPU only)."""
import math

import numpy as np
import pytest
import torch

from gray_amd.camera import CameraInfo, look_at
from gray_amd.config import RaytracerConfig


def test_config_defaults_match_reference():
    # reference gray/config.py:29-151
    c = RaytracerConfig()
    assert (c.exp_power, c.alpha_threshold, c.t_threshold) == (2.0, 0.01, 0.03)
    assert (c.ppll_forward_size, c.ppll_backward_size) == (300_000_000, 120_000_000)
    assert (c.lr_mean_init, c.lr_scale_init, c.lr_sh_dc_init, c.lr_sh_rest) == (0.00016, 0.02, 0.04, 0.000625)
    assert (c.beta_1, c.beta_2, c.epsilon, c.sh_update_laziness) == (0.9, 0.999, 1e-15, 1)
    assert c.sh and c.sh_max_degree == 3
    assert RaytracerConfig(sh=False).sh_max_degree == 0
    assert c.iterations in c.save_iters
    hq = RaytracerConfig.from_json({"exp_power": 1, "alpha_threshold": 0.001, "t_threshold": 0.001})
    assert (hq.exp_power, hq.alpha_threshold) == (1, 0.001)


def test_camera_blender_convention():
    # gray/camera.py:109-119: c2w = -R with column 0 negated back
    cam = CameraInfo(uid=0, R=-np.eye(3), T=np.zeros(3), origin=np.array([0, 0, 1.0]), fov_y=1.5, fov_x=1.5)
    assert np.allclose(cam.rotation_c2w_blender(), np.diag([-1.0, 1.0, 1.0]))


def test_look_at_forward_and_up():
    cam = look_at((0.0, 0.0, 3.0), (0, 0, 0), (0, 1, 0), math.radians(60), 64, 64)
    c2w = cam.rotation_c2w_blender()
    assert np.allclose(-c2w[:, 2], [0, 0, -1], atol=1e-6)  # looks toward the origin
    assert np.allclose(c2w[:, 1], [0, 1, 0], atol=1e-6)   # up stays up
    assert np.allclose(np.cross(c2w[:, 0], c2w[:, 1]), c2w[:, 2], atol=1e-6)


def test_holder_resize_clears_new_rows():
    from gray_amd.native import GaussianDataHolder, hit_capacity_for

    g = GaussianDataHolder(3, False, torch.device("cpu"))
    g.lr_mean.fill_(0.5)
    # poison the allocator-visible storage, then grow
    g.resize(1)
    g.grad_mean.fill_(7.0)
    g.resize(5)
    assert g.grad_mean.shape == (5, 3) and torch.all(g.grad_mean[1:] == 0)
    assert g.mean.grad is g.grad_mean
    assert torch.all(g.lr_mean == 0.5) and g.lr_mean.shape == (5, 1)
    assert g.sh_coeffs_rest.shape == (5, 15, 3)
    g.mean.copy_(torch.arange(15.0).reshape(5, 3))
    g.resize(2)
    assert torch.equal(g.mean, torch.arange(6.0).reshape(2, 3))
    with pytest.raises(RuntimeError):
        g.resize(4)
    assert hit_capacity_for(980 * 545, 120_000_000) == 64
    assert hit_capacity_for(4096, 0) == 0


def test_capacities_hold_the_cap():
    from gray_amd.functional import capacities

    for views, w, h in ((8, 512, 512), (1, 980, 545), (3, 64, 64)):
        stored, lists = capacities(views, w, h, 4 << 30)
        tiles = -(-w // 8) * -(-h // 8)
        lanes = min(views, 8) * tiles * 64
        used = 8 * stored * views * w * h + 16 * lists * lanes + 256 * lanes
        assert used <= 4 << 30 and lists > 0


def test_expon_lr_schedule_closed_forms():
    import math

    from gray_amd.training import expon_lr

    f = expon_lr(1e-2, 1e-4, max_steps=1000)
    assert math.isclose(f(0), 1e-2, rel_tol=1e-12)
    assert math.isclose(f(1000), 1e-4, rel_tol=1e-12)
    assert math.isclose(f(5000), 1e-4, rel_tol=1e-12)  # constant after max_steps
    assert math.isclose(f(500), math.sqrt(1e-2 * 1e-4), rel_tol=1e-12)  # geometric midpoint
    assert all(f(s) > f(s + 1) for s in range(0, 1000, 37))
    assert f(-1) == 0.0 and expon_lr(0.0, 0.0)(10) == 0.0
    d = expon_lr(1e-2, 1e-4, delay_steps=100, delay_mult=0.01, max_steps=1000)
    assert math.isclose(d(0), 0.01 * 1e-2, rel_tol=1e-12)  # eased in from delay_mult * lr
    assert math.isclose(d(100), f(100), rel_tol=1e-12)  # ramp complete
