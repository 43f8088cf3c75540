"""A few iterations of the reference training loop through the drop-in API:
learning-rate schedules, SH degree increments, scale decay, pruning (with
compaction of parameters and Adam state) and BVH rebuilds; after pruning the
render still equals the oracle on the surviving Gaussians."""
import math

import numpy as np
import pytest
import torch

from gray_amd import Raytracer, RaytracerConfig, scenes
from gray_amd.training import TrainingSchedule, expon_lr
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_training_loop_schedules_prune_rebuild():
    scene = scenes.tiny_scene(1024, seed=13)
    cam = scenes.tiny_camera(48, 40)
    cfg = RaytracerConfig(iterations=8, pruning_interval=4, pruning_from_iter=4, pruning_min_weight=1e-5,
                          rebuild_interval=3, sh_increment_interval=2, sh_init_degree=0)
    rt = Raytracer(cfg, scene.count, 48, 40)
    g = rt.cuda_module.get_gaussians()
    with torch.no_grad():
        for name, src in (("mean", scene.mean), ("rotation", scene.rotation), ("scale", scene.scale),
                          ("opacity", scene.opacity), ("sh_coeffs_dc", scene.sh_dc),
                          ("sh_coeffs_rest", scene.sh_rest)):
            getattr(g, name).copy_(src)
        g.current_sh_degree.fill_(0)
    rt.cuda_module.rebuild_bvh()
    sched = TrainingSchedule(cfg, rt, scene_radius=1.5)
    target = torch.rand(3, 40, 48, generator=torch.Generator().manual_seed(4)).cuda()
    counts = []
    for it in range(1, cfg.iterations + 1):
        sched.before_step(it)
        assert math.isclose(float(g.lr_opacity[0]), expon_lr(cfg.lr_opacity_init, cfg.lr_opacity_final,
                            delay_mult=cfg.lr_schedule_delay_mult, max_steps=cfg.iterations)(it - 1), rel_tol=1e-6)
        rt.backward(torch.nn.functional.l1_loss(rt(cam), target))
        rt.step()
        sched.after_step(it)
        counts.append(g.count)
        torch.cuda.synchronize()
        rt.cuda_module.device_status()
    assert int(g.current_sh_degree) == 3  # incremented at 2, 4, 6 (capped at 3)
    assert sched.prunes == 2 and counts[-1] < scene.count  # pruned at 4 and 8
    assert sched.rebuilds >= 3
    # after the last step: a fresh render equals the oracle on the survivors
    with torch.no_grad():
        img = rt(cam).permute(1, 2, 0).reshape(-1, 3).cpu().numpy()
    torch.cuda.synchronize()
    nm = rt.cuda_module
    geo, _ = nm.debug_snapshot()
    fb = nm.get_framebuffer()
    ro, rd = O.pinhole_rays(48, 40, cam.fov_y, cam.origin.astype(np.float32), cam.rotation_c2w_blender())
    k = O.make_knobs(alpha_threshold=cfg.alpha_threshold, t_threshold=cfg.t_threshold, exp_power=cfg.exp_power)
    ref = O.render(geo.cpu().numpy(), g.channels.cpu().numpy(), ro, rd, k)
    assert np.abs(img - ref["C"]).max() < 1e-4
