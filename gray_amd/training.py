"""Training-loop neighbours of the ray tracer: learning-rate schedules, SH
degree increments, scale decay, pruning and periodic BVH rebuilds, applied
around `Raytracer.step()` in the order of the reference training loop
(train.py:218-287):

    sched = TrainingSchedule(cfg, raytracer, scene_radius)
    for it in range(1, cfg.iterations + 1):
        sched.before_step(it)          # SH degree, learning rates
        loss = ...(raytracer(camera), target); raytracer.backward(loss)
        raytracer.step()               # fused Adam + pruning counters + BVH refit
        sched.after_step(it)           # scale decay, pruning, rebuilds
"""
from __future__ import annotations

import math
from typing import Callable

import torch

from .config import RaytracerConfig


def expon_lr(lr_init: float, lr_final: float, delay_steps: int = 0, delay_mult: float = 1.0,
             max_steps: int = 1_000_000) -> Callable[[int], float]:
    """Log-linear learning-rate decay: lr_init at step 0, lr_final at
    max_steps (constant afterwards), optionally eased in from
    delay_mult * lr over the first delay_steps by a quarter sine.  This is the
    continuous exponential schedule of JaxNeRF/Plenoxels that the reference
    uses (gray/utils.py:4-61); a step < 0, or both rates 0, gives 0."""

    def rate(step: int) -> float:
        if step < 0 or (lr_init == 0.0 and lr_final == 0.0):
            return 0.0
        ramp = 1.0
        if delay_steps > 0:
            u = min(max(step / delay_steps, 0.0), 1.0)
            ramp = delay_mult + (1.0 - delay_mult) * math.sin(0.5 * math.pi * u)
        t = min(max(step / max_steps, 0.0), 1.0)
        return ramp * math.exp(math.log(lr_init) * (1.0 - t) + math.log(lr_final) * t)

    return rate


class TrainingSchedule:
    """Per-iteration bookkeeping of the reference training loop for one
    Raytracer (gray_amd.raytracer.Raytracer)."""

    def __init__(self, cfg: RaytracerConfig, raytracer, scene_radius: float = 1.0):
        self.cfg = cfg
        self.rt = raytracer
        mk = lambda a, b: expon_lr(a, b, delay_mult=cfg.lr_schedule_delay_mult, max_steps=cfg.iterations)  # noqa: E731
        # (holder learning-rate tensor, schedule) pairs; a schedule whose end
        # points are equal is left alone, as in the reference
        self.schedules = []
        for name, init, final in (("lr_mean", cfg.lr_mean_init * scene_radius, cfg.lr_mean_final * scene_radius),
                                  ("lr_rotation", cfg.lr_rotation_init, cfg.lr_rotation_final),
                                  ("lr_scale", cfg.lr_scale_init, cfg.lr_scale_final),
                                  ("lr_opacity", cfg.lr_opacity_init, cfg.lr_opacity_final),
                                  ("lr_sh_dc", cfg.lr_sh_dc_init, cfg.lr_sh_dc_final)):
            if final != init:
                self.schedules.append((name, mk(init, final)))
        self.rebuilds = 0
        self.prunes = 0

    def before_step(self, iteration: int) -> None:
        """SH degree increment and learning rates for `iteration` (1-based)."""
        cfg, g = self.cfg, self.rt.cuda_module.get_gaussians()
        if cfg.sh and iteration % cfg.sh_increment_interval == 0:
            g.increment_sh_degree()
        with torch.no_grad():
            for name, sched in self.schedules:
                getattr(g, name).fill_(sched(iteration - 1))

    def after_step(self, iteration: int) -> bool:
        """Scale decay, pruning and periodic rebuilds after the optimiser step;
        returns whether the BVH was rebuilt."""
        cfg, nm = self.cfg, self.rt.cuda_module
        g = nm.get_gaussians()
        if cfg.scale_decay < 1.0:
            with torch.no_grad():
                g.scale.add_(math.log(cfg.scale_decay))
        rebuild = False
        if cfg.pruning and iteration >= cfg.pruning_from_iter and iteration % cfg.pruning_interval == 0:
            self.rt.prune(iteration)
            self.prunes += 1
            rebuild = True
        if cfg.rebuild_interval > 0 and iteration % cfg.rebuild_interval == 0:
            rebuild = True
        if rebuild:
            nm.rebuild_bvh()
            self.rebuilds += 1
        return rebuild
