"""Raytracer: drop-in for the reference's Python API (gray/raytracer.py:30-438).

Same constructor, `__call__`, `backward`, `step`, `set_render_resolution`,
`from_point_cloud`, `from_safetensors`, `save_safetensors`, `prune` and
`update_pruning_stats`; the same Gaussian tensor layout, reached through
`raytracer.cuda_module.get_gaussians()`.  The native side is the gfx950 HIP
library (gray_amd/_lib.py); there is no CPU fallback.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from .camera import CameraInfo
from .config import RaytracerConfig
from .native import NativeRaytracer

SH_C0 = 0.28209479177387814


@dataclass
class BasicPointCloud:
    """Reference gray/scene.py BasicPointCloud fields used by from_point_cloud."""

    points: np.ndarray
    colors: np.ndarray
    distances_to_cam: np.ndarray
    radius: float = 1.0
    normals: Optional[np.ndarray] = None


class Raytracer(torch.nn.Module):
    def __init__(self, cfg: RaytracerConfig, num_points: int, image_width: int, image_height: int,
                 inference_only: bool = False):
        "Call `from_safetensors` or `from_point_cloud` to initialise with actual data."
        super().__init__()
        self.cfg = cfg
        self.image_width = image_width
        self.image_height = image_height
        self.render_width = image_width
        self.render_height = image_height
        self.cuda_module = NativeRaytracer(image_width, image_height, num_points, cfg.sh_max_degree,
                                           cfg.ppll_forward_size, cfg.ppll_backward_size, inference_only)
        self.bg_color = torch.tensor(cfg.bg_color)

        config = self.cuda_module.get_config()
        config.alpha_threshold.fill_(cfg.alpha_threshold)
        config.t_threshold.fill_(cfg.t_threshold)
        config.exp_power.fill_(cfg.exp_power)
        config.render_depth.fill_(cfg.render_depth)
        config.background_channels.copy_(self.bg_color)
        config.enable_sh.fill_(cfg.sh)
        config.needs_ray_output.fill_(cfg.post_mlp)
        # only optimise channels when they are not a view-dependent output colour
        config.update_channels.fill_(not cfg.pre_mlp and not cfg.sh)

        gaussians = self.cuda_module.get_gaussians()
        gaussians.lr_rotation.fill_(cfg.lr_rotation_init)
        gaussians.lr_scale.fill_(cfg.lr_scale_init)
        gaussians.lr_mean.fill_(cfg.lr_mean_init)
        gaussians.lr_opacity.fill_(cfg.lr_opacity_init)
        gaussians.lr_channels.fill_(cfg.lr_channels)
        gaussians.lr_sh_dc.fill_(cfg.lr_sh_dc_init)
        gaussians.lr_sh_rest.fill_(cfg.lr_sh_rest)
        gaussians.beta_1.fill_(cfg.beta_1)
        gaussians.beta_2.fill_(cfg.beta_2)
        gaussians.epsilon.fill_(cfg.epsilon)
        gaussians.sh_update_laziness.fill_(cfg.sh_update_laziness)

        # last output, kept for the backward pass
        self.output_channels = None
        self._loss = None  # FramebufferLoss of backward_l1_ssim

    def __call__(self, cam_info: CameraInfo, znear=0.0, zfar=99999.9, skip_copy=False):
        "Render the scene from cam_info; returns [C, h, w]."
        framebuffer = self.cuda_module.get_framebuffer()
        grad_enabled = torch.is_grad_enabled()
        assert not (skip_copy and grad_enabled), "skip_copy=True is not supported with gradients enabled"
        # render_ellipsoids is a host-memory knob here: the check costs no synchronisation
        assert not (
            grad_enabled and self.cuda_module.get_config().render_ellipsoids.item()
        ), "render_ellipsoids=True is only supported for no-grad display renders"

        # the camera fills and set_pose (gray/raytracer.py:126-132) and the
        # forward's resets, in one launch
        self.cuda_module.prepare_frame(cam_info.origin, cam_info.rotation_c2w_blender(), float(cam_info.fov_y),
                                       znear, zfar)
        self.cuda_module.forward_pass()

        h, w = self.render_height, self.render_width
        output_channels = framebuffer.output_channels.detach()[:h, :w]
        if not skip_copy or grad_enabled:
            output_channels = output_channels.clone()
        output_channels = output_channels.moveaxis(-1, 0)
        if grad_enabled:
            assert self.output_channels is None, "Called the forward pass multiple times without a backward pass"
            output_channels.requires_grad_()
            self.output_channels = output_channels
        return output_channels

    def backward(self, loss):
        loss.backward()
        with torch.no_grad():
            framebuffer = self.cuda_module.get_framebuffer()
            h, w = self.render_height, self.render_width
            framebuffer.grad_output_channels[:h, :w].copy_(self.output_channels.grad.moveaxis(0, -1))
            self.output_channels = None
        self.cuda_module.backward_pass()

    def backward_l1_ssim(self, target: torch.Tensor, lambda_ssim: float = 0.2, scale: float = 1.0) -> torch.Tensor:
        """The reference training step's loss and backward (train.py:255-262)
        with the fused loss kernels: `scale` * gradient of (1 - lambda) *
        l1(render, target) + lambda * (1 - SSIM(render, target)) goes straight
        into the framebuffer's grad_output_channels (no autograd hop and no
        copy, gray/raytracer.py:175-179), then the ray tracer's backward runs.
        target [C, h, w]; returns the device tensor {loss, l1, ssim} (unscaled)."""
        assert self.output_channels is not None, "backward_l1_ssim needs a forward pass with gradients enabled"
        if self._loss is None:
            from .loss import FramebufferLoss

            self._loss = FramebufferLoss(self.cuda_module.device)
        fb = self.cuda_module.get_framebuffer()
        loss = self._loss(fb.output_channels, fb.grad_output_channels, self.render_height, self.render_width,
                          target, lambda_ssim, scale)
        self.output_channels = None
        self.cuda_module.backward_pass()
        return loss

    def step(self):
        # pruning counter update (update_pruning_stats) is fused into the Adam launch
        self.cuda_module.step(update_pruning_stats=True)
        self.cuda_module.update_bvh()

    def set_render_resolution(self, width: int, height: int):
        "Render at a reduced resolution (must not exceed the allocated framebuffer size)."
        self.cuda_module.set_render_resolution(width, height)
        self.render_width = width
        self.render_height = height

    @staticmethod
    def from_point_cloud(cfg: RaytracerConfig, point_cloud: BasicPointCloud, image_width: int, image_height: int,
                         inference_only: bool = False) -> "Raytracer":
        """Reference gray/raytracer.py:205-294 (optional voxel binning, then
        identity rotation, log(distance*init_scale), logit(init_opacity),
        sh_dc = (rgb-0.5)/C0)."""
        if cfg.init_binning:
            points = torch.from_numpy(point_cloud.points).cuda()
            colors = torch.from_numpy(point_cloud.colors).cuda()
            distances = torch.from_numpy(point_cloud.distances_to_cam).cuda().unsqueeze(1)
            rounded = (points / (cfg.init_bin_size * point_cloud.radius)).round()
            _, inverse, counts = rounded.unique(dim=0, sorted=True, return_counts=True, return_inverse=True)
            m = counts.shape[0]
            cs = torch.zeros((m, 3), device=colors.device, dtype=colors.dtype).index_add_(0, inverse, colors)
            ps = torch.zeros((m, 3), device=points.device, dtype=points.dtype).index_add_(0, inverse, points)
            ds = torch.zeros((m, 1), device=points.device, dtype=distances.dtype).index_add_(0, inverse, distances)
            cnt = counts.unsqueeze(1)
            point_cloud = BasicPointCloud(points=(ps / cnt).cpu().numpy(), colors=(cs / cnt).cpu().numpy(),
                                          distances_to_cam=(ds / cnt).squeeze(1).cpu().numpy(),
                                          radius=point_cloud.radius)
        n = point_cloud.points.shape[0]
        rt = Raytracer(cfg, n, image_width, image_height, inference_only=inference_only)
        g = rt.cuda_module.get_gaussians()
        with torch.no_grad():
            g.rotation.copy_(torch.tensor([[1.0, 0.0, 0.0, 0.0]]).repeat(n, 1))
            g.scale.copy_(torch.from_numpy(point_cloud.distances_to_cam * cfg.init_scale).unsqueeze(1).log())
            g.mean.copy_(torch.from_numpy(point_cloud.points))
            g.opacity.copy_(torch.logit(cfg.init_opacity * torch.ones(n, 1)))
            g.channels.copy_(torch.from_numpy(point_cloud.colors))
            g.sh_coeffs_dc.copy_((torch.from_numpy(point_cloud.colors).unsqueeze(1) - 0.5) / SH_C0)
        rt.cuda_module.rebuild_bvh()
        return rt

    @torch.no_grad()
    def prune(self, iteration: int, mask: Optional[torch.Tensor] = None):
        """Reference gray/raytracer.py:296-359: keep Gaussians whose average
        pruning weight reaches cfg.pruning_min_weight (or the given mask)."""
        g = self.cuda_module.get_gaussians()
        names = list(g.PARAMS)
        kept = {}
        for name in names:
            for prefix in ("", "first_moment_", "second_moment_"):
                kept[prefix + name] = getattr(g, prefix + name).clone()
        if mask is None:
            denom = g.pruning_counter.squeeze(1).clamp(min=1)
            average_weight = g.pruning_weight.squeeze(1) / denom
            mask = average_weight >= self.cfg.pruning_min_weight
        self.cuda_module.resize(int(mask.sum().item()))
        for key, val in kept.items():
            getattr(g, key).copy_(val[mask])
        g.pruning_weight.zero_()
        g.pruning_counter.zero_()

    @torch.no_grad()
    def update_pruning_stats(self):
        """Reference gray/raytracer.py:361-366 (Raytracer.step fuses this into
        the Adam launch; call it directly only when stepping by hand)."""
        g = self.cuda_module.get_gaussians()
        mask = g.was_visible.squeeze(1)
        g.pruning_counter[mask] += 1

    @staticmethod
    def from_safetensors(cfg: RaytracerConfig, path: str, image_width: int, image_height: int,
                         inference_only: bool = False) -> "Raytracer":
        """Reference gray/raytracer.py:368-420 (same keys and layouts)."""
        import safetensors.torch

        sd = safetensors.torch.load_file(path)
        for k in ("bg_color", "image_width", "image_height"):
            sd.pop(k, None)
        n = sd["mean"].shape[0]
        rt = Raytracer(cfg, n, image_width, image_height, inference_only=inference_only)
        g = rt.cuda_module.get_gaussians()
        with torch.no_grad():
            g.mean.copy_(sd["mean"])
            g.rotation.copy_(sd["rotation"])
            g.scale.copy_(sd["scale"])
            g.opacity.copy_(sd["opacity"])
            if cfg.sh:
                g.channels.zero_()
                g.sh_coeffs_dc.copy_(sd["sh_coeffs_dc"])
                g.sh_coeffs_rest.copy_(sd["sh_coeffs_rest"])
                g.current_sh_degree.copy_(sd["current_sh_degree"])
            else:
                g.channels.copy_(sd["channels"])
                g.sh_coeffs_dc.zero_()
                g.sh_coeffs_rest.zero_()
                g.current_sh_degree.zero_()
        rt.cuda_module.rebuild_bvh()
        for k in ("mean", "rotation", "scale", "opacity", "channels", "sh_coeffs_dc", "sh_coeffs_rest",
                  "current_sh_degree"):
            sd.pop(k, None)
        rt.load_state_dict(sd)
        return rt

    def save_safetensors(self, model_path: str, iteration: int) -> str:
        """Reference gray/raytracer.py:422-438 (keys: mean, rotation, scale,
        opacity and sh_coeffs_dc/sh_coeffs_rest/current_sh_degree or channels)."""
        import safetensors.torch

        g = self.cuda_module.get_gaussians()
        tensors = {"mean": g.mean, "rotation": g.rotation, "scale": g.scale, "opacity": g.opacity}
        if self.cfg.sh:
            tensors["sh_coeffs_dc"] = g.sh_coeffs_dc
            tensors["sh_coeffs_rest"] = g.sh_coeffs_rest
            tensors["current_sh_degree"] = g.current_sh_degree
        else:
            tensors["channels"] = g.channels
        tensors = {k: v.detach().contiguous() for k, v in tensors.items()}
        path = os.path.join(model_path, f"gaussians_{iteration:05d}.safetensors")
        safetensors.torch.save_file({**self.state_dict(), **tensors}, path)
        return path
