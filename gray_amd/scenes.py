"""Seeded synthetic scenes and cameras for the benchmark workloads.

Parameter distributions (per Gaussian, reference tensor layout):
  mean        random (ball) or on random surfaces with N(0, 0.02^2) jitter
  scale       log-scales ~ N(mu_s, sigma_s^2)
  rotation    normalize(N(0,1)^4)
  opacity     logit stored ~ N(0,1)      (opacity = sigmoid(N(0,1)))
  sh_dc       ~ N(0.5, 0.2^2)
  sh_rest     ~ N(0, 0.05^2), SH degree 3
Workloads:
  tiny(1024)          1,024 Gaussians, means ~ N(0, 0.5^2), log-scale N(-3, 0.5^2);
                      one 64x64 camera (fov 60 deg) at distance 3 looking at the origin
  surfaces(N, S, ...) N Gaussians on S random spheres/planes inside radius 5
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .camera import CameraInfo, look_at


@dataclass
class GaussianScene:
    mean: torch.Tensor      # [N,3]
    rotation: torch.Tensor  # [N,4]
    scale: torch.Tensor     # [N,3] log
    opacity: torch.Tensor   # [N,1] logit
    sh_dc: torch.Tensor     # [N,1,3]
    sh_rest: torch.Tensor   # [N,15,3]

    @property
    def count(self) -> int:
        return int(self.mean.shape[0])

    def channels(self) -> torch.Tensor:
        """Degree-0 colour (used when SH is off)."""
        return (0.28209479177387814 * self.sh_dc[:, 0, :] + 0.5).clamp(min=0.0)

    def to(self, device) -> "GaussianScene":
        return GaussianScene(*(getattr(self, f).to(device) for f in
                               ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")))


def _appearance(n: int, gen: torch.Generator, log_scale_mu: float, log_scale_sigma: float):
    scale = torch.randn(n, 3, generator=gen) * log_scale_sigma + log_scale_mu
    q = torch.randn(n, 4, generator=gen)
    rotation = q / q.norm(dim=1, keepdim=True)
    opacity = torch.randn(n, 1, generator=gen)
    sh_dc = torch.randn(n, 1, 3, generator=gen) * 0.2 + 0.5
    sh_rest = torch.randn(n, 15, 3, generator=gen) * 0.05
    return rotation, scale, opacity, sh_dc, sh_rest


def tiny_scene(n: int = 1024, seed: int = 0) -> GaussianScene:
    gen = torch.Generator().manual_seed(seed)
    mean = torch.randn(n, 3, generator=gen) * 0.5
    rotation, scale, opacity, sh_dc, sh_rest = _appearance(n, gen, -3.0, 0.5)
    return GaussianScene(mean, rotation, scale, opacity, sh_dc, sh_rest)


def surface_points(n: int, surfaces: int, gen: torch.Generator) -> torch.Tensor:
    """n points on `surfaces` random surfaces inside radius 5: even index a
    sphere (centre in a ball of radius 3, radius U(0.5, 2)), odd index a square
    patch (centre in a ball of radius 2, random normal, side U(1, 4)); uniform
    on each surface, then N(0, 0.02^2) jitter."""
    per = [n // surfaces + (1 if i < n % surfaces else 0) for i in range(surfaces)]
    pts = []
    for i, m in enumerate(per):
        if i % 2 == 0:
            c = torch.randn(3, generator=gen)
            c = c / c.norm() * 3.0 * torch.rand(1, generator=gen) ** (1 / 3)
            rad = 0.5 + 1.5 * torch.rand(1, generator=gen)
            u = torch.randn(m, 3, generator=gen)
            p = c + rad * u / u.norm(dim=1, keepdim=True)
        else:
            c = torch.randn(3, generator=gen)
            c = c / c.norm() * 2.0 * torch.rand(1, generator=gen) ** (1 / 3)
            nrm = torch.randn(3, generator=gen)
            nrm = nrm / nrm.norm()
            a = torch.linalg.cross(nrm, torch.tensor([1.0, 0.0, 0.0]) if abs(float(nrm[0])) < 0.9
                                   else torch.tensor([0.0, 1.0, 0.0]))
            a = a / a.norm()
            b = torch.linalg.cross(nrm, a)
            side = 1.0 + 3.0 * torch.rand(1, generator=gen)
            uv = (torch.rand(m, 2, generator=gen) - 0.5) * side
            p = c + uv[:, :1] * a + uv[:, 1:] * b
        pts.append(p)
    pts = torch.cat(pts, 0)
    return pts + torch.randn(pts.shape, generator=gen) * 0.02


def surface_scene(n: int, surfaces: int = 20, log_scale_mu: float = -4.5, log_scale_sigma: float = 0.4,
                  seed: int = 0) -> GaussianScene:
    gen = torch.Generator().manual_seed(seed)
    mean = surface_points(n, surfaces, gen)
    rotation, scale, opacity, sh_dc, sh_rest = _appearance(n, gen, log_scale_mu, log_scale_sigma)
    return GaussianScene(mean, rotation, scale, opacity, sh_dc, sh_rest)


def ring_camera(radius: float, theta: float, width: int, height: int, fov_deg: float = 60.0,
                uid: int = 0) -> CameraInfo:
    origin = (radius * math.sin(theta), 0.0, radius * math.cos(theta))
    return look_at(origin, (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(fov_deg), width, height, uid=uid)


def ring_cameras(count: int, radius: float, width: int, height: int, fov_deg: float = 60.0):
    return [ring_camera(radius, 2 * math.pi * i / count, width, height, fov_deg, uid=i) for i in range(count)]


def sphere_camera(rng: np.random.Generator, r_min: float, r_max: float, width: int, height: int,
                  fov_deg: float = 60.0, uid: int = 0) -> CameraInfo:
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    o = d * rng.uniform(r_min, r_max)
    up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.95 else (1.0, 0.0, 0.0)
    return look_at(o, (0.0, 0.0, 0.0), up, math.radians(fov_deg), width, height, uid=uid)


def tiny_camera(width: int = 64, height: int = 64) -> CameraInfo:
    return look_at((0.0, 0.0, 3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(60.0), width, height)
