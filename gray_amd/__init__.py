"""gray_amd: MI355X-native (gfx950) differentiable Gaussian ray tracer.

Drop-in for the ray-tracing path of graphdeco-inria/gray: `Raytracer` keeps
the reference Python API (gray/raytracer.py) and Gaussian tensor layout; the
work runs in hand-written HIP kernels behind a C ABI (include/gray_hip.h).
"""
from .camera import CameraInfo, look_at
from .config import RaytracerConfig

__all__ = ["CameraInfo", "RaytracerConfig", "Raytracer", "look_at"]


def __getattr__(name):
    if name == "Raytracer":
        from .raytracer import Raytracer

        return Raytracer
    raise AttributeError(name)
