"""CameraInfo with the reference's conventions (gray/camera.py:8-118).

R is the COLMAP camera-to-world rotation (transposed world-to-camera),
origin the camera centre, fov_y the vertical field of view in radians.  The
ray generator uses the Blender-axes camera-to-world matrix
c2w = [R[:,0], -R[:,1], -R[:,2]] (gray/camera.py:109-119); a pixel's
direction is normalize(x*c2w[:,0] + y*c2w[:,1] - c2w[:,2]) (camera.h:17-36).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, fields
from typing import Optional

import numpy as np


@dataclass
class CameraInfo:
    uid: int
    R: np.ndarray
    T: np.ndarray
    origin: np.ndarray
    fov_y: float
    fov_x: float
    image_path: str = ""
    image_name: str = ""
    image_width: int = 0
    image_height: int = 0
    is_test: bool = False

    @staticmethod
    def from_json(json_data: dict) -> "CameraInfo":
        kw = {}
        for f in fields(CameraInfo):
            v = json_data.get(f.name)
            kw[f.name] = np.array(v) if isinstance(v, list) else v
        return CameraInfo(**kw)

    def to_json(self) -> dict:
        out = {}
        for f in fields(self):
            v = getattr(self, f.name)
            out[f.name] = v.tolist() if isinstance(v, np.ndarray) else v
        return out

    def origin_cuda(self):
        import torch

        t = getattr(self, "_origin_cuda", None)
        if t is None:
            t = torch.from_numpy(np.asarray(self.origin, dtype=np.float32).copy()).cuda()
            self._origin_cuda = t
        return t

    def rotation_c2w_blender(self) -> np.ndarray:
        rot = -np.asarray(self.R, dtype=np.float32).copy()
        rot[:, 0] *= -1
        return rot

    def rotation_c2w_blender_cuda(self):
        import torch

        t = getattr(self, "_rotation_c2w_blender_cuda", None)
        if t is None:
            t = torch.from_numpy(self.rotation_c2w_blender()).cuda()
            self._rotation_c2w_blender_cuda = t
        return t


def look_at(origin, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov_y: float = math.radians(60.0),
            width: int = 64, height: int = 64, uid: int = 0, name: Optional[str] = None) -> CameraInfo:
    """Pinhole camera at `origin` looking at `target` with `up` roughly up.

    Builds the COLMAP-convention R (columns: camera x right, y down, z
    forward, in world coordinates) so that the Blender c2w used for rays has
    -c2w[:,2] = forward and c2w[:,1] = up (gray/camera.py:109-119)."""
    o = np.asarray(origin, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - o
    f /= np.linalg.norm(f)
    u = np.asarray(up, dtype=np.float64)
    r = np.cross(f, u)
    if np.linalg.norm(r) < 1e-9:
        u = np.array([0.0, 0.0, 1.0]) if abs(f[2]) < 0.9 else np.array([1.0, 0.0, 0.0])
        r = np.cross(f, u)
    r /= np.linalg.norm(r)
    down = np.cross(f, r)
    R = np.stack([r, down, f], axis=1)
    T = -R.T @ o
    fov_x = 2.0 * math.atan(math.tan(fov_y / 2.0) * width / height)
    return CameraInfo(uid=uid, R=R, T=T, origin=o, fov_y=float(fov_y), fov_x=float(fov_x),
                      image_name=name or f"view_{uid:05d}", image_width=width, image_height=height)
