"""gray_frame_setup: the one-launch holder updates before a frame (camera
fills and set_pose of gray/raytracer.py:126-132, the metadata update and
was_visible reset of raytracer.h:88-93) write exactly what the torch
operations they replace write, and the reset covers any base alignment and
length."""
import ctypes as C
import math

import numpy as np
import pytest
import torch

from gray_amd import _lib as L
from gray_amd import scenes

from gpu_util import setup_native

pytestmark = pytest.mark.gpu


def test_prepare_frame_writes_the_holders():
    scene = scenes.tiny_scene(1000, seed=3)
    cam = scenes.tiny_camera(24, 16)
    nr = setup_native(scene, cam)
    g, c, cfg, meta = nr.get_gaussians(), nr.get_camera(), nr.get_config(), nr.get_metadata()
    g.was_visible.fill_(True)
    cfg.rays_from_python.fill_(True)
    calls = int(meta.total_num_calls.item())
    other = scenes.look_at((1.0, -2.0, 4.0), (0.1, 0.2, 0.0), (0.0, 1.0, 0.0), math.radians(40.0), 24, 16)
    rot = other.rotation_c2w_blender()
    with torch.no_grad():
        nr.prepare_frame(other.origin, rot, other.fov_y, 0.25, 50.0)
    torch.cuda.synchronize()
    assert np.array_equal(c.origin.cpu().numpy(), np.asarray(other.origin, np.float32))
    assert np.array_equal(c.rotation_c2w.cpu().numpy(), rot)
    assert np.array_equal(c.rotation_w2c.cpu().numpy(), rot.T)
    assert c.vertical_fov_radians.item() == np.float32(other.fov_y)
    assert c.znear.item() == np.float32(0.25) and c.zfar.item() == np.float32(50.0)
    assert not cfg.rays_from_python.item()
    assert not meta.grads_enabled.item()
    assert int(meta.total_num_calls.item()) == calls + 1
    assert not g.was_visible.any().item()


@pytest.mark.parametrize("offset,count", [(0, 1), (0, 16), (0, 1000), (3, 1037), (5, 15), (1, 4096)])
def test_was_visible_reset_any_alignment(offset, count):
    buf = torch.ones(offset + count + 7, dtype=torch.uint8, device="cuda")
    a = L.FrameArgs()
    a.was_visible = buf.data_ptr() + offset
    a.count = count
    L.check(L.LIB.gray_frame_setup(C.byref(a), L.stream_ptr()))
    torch.cuda.synchronize()
    out = buf.cpu().numpy()
    assert (out[:offset] == 1).all() and (out[offset:offset + count] == 0).all() and (out[offset + count:] == 1).all()


def test_frame_setup_rejects_negative_count():
    a = L.FrameArgs()
    a.count = -1
    assert L.LIB.gray_frame_setup(C.byref(a), None) == L.GRAY_ERR_INVALID
