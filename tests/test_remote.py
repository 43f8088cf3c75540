"""Remote-viewer front end on the CPU: the wire protocol (packet framing of
viewer/__init__.py:76-165, the image codec of widgets/image.py:27-58), the
camera widget convention, the depth display colours, and a loopback
WebSocket session against a stand-in scene that serves the same submit/poll
interface as SceneServer (the GPU test drives the real one)."""
import asyncio
import math

import numpy as np
import pytest
import torch

from gray_amd import remote as R
from gray_amd import scenes
from gray_amd.serving import Request
from oracle import oracle as O


# ---- packets ---------------------------------------------------------------

def test_packet_round_trip():
    text, blob = R.pack_packet({"viewer": {"a": 1}, "2": {"codec": "raw"}},
                               [(2, b"abc"), ("viewer", np.arange(4, dtype=np.uint8))])
    assert R.needs_binary(text)
    assert blob == b"abc" + bytes([0, 1, 2, 3])
    pkt = R.unpack_packet(text, blob)
    assert pkt["viewer"]["metadata"] == {"a": 1}
    assert pkt[2]["metadata"] == {"codec": "raw"} and bytes(pkt[2]["binary"]) == b"abc"
    assert bytes(pkt["viewer"]["binary"]) == bytes([0, 1, 2, 3])


def test_packet_without_binaries_and_empty():
    text, blob = R.pack_packet({"0": {"res_x": 4}})
    assert blob == b"" and not R.needs_binary(text)
    assert R.unpack_packet(text) == {0: {"metadata": {"res_x": 4}}}
    assert R.pack_packet({}) == (None, b"")


def test_packet_size_mismatch_and_wrong_frame_type():
    text, blob = R.pack_packet({}, [(2, b"abcd")])
    with pytest.raises(ValueError, match="binary sizes"):
        R.unpack_packet(text, blob + b"x")
    with pytest.raises(ValueError, match="binary sizes"):
        R.unpack_packet(text, blob[:2])
    with pytest.raises(ValueError, match="as text"):
        R.unpack_packet(b"{}")


# ---- frames ----------------------------------------------------------------

def test_raw_codec_is_exact():
    img = np.random.default_rng(0).integers(0, 256, (5, 7, 3), dtype=np.uint8)
    b, meta = R.encode_image(img, "raw", 80)
    assert meta == {"codec": "raw", "shape": [5, 7, 3]}
    assert np.array_equal(R.decode_image(b, meta), img)
    b, meta = R.encode_image(img, "jpeg", 100)  # quality 100 sends raw
    assert meta["codec"] == "raw" and np.array_equal(R.decode_image(b, meta), img)


def test_jpeg_codec():
    y, x = np.mgrid[0:48, 0:64]
    img = np.stack([x * 4, y * 5, (x + y) * 2], -1).astype(np.uint8)
    b, meta = R.encode_image(img, "jpeg", 90)
    assert meta == {"codec": "jpeg"} and b[:2] == b"\xff\xd8"
    out = R.decode_image(b, meta)
    assert out.shape == img.shape and np.abs(out.astype(int) - img).mean() < 3.0
    with pytest.raises(ValueError, match="codec"):
        R.encode_image(img, "png", 80)


def test_rgb8_truncates_and_clamps():
    img = torch.tensor([[[0.999, 1.5, -1.0], [0.5, 1.0, 0.0]]])
    assert R.to_rgb8(img).tolist() == [[[254, 255, 0], [127, 255, 0]]]


def test_depth_colours_piecewise_linear():
    d = torch.tensor([0.0, 1.0, 3.0, 2.0, 1.7, 2.9, 5.0, -1.0]).reshape(1, -1, 1)
    rgb = R.depth_to_rgb(d, 1.0, 3.0)[0].numpy()
    cols = np.array(R.VIRIDIS_COLOURS)
    assert np.allclose(rgb[0], 0) and np.allclose(rgb[7], 0), "misses are black"
    assert np.allclose(rgb[1], cols[-1]) and np.allclose(rgb[2], cols[0]), "near is the last knot"
    assert np.allclose(rgb[3], (cols[2] + cols[3]) / 2, atol=1e-6)
    assert np.allclose(rgb[6], cols[0]), "beyond depth_max clamps"
    v = np.clip((3.0 - d[0, :, 0].numpy()) / 2.0, 0, 1)
    ref = np.stack([np.interp(v, R.VIRIDIS_KNOTS, cols[:, c]) for c in range(3)], -1)
    hit = d[0, :, 0].numpy() > 0
    assert np.allclose(rgb[hit], ref[hit], atol=1e-6)


# ---- cameras ---------------------------------------------------------------

def test_camera_widget_convention(oracle_lib):
    """to_world's columns are right, down and forward: the centre pixel looks
    along forward, the top-left pixel up and to the left."""
    cam = scenes.look_at((1.0, 2.0, 5.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), math.radians(50.0), 3, 3)
    state = R.widget_state(cam)
    back = R.camera_from_widget(state)
    assert np.allclose(back.R, cam.R) and np.allclose(back.origin, cam.origin)
    assert (back.image_width, back.image_height) == (3, 3)
    to_world = np.array(state["to_world"])
    right, down, fwd = to_world[:3, 0], to_world[:3, 1], to_world[:3, 2]
    ro, rd = O.pinhole_rays(3, 3, back.fov_y, back.origin, back.rotation_c2w_blender())
    assert np.allclose(rd[4], fwd, atol=1e-6)
    assert rd[0] @ right < 0 and rd[0] @ down < 0
    assert np.allclose(ro[0], cam.origin)


def test_server_resolution_recomputes_fov_x():
    cam = scenes.tiny_camera(64, 64)
    state = R.widget_state(cam)
    state["res_x"], state["res_y"] = 10, 20
    c = R.camera_from_widget(state, (160, 90))
    assert (c.image_width, c.image_height) == (160, 90)
    assert c.fov_y == pytest.approx(cam.fov_y)
    assert c.fov_x == pytest.approx(2 * math.atan(math.tan(cam.fov_y / 2) * 160 / 90))
    with pytest.raises(ValueError):
        R.camera_from_widget(dict(state, res_x=0))


# ---- loopback session ----------------------------------------------------------

class StandInScene:
    """submit/poll/busy/set_global_scale_factor of SceneServer; the image
    encodes the request (red: origin x / 10, green: scale factor / 4, blue:
    a ramp), depth 2 except a miss at pixel (0, 0)."""

    def __init__(self):
        self.queue, self.gsf, self.batches, self.rebuilds = [], 1.0, [], 0

    @property
    def busy(self):
        return bool(self.queue)

    def submit(self, r):
        self.queue.append(r)

    def poll(self):
        batch, self.queue = self.queue, []
        if batch:
            self.batches.append(len(batch))
        for r in batch:
            h, w = r.height, r.width
            img = torch.zeros(h, w, 3)
            img[..., 0] = float(r.camera.origin[0]) / 10
            img[..., 1] = self.gsf / 4
            img[..., 2] = torch.linspace(0, 1, w)
            r.image = img
            if r.mode == "Depth":
                r.depth = torch.full((h, w, 1), 2.0)
                r.depth[0, 0, 0] = 0.0
            r.done = 0.0
        return batch

    def set_global_scale_factor(self, v):
        self.gsf = v
        self.rebuilds += 1


VIEWER = dict(scaling_modifier=1.0, render_mode="Gaussians", znear=0.0, depth_min=0.0, depth_max=10.0,
              ellipsoid_min_opacity=0.05, jpeg_quality=80)


def test_loopback_session():
    aiohttp = pytest.importorskip("aiohttp")
    W, H = 24, 16
    cam0 = scenes.ring_camera(6.0, 0.3, W, H, 60.0)
    cam1 = scenes.ring_camera(6.0, 1.1, W, H, 60.0)
    scene = StandInScene()
    viewer = R.RemoteViewer(scene, cam0, codec="raw", train_cameras=[cam0, cam1])

    async def scenario():
        port = await viewer.start("127.0.0.1", 0)
        try:
            async with aiohttp.ClientSession() as s:
                c = await R.ViewerClient.connect(s, f"ws://127.0.0.1:{port}/")
                img, meta, pkt = await c.receive_frame()
                assert meta["frame_id"] == 1 and img.shape == (H, W, 3)
                vm = pkt["viewer"]["metadata"]
                assert vm["render_modes"] == ["Gaussians", "Depth", "Ellipsoids"] and vm["remote_codec"] == "raw"
                assert [c_["uid"] for c_ in vm["train_cameras"]] == [cam0.uid, cam1.uid] and vm["test_cameras"] == []
                assert pkt[1]["metadata"]["model_saved_views"] == []
                assert img[0, 0, 0] == int(float(cam0.origin[0]) / 10 * 255)
                assert img[0, -1, 2] == 255 and img[0, 0, 2] == 0

                # a new camera: the server keeps its resolution
                st = R.widget_state(cam1)
                st["res_x"], st["res_y"] = 7, 5
                await c.send_state(VIEWER, st)
                img, meta, pkt = await c.receive_frame()
                assert meta["frame_id"] == 2 and img.shape == (H, W, 3)
                assert "train_cameras" not in pkt["viewer"]["metadata"]
                assert img[3, 3, 0] == int(float(cam1.origin[0]) / 10 * 255)

                # depth display
                await c.send_state(dict(VIEWER, render_mode="Depth", depth_min=1.0, depth_max=3.0), st)
                img, meta, _ = await c.receive_frame()
                mid = (np.array(R.VIRIDIS_COLOURS[2]) + np.array(R.VIRIDIS_COLOURS[3])) / 2
                assert np.array_equal(img[0, 0], [0, 0, 0])
                assert np.abs(img[5, 5].astype(float) - mid * 255).max() <= 1.0

                # scaling modifier: one rebuild, then the new factor in the image
                await c.send_state(dict(VIEWER, scaling_modifier=2.0), st)
                img, meta, _ = await c.receive_frame()
                assert scene.rebuilds == 1 and img[2, 2, 1] == int(2.0 / 4 * 255)

                # a second client shares the loop
                c2 = await R.ViewerClient.connect(s, f"ws://127.0.0.1:{port}/")
                img2, meta2, _ = await c2.receive_frame()
                assert meta2["frame_id"] == 1 and img2.shape == (H, W, 3)
                assert len(viewer.sessions) == 2

                # an unoffered mode and a binary frame first end the connection
                await c2.send_state(dict(VIEWER, render_mode="Normals"))
                msg = await c2.ws.receive(timeout=10)
                assert msg.type in (aiohttp.WSMsgType.CLOSE, aiohttp.WSMsgType.CLOSED)
                await c.ws.send_bytes(b"\x00\x01")
                msg = await c.ws.receive(timeout=10)
                assert msg.type in (aiohttp.WSMsgType.CLOSE, aiohttp.WSMsgType.CLOSED)
        finally:
            await viewer.stop()
        assert viewer.frames_sent == 5

    asyncio.run(asyncio.wait_for(scenario(), 60))


def test_single_client_limit():
    aiohttp = pytest.importorskip("aiohttp")
    cam = scenes.tiny_camera(8, 8)
    viewer = R.RemoteViewer(StandInScene(), cam, codec="raw", max_clients=1)

    async def scenario():
        port = await viewer.start("127.0.0.1", 0)
        try:
            async with aiohttp.ClientSession() as s:
                c = await R.ViewerClient.connect(s, f"ws://127.0.0.1:{port}/")
                await c.receive_frame()
                c2 = await R.ViewerClient.connect(s, f"ws://127.0.0.1:{port}/")
                msg = await c2.ws.receive(timeout=10)
                assert msg.type in (aiohttp.WSMsgType.CLOSE, aiohttp.WSMsgType.CLOSED)
                assert len(viewer.sessions) == 1
        finally:
            await viewer.stop()

    asyncio.run(asyncio.wait_for(scenario(), 60))


def test_request_modes():
    cam = scenes.tiny_camera(8, 8)
    for m in ("Gaussians", "Depth", "Ellipsoids"):
        assert Request(0, cam, mode=m).mode == m
    with pytest.raises(ValueError, match="display mode"):
        Request(0, cam, mode="Normals")


class FailingScene(StandInScene):
    def poll(self):
        if self.queue:
            raise RuntimeError("device: error flag 0x2")
        return []


def test_scene_error_closes_connections():
    aiohttp = pytest.importorskip("aiohttp")
    viewer = R.RemoteViewer(FailingScene(), scenes.tiny_camera(8, 8), codec="raw")

    async def scenario():
        port = await viewer.start("127.0.0.1", 0)
        try:
            async with aiohttp.ClientSession() as s:
                c = await R.ViewerClient.connect(s, f"ws://127.0.0.1:{port}/")
                msg = await c.ws.receive(timeout=10)
                assert msg.type in (aiohttp.WSMsgType.CLOSE, aiohttp.WSMsgType.CLOSED)
        finally:
            with pytest.raises(RuntimeError, match="error flag"):
                await viewer.stop()
        assert "error flag" in str(viewer.error)

    asyncio.run(asyncio.wait_for(scenario(), 60))
