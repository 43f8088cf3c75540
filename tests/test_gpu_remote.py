"""Remote-viewer front end on the GPU: frames received over a WebSocket
(raw codec) equal, byte for byte, the serving loop's render of the same
camera and display mode, also with two clients sharing batches."""
import asyncio

import numpy as np
import pytest
import torch

from gray_amd import remote as R
from gray_amd import scenes
from gray_amd.functional import RenderSettings
from gray_amd.serving import Request, SceneServer, lanes_of

pytestmark = pytest.mark.gpu

W, H = 40, 32
VIEWER = dict(scaling_modifier=1.0, render_mode="Gaussians", znear=0.0, depth_min=1.0, depth_max=5.0,
              ellipsoid_min_opacity=0.3, jpeg_quality=80)


def test_remote_frames_equal_served_renders():
    aiohttp = pytest.importorskip("aiohttp")
    scene = scenes.tiny_scene(1024, seed=8)
    t = [getattr(scene, k).cuda() for k in ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")]
    budget = 3 * lanes_of(W, H) * (64 * 16 + 256) * 2
    server = SceneServer(*t, settings=RenderSettings(t_threshold=0.03), slots=2, budget_bytes=budget,
                         list_capacity=64)
    cam0 = scenes.tiny_camera(W, H)
    cam1 = R.camera_from_widget(R.widget_state(scenes.sphere_camera(np.random.default_rng(4), 2.5, 4.0, W, H)),
                                (W, H))

    def rgb8(req):
        img = server.render_alone(req)
        if req.mode == "Depth":
            img = R.depth_to_rgb(req.depth, VIEWER["depth_min"], VIEWER["depth_max"])
        return R.to_rgb8(img).cpu().numpy()

    ref0 = rgb8(Request(0, cam0))
    ref1 = rgb8(Request(0, cam1))
    ref_depth = rgb8(Request(0, cam1, mode="Depth"))
    ref_ell = rgb8(Request(0, cam1, mode="Ellipsoids", min_opacity=0.3))
    assert ref0.any() and ref_depth.any() and ref_ell.any()
    assert not np.array_equal(ref1, ref_ell)
    viewer = R.RemoteViewer(server, cam0, codec="raw")

    async def scenario():
        port = await viewer.start("127.0.0.1", 0)
        try:
            async with aiohttp.ClientSession() as s:
                url = f"ws://127.0.0.1:{port}/"
                a, b = await R.ViewerClient.connect(s, url), await R.ViewerClient.connect(s, url)
                for c in (a, b):
                    img, meta, _ = await c.receive_frame()
                    assert meta["frame_id"] == 1 and np.array_equal(img, ref0)
                st = R.widget_state(cam1)
                await a.send_state(VIEWER, st)
                await b.send_state(dict(VIEWER, render_mode="Depth"), st)
                img_a, _, _ = await a.receive_frame()
                img_b, _, _ = await b.receive_frame()
                assert np.array_equal(img_a, ref1)
                assert np.array_equal(img_b, ref_depth)
                await a.send_state(dict(VIEWER, render_mode="Ellipsoids"), st)
                img_a, meta, _ = await a.receive_frame()
                assert meta["frame_id"] == 3 and np.array_equal(img_a, ref_ell)
                await a.close()
                await b.close()
        finally:
            await viewer.stop()

    print(f"remote viewer: {W}x{H}, two clients", flush=True)
    asyncio.run(asyncio.wait_for(scenario(), 120))
    torch.cuda.synchronize()
