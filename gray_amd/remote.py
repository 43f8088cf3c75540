"""Remote-viewer front end: WebSocket clients as request sources for the
serving loop.

The reference's viewer can run headless as a server that renders for a remote
client over WebSockets (viewer/__init__.py:76-238, view.py:97-190, 340-421).
This module speaks the same wire protocol, so that viewer's client connects
unchanged.  It feeds every connected client's camera into SceneServer's
continuous batching, so several clients share the GPU's launches.  The
reference serves one client at a time (viewer/__init__.py:203-207);
`max_clients=1` restores that.

Packets (viewer/__init__.py:76-165).  A packet is one text frame, JSON
{"metadata": {id: dict}, "binaries": [{"widget_id": id, "length": n}, ...]}.
When "binaries" is present and not empty, one binary frame follows, holding
the binaries back to back in that order.  Ids are "viewer" or the viewer's
widget ids, which are numbered in the order the viewer makes its widgets
(view.py:97-127): 0 the camera, 1 the saved views, 2 the image.

Client to server:
  * "viewer": {scaling_modifier, render_mode, znear, depth_min, depth_max,
    ellipsoid_min_opacity, jpeg_quality} (view.py:340-349);
  * 0: the camera {res_x, res_y, fov_x, fov_y, z_near, z_far, to_world 4x4}
    (widgets/cameras/__init__.py:43-70; OpenCV axes: to_world's columns are
    right, down and forward, its last column the centre);
  * 1: saved-view actions.  They are ignored: this server keeps no saved
    views.

Server to client, with every frame:
  * "viewer": {render_modes, remote_codec, jpeg_quality}, plus train_cameras
    and test_cameras once per connection (view.py:356-366);
  * 1: {saved_view_status}, plus empty model_saved_views and
    source_saved_views once (widgets/saved_views.py:371-378);
  * 2: the frame.  Its metadata is {codec "raw" with shape [h, w, 3] or codec
    "jpeg", frame_id}, its binary the RGB8 image (widgets/image.py:27-46,
    233-240).

As in view.py:107-116, the server keeps its own resolution by default and
recomputes fov_x from the client's fov_y.  `client_resolution=True` renders
at the client's resolution instead.  The scaling modifier is a property of
the scene: a change, from any client, rebuilds the BVH of every slot.
"""
from __future__ import annotations

import argparse
import asyncio
import io
import json
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .camera import CameraInfo
from .serving import MODES, Request

CAMERA_WIDGET, SAVED_VIEWS_WIDGET, IMAGE_WIDGET = 0, 1, 2
CODECS = ("jpeg", "raw")

# Piecewise-linear viridis of the depth display (view.py:424-459): six knots
# on [0, 1] and their colours.
VIRIDIS_KNOTS = (0.0, 0.2, 0.4, 0.6, 0.8, 1.0)
VIRIDIS_COLOURS = (
    (0.267004, 0.004874, 0.329415),
    (0.253935, 0.265254, 0.529983),
    (0.163625, 0.471133, 0.558148),
    (0.134692, 0.658636, 0.517649),
    (0.477504, 0.821444, 0.318195),
    (0.993248, 0.906157, 0.143936),
)


# ---- packets ---------------------------------------------------------------

def pack_packet(metadata: Dict, binaries: Sequence[Tuple[object, bytes]] = ()) -> Tuple[Optional[str], bytes]:
    """(text frame, binary frame) of one packet; text is None for an empty
    packet, which is not sent (viewer/__init__.py:106-108).  The binary frame
    is b"" when there are no binaries."""
    binaries = [(wid, memoryview(b).cast("B")) for wid, b in binaries]
    if not metadata and not binaries:
        return None, b""
    payload: Dict = {"metadata": metadata}
    if binaries:
        payload["binaries"] = [{"widget_id": wid, "length": b.nbytes} for wid, b in binaries]
    return json.dumps(payload, separators=(",", ":")), b"".join(bytes(b) for _, b in binaries)


def _widget_key(widget_id):
    return "viewer" if widget_id == "viewer" else int(widget_id)


def unpack_packet(text, blob: bytes = b"") -> Dict[object, Dict]:
    """{id: {"metadata": dict or None, "binary": memoryview or None}} of one
    packet (viewer/__init__.py:122-165).  Raises ValueError for a text frame
    that is not text, and for binaries whose lengths do not add up to the
    binary frame."""
    if not isinstance(text, str):
        raise ValueError(f"Expected websocket control packet as text, received {type(text).__name__}")
    payload = json.loads(text)
    out: Dict[object, Dict] = {}
    for wid, meta in payload.get("metadata", {}).items():
        out.setdefault(_widget_key(wid), {})["metadata"] = meta
    view = memoryview(blob if blob is not None else b"")
    offset = 0
    for desc in payload.get("binaries", []):
        n = int(desc["length"])
        if offset + n > len(view):
            raise ValueError("Corrupted websocket payload: binary sizes do not match frame data")
        out.setdefault(_widget_key(desc["widget_id"]), {})["binary"] = view[offset:offset + n]
        offset += n
    if offset != len(view):
        raise ValueError("Corrupted websocket payload: binary sizes do not match frame data")
    return out


def needs_binary(text: str) -> bool:
    """Whether a binary frame follows this text frame."""
    return bool(json.loads(text).get("binaries"))


# ---- frames ----------------------------------------------------------------

def encode_image(img: np.ndarray, codec: str, jpeg_quality: int) -> Tuple[bytes, Dict]:
    """RGB8 [h, w, 3] image -> (bytes, metadata) (widgets/image.py:27-46):
    raw bytes with their shape, or a baseline 4:2:0 JPEG; quality 100 and
    above sends raw."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    if img.ndim != 3 or img.shape[2] != 3:
        raise ValueError(f"expected an [h, w, 3] image, got shape {img.shape}")
    if codec == "raw" or (codec == "jpeg" and jpeg_quality >= 100):
        return img.tobytes(), {"codec": "raw", "shape": list(img.shape)}
    if codec == "jpeg":
        from PIL import Image as PILImage

        buf = io.BytesIO()
        PILImage.fromarray(img, "RGB").save(buf, format="JPEG", quality=int(jpeg_quality), optimize=False,
                                            progressive=False, subsampling=2)
        return buf.getvalue(), {"codec": "jpeg"}
    raise ValueError(f"Unsupported remote image codec: {codec}")


def decode_image(binary, meta: Dict) -> np.ndarray:
    """Inverse of encode_image (widgets/image.py:49-58)."""
    codec = meta.get("codec", "raw")
    if codec == "raw":
        return np.frombuffer(bytes(binary), dtype=np.uint8).reshape(meta["shape"])
    if codec == "jpeg":
        from PIL import Image as PILImage

        with PILImage.open(io.BytesIO(bytes(binary))) as im:
            return np.ascontiguousarray(im.convert("RGB"))
    raise ValueError(f"Unsupported remote image codec: {codec}")


def to_rgb8(img: torch.Tensor) -> torch.Tensor:
    """[h, w, 3] float image -> uint8, clamped to [0, 1] and scaled by 255
    with truncation, as the viewer's image widget does (view.py:183-186,
    widgets/image.py:319-321)."""
    return (img.clamp(0.0, 1.0) * 255.0).to(torch.uint8)


def depth_to_rgb(depth: torch.Tensor, depth_min: float, depth_max: float) -> torch.Tensor:
    """[h, w, 1] depth -> [h, w, 3] display colours (view.py:424-459):
    depth_min maps to the last viridis knot and depth_max to the first,
    linearly in between; pixels with depth <= 0 (misses) are black."""
    depth_max = max(depth_max, depth_min + 1e-6)
    v = torch.clamp((depth_max - depth) / (depth_max - depth_min), 0.0, 1.0)[..., 0]
    knots = depth.new_tensor(VIRIDIS_KNOTS)
    colours = depth.new_tensor(VIRIDIS_COLOURS)
    flat = v.reshape(-1)
    seg = torch.clamp(torch.bucketize(flat, knots, right=True) - 1, 0, len(VIRIDIS_KNOTS) - 2)
    t = ((flat - knots[seg]) / (knots[seg + 1] - knots[seg])).unsqueeze(-1)
    rgb = (colours[seg] + t * (colours[seg + 1] - colours[seg])).view(v.shape[0], v.shape[1], 3)
    return torch.where((depth > 0.0).expand_as(rgb), rgb, torch.zeros_like(rgb))


# ---- cameras ---------------------------------------------------------------

def fov_x_of(fov_y: float, width: int, height: int) -> float:
    """The camera widget's compute_fov_x: the horizontal field of view of a
    pinhole with vertical field fov_y at width / height."""
    return 2.0 * math.atan(math.tan(fov_y / 2.0) * (width / height))


def camera_from_widget(state: Dict, resolution: Optional[Tuple[int, int]] = None) -> CameraInfo:
    """CameraInfo of the camera widget's state (view.py:153-166): R is
    to_world's rotation (OpenCV axes, the COLMAP convention of CameraInfo),
    the origin its last column.  With `resolution` the image is that size and
    fov_x follows from fov_y (view.py:107-116)."""
    to_world = np.asarray(state["to_world"], dtype=np.float64).reshape(4, 4)
    w, h = (int(state["res_x"]), int(state["res_y"])) if resolution is None else resolution
    fov_y = float(state["fov_y"])
    fov_x = float(state["fov_x"]) if resolution is None else fov_x_of(fov_y, w, h)
    if w <= 0 or h <= 0 or not (0.0 < fov_y < math.pi):
        raise ValueError(f"invalid camera: {w}x{h}, fov_y {fov_y}")
    return CameraInfo(uid=0, R=to_world[:3, :3].copy(), T=None, origin=to_world[:3, 3].copy(), fov_y=fov_y,
                      fov_x=fov_x, image_width=w, image_height=h)


def widget_state(camera: CameraInfo, z_near: float = 0.001, z_far: float = 100.0) -> Dict:
    """The camera widget's to_json() for a CameraInfo (the inverse of
    camera_from_widget)."""
    to_world = np.identity(4)
    to_world[:3, :3] = np.asarray(camera.R, dtype=np.float64)
    to_world[:3, 3] = np.asarray(camera.origin, dtype=np.float64)
    return {"res_x": int(camera.image_width), "res_y": int(camera.image_height), "fov_x": float(camera.fov_x),
            "fov_y": float(camera.fov_y), "z_near": z_near, "z_far": z_far, "to_world": to_world.tolist()}


# ---- one connection ----------------------------------------------------------

class ViewerSession:
    """Server-side state of one client (view.py:97-150, 351-421)."""

    def __init__(self, viewer: "RemoteViewer", sid: int):
        self.viewer = viewer
        self.sid = sid
        self.camera = viewer.initial_camera
        self.scaling_modifier = 1.0
        self.render_mode = "Gaussians"
        self.znear = 0.0
        self.depth_min, self.depth_max = 0.0, 10.0
        self.ellipsoid_min_opacity = 0.05
        self.jpeg_quality = viewer.jpeg_quality
        self.frame_id = 0
        self.dirty = True            # state changed since the last frame was requested
        self.in_flight: Optional[Request] = None
        self.cameras_sent = False
        self.saved_views_sent = False
        self._pending: Optional[str] = None  # a text frame waiting for its binary frame
        self.closed = False

    # -- from the client --
    def receive(self, frame) -> None:
        """One WebSocket frame.  A packet with binaries completes with the
        binary frame that follows it (viewer/__init__.py:122-136)."""
        if self._pending is not None:
            if isinstance(frame, str):
                raise ValueError("Expected websocket frame payload as binary, received text")
            text, self._pending = self._pending, None
            self.apply(unpack_packet(text, frame))
            return
        if not isinstance(frame, str):
            raise ValueError(f"Expected websocket control packet as text, received {type(frame).__name__}")
        if needs_binary(frame):
            self._pending = frame
        else:
            self.apply(unpack_packet(frame))

    def apply(self, packet: Dict) -> None:
        """view.py:406-421 (viewer state) and the camera widget's server_recv
        (widgets/cameras/__init__.py:43-50)."""
        before = self._key()
        meta = packet.get("viewer", {}).get("metadata")
        if meta:
            self.scaling_modifier = float(meta.get("scaling_modifier", self.scaling_modifier))
            mode = meta.get("render_mode", self.render_mode)
            if mode not in self.viewer.render_modes:
                raise ValueError(f"render mode {mode!r} is not offered (offered: {self.viewer.render_modes})")
            self.render_mode = mode
            self.znear = float(meta.get("znear", self.znear))
            self.depth_min = float(meta.get("depth_min", self.depth_min))
            self.depth_max = float(meta.get("depth_max", self.depth_max))
            self.ellipsoid_min_opacity = float(meta.get("ellipsoid_min_opacity", self.ellipsoid_min_opacity))
            if "jpeg_quality" in meta:
                self.jpeg_quality = max(1, min(100, int(meta["jpeg_quality"])))
            if self.scaling_modifier != self.viewer.scaling_modifier:
                self.viewer.scaling_request = self.scaling_modifier
        cam = packet.get(CAMERA_WIDGET, {}).get("metadata")
        if cam:
            res = None if self.viewer.client_resolution else self.viewer.resolution
            self.camera = camera_from_widget(cam, res)
        if self._key() != before:
            self.dirty = True

    def _key(self):
        c = self.camera
        return (np.asarray(c.R).tobytes(), np.asarray(c.origin).tobytes(), c.fov_y, c.image_width,
                c.image_height, self.render_mode, self.znear, self.depth_min, self.depth_max,
                self.ellipsoid_min_opacity, self.jpeg_quality)

    # -- to the serving loop --
    def request(self) -> Request:
        self.dirty = False
        return Request(self.sid, self.camera, mode=self.render_mode, znear=self.znear,
                       min_opacity=self.ellipsoid_min_opacity, tag=self)

    def frame_rgb8(self, req: Request) -> np.ndarray:
        """The finished request's display image as RGB8 on the host."""
        if req.mode == "Depth":
            img = depth_to_rgb(req.depth, self.depth_min, self.depth_max)
        else:
            img = req.image
        return to_rgb8(img).cpu().numpy()

    # -- to the client --
    def state_metadata(self) -> Dict:
        """The viewer's and the saved-views widget's text of a packet
        (view.py:356-366, widgets/saved_views.py:371-378)."""
        v = self.viewer
        meta = {"viewer": {"render_modes": list(v.render_modes), "remote_codec": v.codec,
                           "jpeg_quality": int(self.jpeg_quality)}}
        if not self.cameras_sent:
            meta["viewer"]["train_cameras"] = [c.to_json() for c in v.train_cameras]
            meta["viewer"]["test_cameras"] = [c.to_json() for c in v.test_cameras]
            self.cameras_sent = True
        sv = {"saved_view_status": ""}
        if not self.saved_views_sent:
            sv["model_saved_views"], sv["source_saved_views"] = [], []
            self.saved_views_sent = True
        meta[str(SAVED_VIEWS_WIDGET)] = sv
        return meta

    def frame_packet(self, rgb8: np.ndarray) -> Tuple[str, bytes]:
        binary, fmeta = encode_image(rgb8, self.viewer.codec, self.jpeg_quality)
        self.frame_id += 1
        fmeta["frame_id"] = self.frame_id
        meta = self.state_metadata()
        meta[str(IMAGE_WIDGET)] = fmeta
        return pack_packet(meta, [(IMAGE_WIDGET, binary)])


# ---- the server --------------------------------------------------------------

class RemoteViewer:
    """WebSocket front end of a scene server.

    `scene` is a SceneServer, or anything with its submit(), poll(), busy and
    set_global_scale_factor().  `initial_camera` fixes the server's
    resolution (unless client_resolution) and the first frame's view."""

    def __init__(self, scene, initial_camera: CameraInfo, *, render_modes: Sequence[str] = MODES,
                 codec: str = "jpeg", jpeg_quality: int = 80, train_cameras: Sequence[CameraInfo] = (),
                 test_cameras: Sequence[CameraInfo] = (), client_resolution: bool = False,
                 continuous: bool = False, max_clients: Optional[int] = None):
        if codec not in CODECS:
            raise ValueError(f"Unsupported remote image codec: {codec}")
        bad = [m for m in render_modes if m not in MODES]
        if bad or not render_modes:
            raise ValueError(f"render modes must be among {MODES}, got {list(render_modes)}")
        self.scene = scene
        self.initial_camera = initial_camera
        self.resolution = (int(initial_camera.image_width), int(initial_camera.image_height))
        self.render_modes = list(render_modes)
        self.codec = codec
        self.jpeg_quality = max(1, min(100, int(jpeg_quality)))
        self.train_cameras = list(train_cameras)
        self.test_cameras = list(test_cameras)
        self.client_resolution = client_resolution
        self.continuous = continuous
        self.max_clients = max_clients
        self.scaling_modifier = 1.0
        self.scaling_request: Optional[float] = None
        self.sessions: Dict[int, Tuple[ViewerSession, object]] = {}
        self.frames_sent = 0
        self.error: Optional[BaseException] = None  # what stopped the render pump
        self._next_sid = 0
        self._pump_task: Optional[asyncio.Task] = None
        self._runner = None
        self._wake: Optional[asyncio.Event] = None

    # -- connections --
    async def handle(self, request):
        from aiohttp import WSMsgType, web

        ws = web.WebSocketResponse(max_msg_size=0, compress=False)
        await ws.prepare(request)
        if self.max_clients is not None and len(self.sessions) >= self.max_clients:
            await ws.close(message=b"Client already connected. Only one client is allowed."
                           if self.max_clients == 1 else b"Too many clients.")
            return ws
        sid = self._next_sid
        self._next_sid += 1
        sess = ViewerSession(self, sid)
        self.sessions[sid] = (sess, ws)
        self._poke()
        try:
            async for msg in ws:
                if msg.type == WSMsgType.TEXT:
                    sess.receive(msg.data)
                elif msg.type == WSMsgType.BINARY:
                    sess.receive(bytes(msg.data))
                else:
                    break
                self._poke()
        except ValueError as e:
            await ws.close(message=str(e).encode()[:120])
        finally:
            sess.closed = True
            self.sessions.pop(sid, None)
        return ws

    def _poke(self) -> None:
        if self._wake is not None:
            self._wake.set()

    # -- the render pump --
    async def pump(self) -> None:
        """Submits a request for every session whose state changed (every
        session in continuous mode) with no frame in flight, polls the
        scene, and sends finished frames.  An error of the scene server
        (a device error, say) closes every connection and ends the pump with
        that error."""
        try:
            await self._pump()
        except asyncio.CancelledError:
            raise
        except Exception as e:
            self.error = e
            print(f"ERROR: remote viewer render loop stopped: {e}", flush=True)
            for _, ws in list(self.sessions.values()):
                await ws.close(message=b"server error")
            raise

    async def _pump(self) -> None:
        loop = asyncio.get_running_loop()
        self._wake = asyncio.Event()
        while True:
            if self.scaling_request is not None:
                value, self.scaling_request = self.scaling_request, None
                self.scene.set_global_scale_factor(value)
                self.scaling_modifier = value
                for sess, _ in self.sessions.values():
                    sess.dirty = True
            for sess, _ in list(self.sessions.values()):
                if sess.in_flight is None and (sess.dirty or self.continuous):
                    sess.in_flight = sess.request()
                    self.scene.submit(sess.in_flight)
            finished = self.scene.poll()
            for req in finished:
                sess = req.tag
                sess.in_flight = None
                if sess.closed or sess.sid not in self.sessions:
                    continue
                ws = self.sessions[sess.sid][1]
                if req.error is not None:
                    text, _ = pack_packet({"viewer": {"error": req.error}})
                    try:
                        await ws.send_str(text)
                    except (ConnectionError, RuntimeError):
                        sess.closed = True  # only this client's socket is affected
                    continue
                rgb8 = sess.frame_rgb8(req)
                text, blob = await loop.run_in_executor(None, sess.frame_packet, rgb8)
                try:
                    await ws.send_str(text)
                    await ws.send_bytes(blob)
                    self.frames_sent += 1
                except (ConnectionError, RuntimeError):
                    sess.closed = True
            if self.scene.busy:
                await asyncio.sleep(0.0002)
            else:
                self._wake.clear()
                try:
                    await asyncio.wait_for(self._wake.wait(), timeout=0.05)
                except asyncio.TimeoutError:
                    pass

    def app(self):
        from aiohttp import web

        app = web.Application()
        app.router.add_get("/", self.handle)
        return app

    async def start(self, host: str = "127.0.0.1", port: int = 6009) -> int:
        """Starts listening and the pump; returns the bound port (port 0 picks
        a free one)."""
        from aiohttp import web

        self._runner = web.AppRunner(self.app())
        await self._runner.setup()
        site = web.TCPSite(self._runner, host, port)
        await site.start()
        self._pump_task = asyncio.get_running_loop().create_task(self.pump())
        return site._server.sockets[0].getsockname()[1]

    async def stop(self) -> None:
        """Stops the pump and the server, then re-raises the error that
        stopped the pump, if one did."""
        err = None
        if self._pump_task is not None:
            self._pump_task.cancel()
            try:
                await self._pump_task
            except asyncio.CancelledError:
                pass
            except Exception as e:  # the pump's own error, already reported
                err = e
            self._pump_task = None
        if self._runner is not None:
            await self._runner.cleanup()
            self._runner = None
        if err is not None:
            raise err

    def run(self, host: str = "0.0.0.0", port: int = 6009) -> None:
        """Serves until interrupted."""
        async def main():
            bound = await self.start(host, port)
            print(f"INFO: serving the remote viewer on ws://{host}:{bound}", flush=True)
            await asyncio.Event().wait()

        asyncio.run(main())


# ---- a client for tests and scripts ------------------------------------------

class ViewerClient:
    """Minimal client of the protocol (the reference viewer's client side,
    viewer/__init__.py:238-300): sends viewer and camera state, receives
    packets."""

    def __init__(self, ws):
        self.ws = ws

    @staticmethod
    async def connect(session, url: str) -> "ViewerClient":
        return ViewerClient(await session.ws_connect(url, max_msg_size=0, compress=0))

    async def send_state(self, viewer: Optional[Dict] = None, camera: Optional[Dict] = None) -> None:
        meta = {}
        if viewer is not None:
            meta["viewer"] = viewer
        if camera is not None:
            meta[str(CAMERA_WIDGET)] = camera
        text, _ = pack_packet(meta)
        if text is not None:
            await self.ws.send_str(text)

    async def receive(self, timeout: float = 30.0) -> Dict:
        msg = await self.ws.receive(timeout=timeout)
        if not isinstance(msg.data, str):
            raise ValueError(f"expected a text frame, got {msg.type}")
        text, blob = msg.data, b""
        if needs_binary(text):
            bmsg = await self.ws.receive(timeout=timeout)
            blob = bytes(bmsg.data)
        return unpack_packet(text, blob)

    async def receive_frame(self, timeout: float = 30.0) -> Tuple[np.ndarray, Dict, Dict]:
        """(image, frame metadata, whole packet) of the next packet that
        carries a frame."""
        while True:
            pkt = await self.receive(timeout)
            img = pkt.get(IMAGE_WIDGET)
            if img and img.get("binary") is not None:
                return decode_image(img["binary"], img["metadata"]), img["metadata"], pkt

    async def close(self) -> None:
        await self.ws.close()


def _load_scene(args):
    from .functional import RenderSettings
    from .serving import SceneServer

    if args.scene:
        import safetensors.torch

        sd = safetensors.torch.load_file(args.scene)
        t = [sd[k] for k in ("mean", "rotation", "scale", "opacity", "sh_coeffs_dc", "sh_coeffs_rest")]
        degree = int(sd["current_sh_degree"].reshape(-1)[0]) if "current_sh_degree" in sd else 3
    else:
        from . import scenes

        sc = scenes.surface_scene(args.synthetic, 20, -4.5, 0.4, seed=0)
        t = [getattr(sc, k) for k in ("mean", "rotation", "scale", "opacity", "sh_dc", "sh_rest")]
        degree = 3
    settings = RenderSettings(sh_degree=degree)
    return SceneServer(*t, settings=settings, slots=args.slots, budget_bytes=int(args.budget_gb * (1 << 30)))


def main(argv: Optional[List[str]] = None) -> None:
    from . import scenes

    ap = argparse.ArgumentParser(description="Serve a resident scene to remote viewers over WebSockets.")
    src = ap.add_mutually_exclusive_group()
    src.add_argument("--scene", help="gaussians_*.safetensors file (the reference's checkpoint format)")
    src.add_argument("--synthetic", type=int, default=1_000_000, help="Gaussians of a synthetic surface scene")
    ap.add_argument("--ip", default="0.0.0.0")
    ap.add_argument("--port", type=int, default=6009)
    ap.add_argument("--codec", choices=CODECS, default="jpeg")
    ap.add_argument("--jpeg-quality", type=int, default=80)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--fov-y", type=float, default=60.0, help="degrees")
    ap.add_argument("--client-resolution", action="store_true", help="render at each client's resolution")
    ap.add_argument("--max-clients", type=int, default=None)
    ap.add_argument("--slots", type=int, default=3)
    ap.add_argument("--budget-gb", type=float, default=32.0)
    a = ap.parse_args(argv)
    scene = _load_scene(a)
    cam = scenes.ring_camera(6.0, 0.0, a.width, a.height, a.fov_y)
    RemoteViewer(scene, cam, codec=a.codec, jpeg_quality=a.jpeg_quality, client_resolution=a.client_resolution,
                 max_clients=a.max_clients).run(a.ip, a.port)


if __name__ == "__main__":
    main()
