"""ctypes binding of libgray_hip.so (the C ABI declared in include/gray_hip.h).

The product path has no fallback: if the HIP library is missing this module
raises on import, naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GRAY_HIP_LIB", os.path.join(HERE, "libgray_hip.so"))

GRAY_OK = 0
GRAY_ERR_INVALID = 1
GRAY_ERR_HIP = 2
GRAY_ERR_OVERFLOW = 3
GRAY_ERR_NOMEM = 4
GRAY_ERR_STATE = 5

MAX_ALPHA = 0.9999  # reference CMakeLists.txt:68-71
NUM_CHANNELS = 3
BATCH = 32

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64


class FrameArgs(C.Structure):
    _fields_ = [
        ("origin", vp), ("rotation_c2w", vp), ("rotation_w2c", vp), ("vertical_fov", vp), ("znear", vp),
        ("zfar", vp), ("rays_from_python", vp), ("grads_enabled", vp), ("total_num_calls", vp),
        ("was_visible", vp), ("count", i64), ("origin_value", C.c_float * 3), ("c2w_value", C.c_float * 9),
        ("fov_value", C.c_float), ("znear_value", C.c_float), ("zfar_value", C.c_float),
        ("grads_enabled_value", i32),
    ]


class Knobs(C.Structure):
    _fields_ = [
        ("exp_power", vp), ("alpha_threshold", vp), ("t_threshold", vp), ("global_scale_factor", vp),
        ("eps_forward_normalization", vp), ("eps_scale_grad", vp), ("background_channels", vp),
        ("render_depth", vp), ("enable_sh", vp), ("jitter_primary_rays", vp), ("call_index", vp),
        ("needs_ray_output", vp),
    ]


class Params(C.Structure):
    _fields_ = [
        ("count", i32), ("num_sh_coeffs", i32), ("sh_is_fp16", i32), ("current_sh_degree", vp),
        ("mean", vp), ("rotation", vp), ("scale", vp), ("opacity", vp), ("sh_dc", vp), ("sh_rest", vp),
    ]


class View(C.Structure):
    _fields_ = [
        ("origin", vp), ("rotation_c2w", vp), ("vertical_fov", vp), ("znear", vp), ("zfar", vp),
        ("rays_origin", vp), ("rays_direction", vp), ("rays_enable", vp),
        ("width", i32), ("height", i32), ("stride", i32), ("alloc_height", i32),
        ("rgb", vp),
        ("out_channels", vp), ("out_depth", vp), ("out_transmittance", vp), ("out_full_transmittance", vp),
        ("out_remaining_est", vp), ("out_ray_origin", vp), ("out_ray_direction", vp),
        ("stat_hits", vp), ("stat_accumulated", vp),
        ("hit_ids", vp), ("hit_trans", vp), ("n_comp", vp), ("comp_rgb", vp),
        ("grad_out", vp), ("grad_rgb", vp),
    ]


class RenderArgs(C.Structure):
    _fields_ = [
        ("knobs", Knobs), ("num_views", i32), ("views", C.POINTER(View)), ("save", i32),
        ("hit_capacity", i32), ("use_bvh", i32), ("tile_stats", vp), ("list_capacity", i32),
        ("collect_budget", i32), ("fp16_hits", i32), ("wave_list_min", i32), ("deterministic", i32),
        ("sh_params", vp), ("collect_mode", i32),
    ]


class Grads(C.Structure):
    _fields_ = [
        ("grad_mean", vp), ("grad_rotation", vp), ("grad_scale", vp), ("grad_opacity", vp),
        ("pruning_weight", vp), ("was_visible", vp),
    ]


class Adam(C.Structure):
    _fields_ = [
        ("param", vp * 7), ("grad", vp * 7), ("m1", vp * 7), ("m2", vp * 7), ("lr", vp * 7),
        ("numel", i64 * 7), ("beta1", vp), ("beta2", vp), ("epsilon", vp),
        ("was_visible", vp), ("pruning_counter", vp), ("count", i32),
        ("update_channels", vp), ("zero_grads", vp), ("enable_sh", vp), ("sh_update_laziness", vp),
        ("step_device", vp),
    ]


class LossArgs(C.Structure):
    _fields_ = [
        ("planes", i32), ("height", i32), ("width", i32), ("valid_padding", i32), ("train", i32), ("pad", i32),
        ("render", vp), ("render_stride", i64 * 3), ("target", vp), ("target_stride", i64 * 3),
        ("grad", vp), ("grad_stride", i64 * 3), ("lambda_ssim", C.c_float), ("grad_scale", C.c_float),
        ("grad_scale_device", vp), ("workspace", vp), ("loss", vp),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libgray_hip.so not found at {LIB_PATH}; build it with `python -m gray_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback."
        )
    lib = C.CDLL(LIB_PATH)
    sig = {
        "gray_ctx_create": (i32, [C.POINTER(vp)]),
        "gray_ctx_destroy": (None, [vp]),
        "gray_ctx_share_scene": (i32, [vp, vp]),
        "gray_last_error": (C.c_char_p, []),
        "gray_abi_version": (i32, []),
        "gray_device_status": (i32, [vp, vp]),
        "gray_frame_setup": (i32, [C.POINTER(FrameArgs), vp]),
        "gray_bvh_update": (i32, [vp, C.POINTER(Params), C.POINTER(Knobs), i32, vp]),
        "gray_bvh_set_timing": (i32, [vp, i32]),
        "gray_bvh_last_build_ms": (C.c_float, [vp]),
        "gray_bvh_info": (i32, [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]),
        "gray_sh_forward": (i32, [vp, C.POINTER(Params), C.POINTER(Knobs), i32, C.POINTER(vp), C.POINTER(vp), vp]),
        "gray_sh_backward": (i32, [vp, C.POINTER(Params), C.POINTER(Knobs), i32, C.POINTER(vp), C.POINTER(vp),
                                   C.POINTER(vp), vp, vp, vp, vp, i32, vp]),
        "gray_render_forward": (i32, [vp, C.POINTER(RenderArgs), vp]),
        "gray_render_backward": (i32, [vp, C.POINTER(RenderArgs), C.POINTER(Params), C.POINTER(Grads), vp]),
        "gray_adam_step": (i32, [vp, C.POINTER(Adam), i32, i32, vp]),
        "gray_debug_hitlists": (i32, [vp, C.POINTER(RenderArgs), i32, vp, vp, vp, vp]),
        "gray_debug_snapshot": (i32, [vp, vp, vp, vp]),
        "gray_debug_bvh": (i32, [vp, vp, vp, vp, vp]),
        "gray_debug_wide": (i32, [vp, vp, vp, vp]),
        "gray_debug_list_counts": (i32, [vp, vp, i64, C.POINTER(i64), vp]),
        "gray_render_stage_ms": (i32, [vp, C.POINTER(C.c_float), i32]),
        "gray_render_reserve": (i32, [vp, i64, i32]),
        "gray_render_ellipsoids": (i32, [vp, C.POINTER(RenderArgs), vp, vp]),
        "gray_loss_workspace_bytes": (i32, [i32, i32, i32, C.POINTER(i64)]),
        "gray_loss_forward": (i32, [C.POINTER(LossArgs), vp]),
        "gray_loss_backward": (i32, [C.POINTER(LossArgs), vp]),
        "gray_loss_noalloc": (i32, [C.POINTER(LossArgs), vp]),
        "gray_debug_stream_copy": (i32, [vp, vp, i64, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


ABI_VERSION = 14  # GRAY_ABI_VERSION of include/gray_hip.h that these structures mirror
LIB = _load()
if LIB.gray_abi_version() != ABI_VERSION:
    raise ImportError(f"libgray_hip.so has ABI {LIB.gray_abi_version()}, the binding expects {ABI_VERSION}; "
                      "rebuild with python -m gray_amd.build")
EXPORTED = (
    "gray_ctx_create", "gray_ctx_destroy", "gray_last_error", "gray_abi_version", "gray_device_status",
    "gray_bvh_update", "gray_bvh_set_timing", "gray_bvh_last_build_ms", "gray_bvh_info", "gray_sh_forward",
    "gray_sh_backward", "gray_render_forward", "gray_render_backward", "gray_adam_step", "gray_debug_hitlists",
    "gray_debug_snapshot", "gray_debug_bvh", "gray_debug_list_counts", "gray_render_stage_ms",
    "gray_render_reserve", "gray_loss_workspace_bytes", "gray_loss_forward", "gray_loss_backward",
    "gray_render_ellipsoids", "gray_debug_wide", "gray_frame_setup", "gray_loss_noalloc", "gray_ctx_share_scene",
    "gray_debug_stream_copy",
)
LOSS_NOALLOC_MAX_BYTES = 80 << 20  # GRAY_LOSS_NOALLOC_MAX_BYTES


class GrayError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def check(rc: int) -> None:
    if rc != GRAY_OK:
        raise GrayError(rc, LIB.gray_last_error().decode(errors="replace"))


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None)."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class Context:
    """Owns a gray_ctx (the LBVH and its scratch) on the current device."""

    def __init__(self):
        h = vp()
        check(LIB.gray_ctx_create(C.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                LIB.gray_ctx_destroy(h)
            except Exception:
                pass
            self.handle = None

    def status(self, stream=None) -> None:
        check(LIB.gray_device_status(self.handle, stream_ptr(stream)))
