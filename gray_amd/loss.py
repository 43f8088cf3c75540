"""Fused L1 + SSIM training loss on the device.

The reference trains on (train.py:255-260)

    loss = (1 - lambda_ssim) * F.l1_loss(render, target) + lambda_ssim * (1 - fused_ssim(render[None], target[None]))

with the third-party fused_ssim extension (git a7c48d6), then copies the
autograd gradient of the [C,H,W] render into the [H,W,C] framebuffer
(gray/raytracer.py:175-179).  Here the loss and its gradient are the two
kernels of gray_loss_forward / gray_loss_backward (include/gray_hip.h), which
read any strided layout and write the gradient straight into the
framebuffer:

* `l1_ssim_loss(render, target, lambda_ssim)` and `fused_ssim(img1, img2,
  padding, train)` are differentiable w.r.t. the first image (autograd);
* `FramebufferLoss` is the path of `Raytracer.backward_l1_ssim`: no autograd
  hop, no [C,H,W] copy of the render or of its gradient.
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple

import torch

from . import _lib as L

F32 = torch.float32


def _planes(t: torch.Tensor) -> Tuple[int, int, int, Tuple[int, int, int]]:
    """(planes, height, width, (plane, row, column) strides) of a [C,H,W] or
    [B,C,H,W] fp32 CUDA tensor whose planes are uniformly strided."""
    if t.dtype != F32 or not t.is_cuda:
        raise TypeError(f"expected a float32 CUDA tensor, got {t.dtype} on {t.device}")
    if t.dim() == 3:
        return t.shape[0], t.shape[1], t.shape[2], (t.stride(0), t.stride(1), t.stride(2))
    if t.dim() == 4:
        if t.shape[0] > 1 and t.stride(0) != t.shape[1] * t.stride(1):
            raise ValueError("[B,C,H,W] input whose batch stride is not C times its channel stride")
        return t.shape[0] * t.shape[1], t.shape[2], t.shape[3], (t.stride(1), t.stride(2), t.stride(3))
    raise ValueError(f"expected [C,H,W] or [B,C,H,W], got shape {tuple(t.shape)}")


def workspace_bytes(planes: int, height: int, width: int) -> int:
    n = C.c_int64()
    L.check(L.LIB.gray_loss_workspace_bytes(planes, height, width, C.byref(n)))
    return n.value


def _args(planes, height, width, render, rs, target, ts, lambda_ssim, valid, train, workspace, loss,
          grad=None, gs=(0, 0, 0), grad_scale=1.0, grad_scale_device=None) -> L.LossArgs:
    a = L.LossArgs()
    a.planes, a.height, a.width = planes, height, width
    a.valid_padding, a.train = int(valid), int(train)
    a.render, a.target = render.data_ptr(), target.data_ptr()
    a.render_stride[:] = rs
    a.target_stride[:] = ts
    a.grad = grad.data_ptr() if grad is not None else None
    a.grad_stride[:] = gs
    a.lambda_ssim, a.grad_scale = float(lambda_ssim), float(grad_scale)
    a.grad_scale_device = grad_scale_device.data_ptr() if grad_scale_device is not None else None
    a.workspace, a.loss = workspace.data_ptr(), loss.data_ptr()
    return a


def _check_pair(x: torch.Tensor, y: torch.Tensor, padding: str):
    if padding not in ("same", "valid"):
        raise ValueError(f"padding must be 'same' or 'valid', got {padding!r}")
    if x.shape != y.shape:
        raise ValueError(f"shape mismatch {tuple(x.shape)} vs {tuple(y.shape)}")
    if x.device != y.device:
        raise ValueError("inputs on different devices")
    px, py = _planes(x), _planes(y)
    return px, py


class _LossFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, lambda_ssim: float, padding: str, ssim_only: bool, train: bool):
        (P, H, W, xs), (_, _, _, ys) = _check_pair(x, y, padding)
        ws = torch.empty(workspace_bytes(P, H, W), dtype=torch.uint8, device=x.device)
        loss = torch.empty(3, dtype=F32, device=x.device)
        a = _args(P, H, W, x, xs, y, ys, lambda_ssim, padding == "valid", train, ws, loss)
        L.check(L.LIB.gray_loss_forward(C.byref(a), L.stream_ptr()))
        ctx.ws, ctx.meta = ws, (P, H, W, xs, ys, lambda_ssim, padding, ssim_only, train)
        ctx.save_for_backward(x, y)
        return (loss[2] if ssim_only else loss[0]).clone()

    @staticmethod
    def backward(ctx, g):
        x, y = ctx.saved_tensors
        P, H, W, xs, ys, lambda_ssim, padding, ssim_only, train = ctx.meta
        if not train:
            raise RuntimeError("backward through a loss evaluated with train=False")
        grad = torch.empty(x.shape, dtype=F32, device=x.device)
        gs = _planes(grad)[3]
        g = g.detach().to(F32).contiguous()
        loss = torch.empty(3, dtype=F32, device=x.device)
        # ssim_only: the kernels differentiate 1 - SSIM (lambda 1), so the sign flips
        a = _args(P, H, W, x, xs, y, ys, lambda_ssim, padding == "valid", True, ctx.ws, loss, grad, gs,
                  -1.0 if ssim_only else 1.0, g)
        L.check(L.LIB.gray_loss_backward(C.byref(a), L.stream_ptr()))
        ctx.ws = None
        return grad, None, None, None, None, None


def l1_ssim_loss(render: torch.Tensor, target: torch.Tensor, lambda_ssim: float = 0.2,
                 padding: str = "same") -> torch.Tensor:
    """(1 - lambda) * l1_loss(render, target) + lambda * (1 - SSIM(render, target)),
    render/target [C,H,W] or [B,C,H,W] fp32 on the device; differentiable
    w.r.t. render (train.py:255-260)."""
    return _LossFn.apply(render, target.detach(), float(lambda_ssim), padding, False, True)


def fused_ssim(img1: torch.Tensor, img2: torch.Tensor, padding: str = "same", train: bool = True) -> torch.Tensor:
    """Mean SSIM of [B,C,H,W] images, differentiable w.r.t. img1: the call
    signature of fused_ssim.fused_ssim (git a7c48d6) used at train.py:258."""
    return _LossFn.apply(img1, img2.detach(), 1.0, padding, True, bool(train))


class FramebufferLoss:
    """The loss of one render read from the [H,W,C] framebuffer, its
    gradient written into the framebuffer's grad_output_channels (the copy of
    gray/raytracer.py:175-179 disappears).  Keeps its workspace."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.ws = torch.empty(0, dtype=torch.uint8, device=self.device)

    def __call__(self, output_channels: torch.Tensor, grad_output_channels: torch.Tensor, height: int, width: int,
                 target: torch.Tensor, lambda_ssim: float = 0.2, grad_scale: float = 1.0) -> torch.Tensor:
        """output_channels / grad_output_channels: the framebuffer [H_alloc,
        W_alloc, C]; target [C, height, width].  Returns the device tensor
        {loss, mean |render - target|, mean SSIM}; the gradient is grad_scale
        * dloss/drender."""
        ch = output_channels.shape[-1]
        if tuple(target.shape) != (ch, height, width):
            raise ValueError(f"target shape {tuple(target.shape)} != {(ch, height, width)}")
        for t in (output_channels, grad_output_channels):
            if t.dim() != 3 or t.shape[0] < height or t.shape[1] < width or t.shape[2] != ch or t.stride(2) != 1:
                raise ValueError(f"framebuffer tensor of shape {tuple(t.shape)} does not hold {height}x{width}x{ch}")
        _, _, _, ts = _planes(target)
        need = workspace_bytes(ch, height, width)
        if self.ws.numel() < need:
            self.ws = torch.empty(need, dtype=torch.uint8, device=self.device)
        fs = (1, output_channels.stride(0), output_channels.stride(1))
        gs = (1, grad_output_channels.stride(0), grad_output_channels.stride(1))
        loss = torch.empty(3, dtype=F32, device=self.device)
        a = _args(ch, height, width, output_channels, fs, target, ts, lambda_ssim, False, True, self.ws, loss,
                  grad_output_channels, gs, grad_scale)
        sp = L.stream_ptr()
        L.check(L.LIB.gray_loss_forward(C.byref(a), sp))
        L.check(L.LIB.gray_loss_backward(C.byref(a), sp))
        return loss
