"""gray_loss_noalloc (the loss entry point with a library-held static
workspace) against gray_loss_forward + gray_loss_backward with a caller
workspace, through the C ABI: identical loss and gradient (the same kernels
in the same order), on two calls in a row, the second on a smaller image,
so stale contents of the shared array cannot leak into a result."""
import ctypes as C

import pytest
import torch

from gray_amd import _lib as L
from gray_amd.loss import _args, workspace_bytes

pytestmark = pytest.mark.gpu


def images(planes, h, w, seed):
    g = torch.Generator().manual_seed(seed)
    x = (torch.rand(planes, h, w, generator=g) * 1.2 - 0.1).cuda()
    y = torch.rand(planes, h, w, generator=g).cuda()
    return x, y


def strides(t):
    return (t.stride(0), t.stride(1), t.stride(2))


def with_workspace(x, y, lam, valid):
    P, H, W = x.shape
    ws = torch.empty(workspace_bytes(P, H, W), dtype=torch.uint8, device="cuda")
    loss = torch.empty(3, dtype=torch.float32, device="cuda")
    grad = torch.empty_like(x)
    a = _args(P, H, W, x, strides(x), y, strides(y), lam, valid, True, ws, loss, grad, strides(grad), 0.5)
    sp = L.stream_ptr()
    L.check(L.LIB.gray_loss_forward(C.byref(a), sp))
    L.check(L.LIB.gray_loss_backward(C.byref(a), sp))
    return loss, grad


def no_alloc(x, y, lam, valid, with_grad=True):
    P, H, W = x.shape
    loss = torch.full((3,), float("nan"), dtype=torch.float32, device="cuda")
    grad = torch.full_like(x, float("nan")) if with_grad else None
    dummy = torch.empty(1, dtype=torch.uint8, device="cuda")
    a = _args(P, H, W, x, strides(x), y, strides(y), lam, valid, False, dummy, loss, grad,
              strides(grad) if with_grad else (0, 0, 0), 0.5)
    a.workspace = None
    L.check(L.LIB.gray_loss_noalloc(C.byref(a), L.stream_ptr()))
    return loss, grad


def test_noalloc_equals_workspace_entry_points_on_two_calls():
    cases = [((3, 545, 980), 0.2, False, 1), ((3, 37, 23), 0.2, True, 2)]  # second call: a smaller image
    for (P, H, W), lam, valid, seed in cases:
        print(f"noalloc case {P}x{H}x{W} valid={valid}", flush=True)
        x, y = images(P, H, W, seed)
        ref_loss, ref_grad = with_workspace(x, y, lam, valid)
        loss, grad = no_alloc(x, y, lam, valid)
        torch.cuda.synchronize()
        assert torch.equal(loss, ref_loss), (loss, ref_loss)
        assert torch.equal(grad, ref_grad)
        assert float(ref_loss[0]) > 0 and grad.abs().max().item() > 0
    # loss only (no gradient output): the same loss
    x, y = images(3, 64, 48, 3)
    ref_loss, _ = with_workspace(x, y, 0.2, False)
    loss, _ = no_alloc(x, y, 0.2, False, with_grad=False)
    torch.cuda.synchronize()
    assert torch.equal(loss, ref_loss)


def test_noalloc_rejects_oversize_and_a_workspace():
    x, y = images(3, 8, 8, 4)
    loss = torch.empty(3, dtype=torch.float32, device="cuda")
    dummy = torch.empty(1, dtype=torch.uint8, device="cuda")
    a = _args(3, 8, 8, x, strides(x), y, strides(y), 0.2, False, False, dummy, loss)
    # a caller workspace is refused: the library holds the only one
    assert L.LIB.gray_loss_noalloc(C.byref(a), L.stream_ptr()) == L.GRAY_ERR_INVALID
    assert b"workspace must be NULL" in L.LIB.gray_last_error()
    # shapes beyond the static maximum are refused before any device work
    a.workspace = None
    a.planes, a.height, a.width = 3, 2160, 3840
    assert workspace_bytes(3, 2160, 3840) > L.LOSS_NOALLOC_MAX_BYTES
    assert L.LIB.gray_loss_noalloc(C.byref(a), L.stream_ptr()) == L.GRAY_ERR_INVALID
    assert b"above the static" in L.LIB.gray_last_error()
