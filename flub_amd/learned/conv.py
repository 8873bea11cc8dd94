"""The U-Net's 3x3x3 convolutions as a torch autograd Function over the HIP
matrix-core kernels of libflub.so (flub_conv3d_bf16 / flub_conv3d_wgrad_bf16).

Torch provides memory, streams and the optimizer; every convolution of the
forward and backward pass (output, data gradient, weight gradient) runs in
the project's BF16 kernels with FP32 accumulation.  Tensors are channels-last
(N, D, H, W, C) BF16; weights are FP32 masters of shape (Cout, 27*Cin) cast to
BF16 per call.
"""
from __future__ import annotations

import ctypes

import torch

from .. import _lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def conv3d(x: torch.Tensor, wt: torch.Tensor, bias: torch.Tensor | None, relu: bool,
           out_fp32: bool = False) -> torch.Tensor:
    """y = [ReLU](conv3x3x3(x, wt) + bias); x BF16 (N,D,H,W,Cin), wt BF16 (Cout, 27*Cin)."""
    assert x.dtype == torch.bfloat16 and wt.dtype == torch.bfloat16 and x.is_contiguous() and wt.is_contiguous()
    n, d, h, w, cin = x.shape
    cout = wt.shape[0]
    assert wt.shape[1] == 27 * cin
    if out_fp32:
        y = torch.empty((n, d, h, w, cout), device=x.device, dtype=torch.float32)
        _lib.call("flub_conv3d_bf16", _p(x), _p(wt), _p(bias), None, _p(y), n, d, h, w, cin, cout, int(relu),
                  _stream())
    else:
        y = torch.empty((n, d, h, w, cout), device=x.device, dtype=torch.bfloat16)
        _lib.call("flub_conv3d_bf16", _p(x), _p(wt), _p(bias), _p(y), None, n, d, h, w, cin, cout, int(relu),
                  _stream())
    return y


_WORK = {}


def conv3d_wgrad(x: torch.Tensor, dy: torch.Tensor) -> torch.Tensor:
    """dL/dW (Cout, 27*Cin) FP32 from the BF16 input and output gradient."""
    assert x.dtype == torch.bfloat16 and dy.dtype == torch.bfloat16
    x = x.contiguous()
    dy = dy.contiguous()
    n, d, h, w, cin = x.shape
    cout = dy.shape[-1]
    need = 256 * cout * 27 * cin
    key = x.device
    work = _WORK.get(key)
    if work is None or work.numel() < need:
        work = torch.empty(need, device=x.device, dtype=torch.float32)
        _WORK[key] = work
    dw = torch.empty((cout, 27 * cin), device=x.device, dtype=torch.float32)
    _lib.call("flub_conv3d_wgrad_bf16", _p(x), _p(dy), _p(dw), _p(work), work.numel(), n, d, h, w, cin, cout,
              _stream())
    return dw


def flip_for_dgrad(w32: torch.Tensor, cin: int) -> torch.Tensor:
    """Weights of the adjoint convolution: taps mirrored, channel roles swapped:
    (Cout, 27*Cin) -> (Cin, 27*Cout)."""
    cout = w32.shape[0]
    return w32.view(cout, 27, cin).flip(1).permute(2, 1, 0).contiguous().view(cin, 27 * cout)


class Conv3dFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w32, b32, relu):
        y = conv3d(x, w32.to(torch.bfloat16), b32, relu)
        ctx.save_for_backward(x, w32, y)
        ctx.relu = relu
        return y

    @staticmethod
    def backward(ctx, gy):
        x, w32, y = ctx.saved_tensors
        gy = gy.to(torch.bfloat16)
        if ctx.relu:
            gy = torch.where(y > 0, gy, torch.zeros_like(gy))
        gy = gy.contiguous()
        db = gy.float().sum(dim=(0, 1, 2, 3))
        dw = conv3d_wgrad(x, gy)
        dx = None
        if ctx.needs_input_grad[0]:
            wd = flip_for_dgrad(w32, x.shape[-1]).to(torch.bfloat16)
            dx = conv3d(gy, wd, None, False)
        return dx, dw, db, None
