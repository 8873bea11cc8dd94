"""3-level 3D U-Net preconditioner: channels 16/32/64, two 3x3x3 BF16 convs
with ReLU per level (FP32 accumulation, flub_conv3d_bf16), 2x2x2 average
pooling down, nearest 2x upsampling and skip concatenation up, a 1x1 FP32
head.  Input channels: the normalised residual and the FLUID indicator
(padded to 16 channels for the matrix-core K width).  Spatial sizes must be
multiples of 4.

Forward cost per voxel: 2*27*(16*16 + 16*16 + (16*32 + 32*32)/8 + (32*64 +
64*64)/64 + (96*32 + 32*32)/8 + 48*16 + 16*16) ~= 1.1e5 flop.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _lib
from .conv import Conv3dFn, conv3d

# (name, cin, cout); every conv is followed by ReLU
LAYERS = [("e0a", 16, 16), ("e0b", 16, 16), ("e1a", 16, 32), ("e1b", 32, 32), ("e2a", 32, 64), ("e2b", 64, 64),
          ("d1a", 96, 32), ("d1b", 32, 32), ("d0a", 48, 16), ("d0b", 16, 16)]
IN_CH = 2
PAD_CH = 16


def flops_per_voxel():
    f = 0.0
    scale = {"e0": 1, "e1": 1 / 8, "e2": 1 / 64, "d1": 1 / 8, "d0": 1}
    for name, cin, cout in LAYERS:
        f += 2 * 27 * cin * cout * scale[name[:2]]
    return f + 2 * 16


def pool2(x):
    n, d, h, w, c = x.shape
    return x.view(n, d // 2, 2, h // 2, 2, w // 2, 2, c).float().mean(dim=(2, 4, 6)).to(x.dtype).contiguous()


def up2(x):
    n, d, h, w, c = x.shape
    return (x.view(n, d, 1, h, 1, w, 1, c).expand(n, d, 2, h, 2, w, 2, c).reshape(n, 2 * d, 2 * h, 2 * w, c)
            .contiguous())


class UNet(torch.nn.Module):
    def __init__(self, seed: int = 0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.w = torch.nn.ParameterDict()
        self.b = torch.nn.ParameterDict()
        for name, cin, cout in LAYERS:
            fan_in = 27 * (IN_CH if name == "e0a" else cin)
            std = math.sqrt(2.0 / fan_in)
            w = torch.randn(cout, 27, cin, generator=g) * std
            if name == "e0a":
                w[:, :, IN_CH:] = 0.0
            self.w[name] = torch.nn.Parameter(w.reshape(cout, 27 * cin))
            self.b[name] = torch.nn.Parameter(torch.zeros(cout))
        self.head_w = torch.nn.Parameter(torch.randn(16, generator=g) * math.sqrt(1.0 / 16))
        self.head_b = torch.nn.Parameter(torch.zeros(1))
        self.out_scale = 1.0   # the network predicts target / out_scale

    def _c(self, x, name):
        return Conv3dFn.apply(x, self.w[name], self.b[name], True)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """x: (N, D, H, W, 2) float -> (N, D, H, W) float."""
        n, d, h, w, c = x.shape
        assert c == IN_CH and d % 4 == 0 and h % 4 == 0 and w % 4 == 0
        x16 = torch.zeros((n, d, h, w, PAD_CH), device=x.device, dtype=torch.bfloat16)
        x16[..., :IN_CH] = x.to(torch.bfloat16)
        e0 = self._c(self._c(x16, "e0a"), "e0b")
        e1 = self._c(self._c(pool2(e0), "e1a"), "e1b")
        e2 = self._c(self._c(pool2(e1), "e2a"), "e2b")
        d1 = self._c(self._c(torch.cat([up2(e2), e1], dim=-1), "d1a"), "d1b")
        d0 = self._c(self._c(torch.cat([up2(d1), e0], dim=-1), "d0a"), "d0b")
        return (d0.float() @ self.head_w) + self.head_b

    def _bf16_weights(self):
        key = tuple(int(v._version) for v in self.w.values())
        if getattr(self, "_wt_key", None) != key:
            self._wt = {k: v.detach().to(torch.bfloat16).contiguous() for k, v in self.w.items()}
            self._wt_key = key
        return self._wt

    def _trunk(self, x16: torch.Tensor) -> torch.Tensor:
        """The network up to the head on a 16-channel BF16 input: the LDS-tiled
        convolutions, BF16 pooling, and the decoder's up2 + cat folded into its
        first convolution of each level (flub_conv3d_upcat_bf16) where the
        volume allows, the unfused sequence elsewhere."""
        n, d, h, w, _ = x16.shape
        wt = self._bf16_weights()
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

        def c3(t, name):
            return conv3d(t, wt[name], self.b[name].detach(), True)

        def pool(t):
            nn, dd, hh, ww, cc = t.shape
            y = torch.empty((nn, dd // 2, hh // 2, ww // 2, cc), device=t.device, dtype=torch.bfloat16)
            _lib.call("flub_pool2_bf16", ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(y.data_ptr()), nn, dd, hh, ww,
                      cc, st)
            return y

        def upcat(lo, hi, name):
            nn, dd, hh, ww, ch = hi.shape
            cin, cout = lo.shape[-1] + ch, wt[name].shape[0]
            if ww % 16 == 0 and hh % 8 == 0 and dd % 8 == 0:
                y = torch.empty((nn, dd, hh, ww, cout), device=hi.device, dtype=torch.bfloat16)
                _lib.call("flub_conv3d_upcat_bf16", ctypes.c_void_p(lo.data_ptr()), lo.shape[-1],
                          ctypes.c_void_p(hi.data_ptr()), ctypes.c_void_p(wt[name].data_ptr()),
                          ctypes.c_void_p(self.b[name].detach().data_ptr()), ctypes.c_void_p(y.data_ptr()), None, nn,
                          dd, hh, ww, cin, cout, 1, st)
                return y
            return c3(torch.cat([up2(lo), hi], dim=-1), name)

        e0 = c3(c3(x16, "e0a"), "e0b")
        e1 = c3(c3(pool(e0), "e1a"), "e1b")
        e2 = c3(c3(pool(e1), "e2a"), "e2b")
        d1 = c3(upcat(e2, e1, "d1a"), "d1b")
        del e1, e2
        return c3(upcat(d1, e0, "d0a"), "d0b")

    @torch.no_grad()
    def infer(self, x: torch.Tensor) -> torch.Tensor:
        """Forward without autograd bookkeeping: x (N, D, H, W, 2) float -> (N, D, H, W)."""
        n, d, h, w, c = x.shape
        x16 = torch.zeros((n, d, h, w, PAD_CH), device=x.device, dtype=torch.bfloat16)
        x16[..., :IN_CH] = x.to(torch.bfloat16)
        d0 = self._trunk(x16)
        y = torch.empty((n, d, h, w), device=x.device, dtype=torch.float32)
        hw = self.head_w.detach().float().contiguous()
        _lib.call("flub_head_fwd", ctypes.c_void_p(d0.data_ptr()), ctypes.c_void_p(hw.data_ptr()),
                  float(self.head_b.detach()[0]), ctypes.c_void_p(y.data_ptr()), y.numel(), d0.shape[-1],
                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        return y

    @torch.no_grad()
    def precondition(self, r: torch.Tensor, fluid_u8: torch.Tensor, s: float) -> torch.Tensor:
        """z = s * out_scale * U-Net([r/s, FLUID]) on FLUID cells, 0 elsewhere, for
        one (D, H, W) FP32 residual; the input packing and the output scaling
        and masking are fused into the first and last kernels."""
        d, h, w = r.shape
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        x16 = torch.empty((1, d, h, w, PAD_CH), device=r.device, dtype=torch.bfloat16)
        _lib.call("flub_pack_input_bf16", ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(fluid_u8.data_ptr()),
                  float(s), ctypes.c_void_p(x16.data_ptr()), r.numel(), st)
        d0 = self._trunk(x16)
        del x16
        z = torch.empty((d, h, w), device=r.device, dtype=torch.float32)
        hw = self.head_w.detach().float().contiguous()
        _lib.call("flub_head_fwd_masked", ctypes.c_void_p(d0.data_ptr()), ctypes.c_void_p(hw.data_ptr()),
                  float(self.head_b.detach()[0]), ctypes.c_void_p(z.data_ptr()), z.numel(), d0.shape[-1],
                  ctypes.c_void_p(fluid_u8.data_ptr()), float(s * self.out_scale), st)
        return z

    def state_numpy(self):
        out = {f"w_{k}": v.detach().cpu().numpy() for k, v in self.w.items()}
        out.update({f"b_{k}": v.detach().cpu().numpy() for k, v in self.b.items()})
        out["head_w"] = self.head_w.detach().cpu().numpy()
        out["head_b"] = self.head_b.detach().cpu().numpy()
        out["out_scale"] = np.float64(self.out_scale)
        return out

    def load_numpy(self, d):
        with torch.no_grad():
            for k in self.w:
                self.w[k].copy_(torch.from_numpy(np.asarray(d[f"w_{k}"])))
                self.b[k].copy_(torch.from_numpy(np.asarray(d[f"b_{k}"])))
            self.head_w.copy_(torch.from_numpy(np.asarray(d["head_w"])))
            self.head_b.copy_(torch.from_numpy(np.asarray(d["head_b"])))
        self.out_scale = float(d.get("out_scale", 1.0))
