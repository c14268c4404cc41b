"""Learned depth prior: a small U-Net over the plain RunPatchMatch output.

No counterpart in the reference, whose prior is the planar/KNN construction
(HPM.cpp:759-1179). The network reads, per pixel, the grey image and the
plain pass's depth, camera-frame normal and cost, and proposes a depth and
normal; the pipeline's judge keeps a proposal only where its multi-view cost
is no worse than the plain hypothesis, and falls back to the planar prior
when too few pixels pass (hpm_pipeline.cpp).

Network (NHWC, channel counts as the BF16 MFMA kernels need them): input 32
channels (6 used), encoder 32-64-96-128 with two 3x3 convolutions + ReLU per
level and 2x2 average pooling between levels, decoder with nearest 2x
upsampling, concatenation of the skip, one 3x3 convolution + ReLU per level,
and a 1x1 head of 32 channels of which 5 are used (depth residual, normal
residual, and the logit of the proposal's confidence: the probability that
its depth lies within 2% of the label). 0.78 M parameters. Weight files of
networks trained without the confidence channel carry MAGIC_NO_CONF; the
pipeline then judges every proposal.

Training runs the convolutions (forward, input gradient, weight gradient) on
the hand-written kernels of include/hpm_unet.h through a torch autograd
function; torch supplies the memory, the elementwise glue and the optimizer.
Inference in the pipeline is the C++ restatement over the same kernels.
"""
from __future__ import annotations

import ctypes as C
import math
import struct
from typing import List, Sequence, Tuple

import numpy as np

from . import hpm

# (name, ci, co, ks, relu); the forward order of the graph below
LAYERS: List[Tuple[str, int, int, int, int]] = [
    ("e0a", 32, 32, 3, 1), ("e0b", 32, 32, 3, 1),
    ("e1a", 32, 64, 3, 1), ("e1b", 64, 64, 3, 1),
    ("e2a", 64, 96, 3, 1), ("e2b", 96, 96, 3, 1),
    ("e3a", 96, 128, 3, 1), ("e3b", 128, 128, 3, 1),
    ("d2", 224, 96, 3, 1), ("d1", 160, 64, 3, 1), ("d0", 96, 32, 3, 1),
    ("head", 32, 32, 1, 0),
]
IN_CHANNELS = 32
MAGIC = b"HPMUNET2"          # head channel 4 is the confidence logit
MAGIC_NO_CONF = b"HPMUNET1"  # no confidence channel


def num_params() -> int:
    return sum(co * ks * ks * ci + co for _, ci, co, ks, _ in LAYERS)


def features(image: np.ndarray, depth: np.ndarray, normal_cam: np.ndarray, cost: np.ndarray, dmin: float,
             dmax: float) -> np.ndarray:
    """Network input (H, W, 32) float32: grey/255, (depth - dmin)/(dmax - dmin),
    camera-frame normal, cost/2, zeros. The pipeline builds the same on the
    device (hpm_k_unet_features)."""
    h, w = depth.shape
    x = np.zeros((h, w, IN_CHANNELS), np.float32)
    x[..., 0] = image.astype(np.float32) / 255.0
    x[..., 1] = (depth - dmin) / (dmax - dmin)
    x[..., 2:5] = normal_cam
    x[..., 5] = cost / 2.0
    return x


def save(path: str, weights: Sequence, biases: Sequence, confidence: bool = True):
    """Weight file: magic, layer count, then per layer (ci, co, ks, relu) as
    int32, BF16 weights [co][ks*ks][ci] and FP32 biases [co]."""
    import torch

    with open(path, "wb") as f:
        f.write(MAGIC if confidence else MAGIC_NO_CONF)
        f.write(struct.pack("<i", len(LAYERS)))
        for (_, ci, co, ks, relu), w, b in zip(LAYERS, weights, biases):
            f.write(struct.pack("<4i", ci, co, ks, relu))
            wb = w.detach().to(torch.bfloat16).contiguous().cpu().view(torch.int16).numpy()
            f.write(wb.astype("<i2").tobytes())
            f.write(b.detach().float().contiguous().cpu().numpy().astype("<f4").tobytes())


def load(path: str):
    """(weights as BF16 torch tensors [co][k*k][ci], biases FP32) on the CPU"""
    import torch

    ws, bs = [], []
    with open(path, "rb") as f:
        if f.read(8) not in (MAGIC, MAGIC_NO_CONF):
            raise ValueError(f"{path}: not a learned-prior weight file")
        (n,) = struct.unpack("<i", f.read(4))
        if n != len(LAYERS):
            raise ValueError(f"{path}: {n} layers, expected {len(LAYERS)}")
        for name, ci, co, ks, relu in LAYERS:
            hdr = struct.unpack("<4i", f.read(16))
            if hdr != (ci, co, ks, relu):
                raise ValueError(f"{path}: layer {name} is {hdr}")
            raw = np.frombuffer(f.read(2 * co * ks * ks * ci), dtype="<i2").copy()
            ws.append(torch.from_numpy(raw).view(torch.bfloat16).reshape(co, ks * ks, ci))
            bs.append(torch.from_numpy(np.frombuffer(f.read(4 * co), dtype="<f4").copy()))
    return ws, bs


# ----------------------------------------------------------------- training

def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _conv_function():
    import torch

    class ConvBF16(torch.autograd.Function):
        """y = conv(x, w) + b on the BF16 MFMA kernels; x (H, W, ci) BF16,
        w (co, k*k, ci) FP32 master weights, y FP32."""

        @staticmethod
        def forward(ctx, x, w, b, ks):
            h, wd, ci = x.shape
            co = w.shape[0]
            wb = w.to(torch.bfloat16).contiguous()
            y = torch.empty(h, wd, co, dtype=torch.float32, device=x.device)
            rc = hpm.lib().hpm_unet_conv_fwd(_ptr(x), ci, ci, _ptr(wb), _ptr(b), _ptr(y), None, co, co, h, wd, ks, 0,
                                             _stream())
            if rc != 0:
                raise RuntimeError(f"hpm_unet_conv_fwd: {rc}")
            ctx.save_for_backward(x, wb)
            ctx.ks = ks
            return y

        @staticmethod
        def backward(ctx, dy):
            x, wb = ctx.saved_tensors
            ks = ctx.ks
            h, wd, ci = x.shape
            co = wb.shape[0]
            dyb = dy.to(torch.bfloat16).contiguous()
            dx = None
            if ctx.needs_input_grad[0]:
                # dX = conv(dY, W'), W'[ci][tap'][co] = W[co][k*k-1-tap'][ci]
                wt = wb.flip(1).permute(2, 1, 0).contiguous()
                dx32 = torch.empty(h, wd, ci, dtype=torch.float32, device=x.device)
                rc = hpm.lib().hpm_unet_conv_fwd(_ptr(dyb), co, co, _ptr(wt), None, _ptr(dx32), None, ci, ci, h, wd,
                                                 ks, 0, _stream())
                if rc != 0:
                    raise RuntimeError(f"hpm_unet_conv_fwd (input gradient): {rc}")
                dx = dx32.to(torch.bfloat16)
            dw = torch.zeros(co, ks * ks, ci, dtype=torch.float32, device=x.device)
            dy_cm = dyb.permute(2, 0, 1).contiguous()
            rc = hpm.lib().hpm_unet_conv_wgrad(_ptr(dy_cm), co, _ptr(x), ci, ci, _ptr(dw), h, wd, ks, _stream())
            if rc != 0:
                raise RuntimeError(f"hpm_unet_conv_wgrad: {rc}")
            db = dy.sum(dim=(0, 1))
            return dx, dw, db, None

    return ConvBF16


def init_params(seed: int = 0):
    import torch

    g = torch.Generator().manual_seed(seed)
    ws, bs = [], []
    for _, ci, co, ks, relu in LAYERS:
        std = math.sqrt(2.0 / (ci * ks * ks)) if relu else 1e-3
        ws.append((std * torch.randn(co, ks * ks, ci, generator=g)).cuda().requires_grad_(True))
        bs.append(torch.zeros(co, device="cuda").requires_grad_(True))
    return ws, bs


def forward(x, ws, bs, conv=None):
    """x (H, W, 32) BF16 CUDA tensor -> head output (H, W, 32) FP32, by the
    graph of the module docstring (torch glue around the MFMA convolutions)."""
    import torch
    import torch.nn.functional as F

    conv = conv or _conv_function()
    L = {name: i for i, (name, *_) in enumerate(LAYERS)}

    def c(name, t):
        i = L[name]
        _, ci, co, ks, relu = LAYERS[i]
        y = conv.apply(t.to(torch.bfloat16).contiguous(), ws[i], bs[i], ks)
        return torch.relu(y) if relu else y

    def pool(t):
        h, w, ch = t.shape
        tt = t.permute(2, 0, 1)[None]
        return F.avg_pool2d(tt, 2, ceil_mode=True, count_include_pad=False)[0].permute(1, 2, 0)

    def up(t, h, w):
        return t.repeat_interleave(2, dim=0)[:h].repeat_interleave(2, dim=1)[:, :w]

    h0, w0, _ = x.shape
    s0 = c("e0b", c("e0a", x))
    p1 = pool(s0)
    s1 = c("e1b", c("e1a", p1))
    p2 = pool(s1)
    s2 = c("e2b", c("e2a", p2))
    p3 = pool(s2)
    b3 = c("e3b", c("e3a", p3))
    d2 = c("d2", torch.cat([up(b3, s2.shape[0], s2.shape[1]), s2], dim=2))
    d1 = c("d1", torch.cat([up(d2, s1.shape[0], s1.shape[1]), s1], dim=2))
    d0 = c("d0", torch.cat([up(d1, h0, w0), s0], dim=2))
    return c("head", d0)


def predict(x, head):
    """(depth_norm, camera normal) from the input features and the head:
    residuals on the plain pass's depth and normal"""
    import torch

    d = torch.clamp(x[..., 1].float() + head[..., 0], 0.0, 1.0)
    n = x[..., 2:5].float() + head[..., 1:4]
    n = n / torch.clamp(torch.linalg.norm(n, dim=-1, keepdim=True), min=1e-6)
    return d, n
