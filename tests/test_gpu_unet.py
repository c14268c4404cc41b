"""BF16 MFMA building blocks of the learned prior (include/hpm_unet.h)
against FP64 CPU references (torch on the CPU in float64 from the same
BF16-rounded values).

Bars: the MFMA lane layout is exactly the documented one; convolution
forward (3x3 and 1x1, ReLU, channel-strided slices) within 1e-4 relative of
the FP64 result (FP32 accumulation of exact BF16 products); the weight
gradient within 1e-3 relative (FP32 atomics); pooling and upsampling exact
up to BF16 rounding of the mean.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(msg, flush=True)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_mfma_lane_layout(hpm):
    say("case: MFMA 32x32x16 bf16 lane layout probe")
    A = torch.zeros(32, 16, dtype=torch.float64)
    for m in range(16):
        A[m, m] = 1.0
    kk, nn = torch.meshgrid(torch.arange(16), torch.arange(32), indexing="ij")
    B = ((kk * 37 + nn * 11) % 200).double()
    D = A @ B
    lanes = torch.arange(64)
    a = torch.stack([A[lanes % 32, 8 * (lanes // 32) + e] for e in range(8)], dim=1)
    b = torch.stack([B[8 * (lanes // 32) + e, lanes % 32] for e in range(8)], dim=1)
    a_d = a.to(torch.bfloat16).cuda().contiguous()
    b_d = b.to(torch.bfloat16).cuda().contiguous()
    d_d = torch.zeros(64, 16, dtype=torch.float32, device="cuda")
    assert hpm.lib().hpm_unet_mfma_probe(_p(a_d), _p(b_d), _p(d_d), _stream()) == 0
    torch.cuda.synchronize()
    d = d_d.cpu().double()
    want = torch.empty(64, 16, dtype=torch.float64)
    for q in range(16):
        want[:, q] = D[8 * (q // 4) + 4 * (lanes // 32) + q % 4, lanes % 32]
    assert torch.equal(d, want)


def _conv_ref(x_bf, w_bf, bias, ks, relu):
    """x (H, W, Ci), w (Co, ks*ks, Ci) -> (H, W, Co) in float64"""
    x = x_bf.double().permute(2, 0, 1)[None]
    co, taps, ci = w_bf.shape
    w = w_bf.double().reshape(co, ks, ks, ci).permute(0, 3, 1, 2)
    y = torch.nn.functional.conv2d(x, w, bias.double(), padding=ks // 2)[0].permute(1, 2, 0)
    return torch.relu(y) if relu else y


@pytest.mark.parametrize("ks,relu,ci,co,h,w", [(3, 1, 32, 64, 13, 45), (3, 0, 16, 32, 7, 33), (1, 0, 96, 32, 5, 70),
                                                (3, 1, 96, 96, 9, 131), (3, 0, 128, 128, 6, 257), (3, 1, 224, 96, 5, 140),
                                                (3, 1, 64, 96, 3, 300), (3, 0, 160, 64, 4, 129), (3, 1, 96, 32, 7, 127),
                                                (3, 1, 32, 128, 37, 1700), (3, 0, 64, 96, 5, 1601),
                                                (3, 1, 32, 128, 257, 200)])
def test_conv_forward(hpm, ks, relu, ci, co, h, w):
    say(f"case: conv forward ks={ks} relu={relu} ci={ci} co={co} {h}x{w}")
    g = torch.Generator().manual_seed(ks * 100 + ci)
    ldi, ldo = ci + 16, co + 32  # channel slices of wider buffers
    x = torch.randn(h, w, ldi, generator=g).to(torch.bfloat16)
    wt = (0.1 * torch.randn(co, ks * ks, ci, generator=g)).to(torch.bfloat16)
    bias = torch.randn(co, generator=g).float()
    out = torch.full((h, w, ldo), -7.0, dtype=torch.float32, device="cuda")
    x_d, w_d, b_d = x.cuda(), wt.cuda(), bias.cuda()
    # input slice starts 16 channels in, output slice 32 channels in
    rc = hpm.lib().hpm_unet_conv_fwd(C.c_void_p(x_d.data_ptr() + 16 * 2), ldi, ci, _p(w_d), _p(b_d),
                                     C.c_void_p(out.data_ptr() + 32 * 4), None, ldo, co, h, w, ks, relu, _stream())
    assert rc == 0
    torch.cuda.synchronize()
    got = out.cpu().double()
    ref = _conv_ref(x[:, :, 16:16 + ci], wt, bias, ks, relu)
    assert torch.all(got[:, :, :32] == -7.0)  # untouched outside the slice
    err = (got[:, :, 32:32 + co] - ref).abs().max().item()
    scale = ref.abs().max().item()
    say(f"max abs err {err:.3e} of {scale:.3e}")
    assert err <= 1e-4 * max(scale, 1.0)


@pytest.mark.parametrize("ci,cup,co,h,w", [(224, 128, 96, 9, 131), (160, 96, 64, 8, 130), (96, 64, 32, 7, 257),
                                           (96, 64, 32, 6, 129)])
def test_conv_with_fused_upsampling(hpm, ci, cup, co, h, w):
    """hpm_unet_conv_fwd_up (channels [0, cup) read as the nearest 2x
    upsampling of a coarser map) equals hpm_unet_conv_fwd on the explicitly
    upsampled concatenation, bit for bit"""
    say(f"case: conv with fused upsampling ci={ci} cup={cup} co={co} {h}x{w}")
    g = torch.Generator().manual_seed(ci + cup)
    hu, wu = (h + 1) // 2, (w + 1) // 2
    up = torch.randn(hu, wu, cup, generator=g).to(torch.bfloat16)
    skip = torch.randn(h, w, ci - cup, generator=g).to(torch.bfloat16)
    cat = torch.cat([up.repeat_interleave(2, 0).repeat_interleave(2, 1)[:h, :w], skip], dim=2).contiguous()
    wt = (0.1 * torch.randn(co, 9, ci, generator=g)).to(torch.bfloat16)
    bias = torch.randn(co, generator=g).float()
    a = torch.zeros(h, w, co, dtype=torch.bfloat16, device="cuda")
    b = torch.zeros(h, w, co, dtype=torch.bfloat16, device="cuda")
    cat_d, up_d, w_d, b_d = cat.cuda(), up.cuda(), wt.cuda(), bias.cuda()
    # the fused call reads the skip channels from a buffer whose first cup
    # channels were never written
    holey = cat.clone()
    holey[:, :, :cup] = float("nan")
    holey_d = holey.cuda()
    assert hpm.lib().hpm_unet_conv_fwd(_p(cat_d), ci, ci, _p(w_d), _p(b_d), None, _p(a), co, co, h, w, 3, 1,
                                       _stream()) == 0
    assert hpm.lib().hpm_unet_conv_fwd_up(_p(holey_d), ci, ci, _p(up_d), cup, cup, _p(w_d), _p(b_d), None, _p(b), co,
                                          co, h, w, 1, _stream()) == 0
    torch.cuda.synchronize()
    assert torch.equal(a.cpu().view(torch.int16), b.cpu().view(torch.int16))


def test_conv_weight_gradient(hpm):
    say("case: conv weight gradient 3x3 ci=32 co=64 19x37")
    h, w, ci, co, ks = 19, 37, 32, 64, 3
    g = torch.Generator().manual_seed(5)
    x = torch.randn(h, w, ci, generator=g).to(torch.bfloat16)
    dy = torch.randn(co, h * w, generator=g).to(torch.bfloat16)
    dw = torch.zeros(co, ks * ks, ci, dtype=torch.float32, device="cuda")
    x_d, dy_d = x.cuda(), dy.cuda()
    assert hpm.lib().hpm_unet_conv_wgrad(_p(dy_d), co, _p(x_d), ci, ci, _p(dw), h, w, ks, _stream()) == 0
    torch.cuda.synchronize()
    # reference: d/dw of sum(dy * conv(x, w))
    xw = x.double().permute(2, 0, 1)[None]
    wv = torch.zeros(co, ci, ks, ks, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.conv2d(xw, wv, padding=1)[0].reshape(co, h * w)
    (y * dy.double()).sum().backward()
    ref = wv.grad.permute(0, 2, 3, 1).reshape(co, ks * ks, ci)
    err = (dw.cpu().double() - ref).abs().max().item()
    scale = ref.abs().max().item()
    say(f"max abs err {err:.3e} of {scale:.3e}")
    assert err <= 1e-3 * scale


def test_pool_upsample(hpm):
    say("case: 2x2 average pooling and nearest upsampling, 9x13x32")
    h, w, c = 9, 13, 32
    g = torch.Generator().manual_seed(7)
    x = torch.randn(h, w, c, generator=g).to(torch.bfloat16)
    oh, ow = (h + 1) // 2, (w + 1) // 2
    pooled = torch.empty(oh, ow, c, dtype=torch.bfloat16, device="cuda")
    x_d = x.cuda()
    assert hpm.lib().hpm_unet_avgpool2(_p(x_d), c, _p(pooled), c, c, h, w, _stream()) == 0
    up = torch.empty(h, w, c, dtype=torch.bfloat16, device="cuda")
    assert hpm.lib().hpm_unet_upsample2(_p(pooled), c, _p(up), c, c, h, w, _stream()) == 0
    torch.cuda.synchronize()
    xp = torch.nn.functional.pad(x.double().permute(2, 0, 1), (0, ow * 2 - w, 0, oh * 2 - h))
    cnt = torch.nn.functional.pad(torch.ones(1, h, w, dtype=torch.float64), (0, ow * 2 - w, 0, oh * 2 - h))
    s = torch.nn.functional.avg_pool2d(xp[None], 2, divisor_override=1)[0]
    n = torch.nn.functional.avg_pool2d(cnt[None], 2, divisor_override=1)[0]
    ref = (s / n).permute(1, 2, 0).float().to(torch.bfloat16)
    assert torch.equal(pooled.cpu().float(), ref.float())
    ref_up = pooled.cpu()[torch.arange(h) // 2][:, torch.arange(w) // 2]
    assert torch.equal(up.cpu(), ref_up)
