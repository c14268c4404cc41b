"""CPU restatement of the learned prior's network graph
(hpm-mvs_amd/csrc/hpm_lprior.cpp: forward; hpm-mvs_amd/unet_prior.py).

TEST INFRASTRUCTURE: the CPU checker of the HIP path. Only tests/, the smoke
check and bench.py's CPU-baseline leg use it.

In float64 with BF16 rounding where the device stores BF16 (every convolution
output and pooled map) it is the parity reference of the device forward; in
float32 without the rounding it is the CPU baseline the device is timed
against. No reference counterpart (the reference has no learned prior).
"""
import torch
import torch.nn.functional as F


def forward(x, ws, bs, layers, dtype=torch.float64, store_bf16=True):
    """x (H, W, 32) CPU tensor; ws[i] (co, ks*ks, ci), bs[i] (co); layers the
    (name, ci, co, ks, relu) table of unet_prior.LAYERS. Returns the head
    (H, W, 32) in dtype."""
    L = {name: i for i, (name, *_) in enumerate(layers)}

    def bf(t):
        return t.to(torch.bfloat16).to(dtype) if store_bf16 else t

    def conv(name, t, last=False):
        i = L[name]
        _, ci, co, ks, relu = layers[i]
        w = ws[i].detach().cpu().to(dtype).reshape(co, ks, ks, ci).permute(0, 3, 1, 2)
        b = bs[i].detach().cpu().to(dtype)
        y = F.conv2d(t.permute(2, 0, 1)[None], w, b, padding=ks // 2)[0].permute(1, 2, 0)
        if relu:
            y = torch.relu(y)
        return y if last else bf(y)

    def pool(t):
        # 2x2 mean over the pixels inside the image (ceil size, partial
        # windows at odd edges average what they hold)
        h, w, c = t.shape
        oh, ow = (h + 1) // 2, (w + 1) // 2
        tp = F.pad(t.permute(2, 0, 1), (0, 2 * ow - w, 0, 2 * oh - h))
        n = F.pad(torch.ones(1, h, w, dtype=dtype), (0, 2 * ow - w, 0, 2 * oh - h))
        s = F.avg_pool2d(tp[None], 2, divisor_override=1)[0]
        cnt = F.avg_pool2d(n[None], 2, divisor_override=1)[0]
        return bf((s / cnt).permute(1, 2, 0))

    def upsample(t, h, w):
        return t[torch.arange(h) // 2][:, torch.arange(w) // 2]

    x = x.to(dtype)
    s0 = conv("e0b", conv("e0a", x))
    s1 = conv("e1b", conv("e1a", pool(s0)))
    s2 = conv("e2b", conv("e2a", pool(s1)))
    b3 = conv("e3b", conv("e3a", pool(s2)))
    d2 = conv("d2", torch.cat([upsample(b3, *s2.shape[:2]), s2], dim=2))
    d1 = conv("d1", torch.cat([upsample(d2, *s1.shape[:2]), s1], dim=2))
    d0 = conv("d0", torch.cat([upsample(d1, *s0.shape[:2]), s0], dim=2))
    return conv("head", d0, last=True)


def flops(layers, height, width):
    """2 x H_l x W_l x C_in x C_out x k^2 summed over the layers: level l of
    the encoder/decoder runs at ceil(H / 2^l) x ceil(W / 2^l)."""
    level = {"e0a": 0, "e0b": 0, "e1a": 1, "e1b": 1, "e2a": 2, "e2b": 2, "e3a": 3, "e3b": 3, "d2": 2, "d1": 1,
             "d0": 0, "head": 0}
    dims = [(height, width)]
    for _ in range(3):
        h, w = dims[-1]
        dims.append(((h + 1) // 2, (w + 1) // 2))
    total = 0
    for name, ci, co, ks, _ in layers:
        h, w = dims[level[name]]
        total += 2 * h * w * ci * co * ks * ks
    return total
