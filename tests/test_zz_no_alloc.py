"""hpm_lprior_forward_noalloc (the network forward over the library's fixed
per-GPU scratch array, no allocation) against hpm_lprior_forward (a fresh
workspace per call), bit for bit: two calls in a row, the second on a
smaller input, so the second call runs over what the first left in the
array; and an input beyond the array's capacity is refused."""
import ctypes as C
import os

import numpy as np
import pytest

from test_gpu_parity import say

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "..", "hpm-mvs_amd", "unet_prior.bin")


def _x(h, w, seed):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand((h, w, 32), generator=g) * 2 - 1).to(torch.bfloat16).cuda().contiguous()


def _run(hpm, fn, lp, x):
    h, w, _ = x.shape
    head = torch.full((h, w, 32), float("nan"), dtype=torch.float32, device="cuda")
    rc = fn(lp, C.c_void_p(x.data_ptr()), h, w, C.c_void_p(head.data_ptr()),
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, head


def test_noalloc_forward_matches_allocating_forward(hpm):
    L = hpm.lib()
    lp = L.hpm_lprior_load(WEIGHTS.encode(), 0)
    assert lp
    cap = L.hpm_lprior_noalloc_capacity()
    say(f"scratch capacity {cap} bytes")
    for i, (h, w) in enumerate([(300, 400), (97, 131)]):
        say(f"case: call {i}: {h}x{w}")
        x = _x(h, w, i)
        rc_a, want = _run(hpm, L.hpm_lprior_forward, lp, x)
        rc_b, got = _run(hpm, L.hpm_lprior_forward_noalloc, lp, x)
        assert rc_a == 0 and rc_b == 0
        assert torch.isfinite(got).all()
        assert torch.equal(got.view(torch.int32), want.view(torch.int32))
    say("case: 1024x1024 beyond the capacity")
    x = _x(1024, 1024, 2)
    rc, _ = _run(hpm, L.hpm_lprior_forward_noalloc, lp, x)
    assert rc == -5
    L.hpm_lprior_destroy(lp)
