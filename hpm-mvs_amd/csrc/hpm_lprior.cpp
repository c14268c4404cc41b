/*
 * hpm_lprior.cpp — the learned depth prior in the pipeline: loads the
 * U-Net weight file, runs the network on a finished plain pass (BF16 MFMA
 * convolutions of hpm_unet.hip) and the judge (hpm_k_lprior_judge), and
 * leaves the kept proposals as the handle's planar prior.
 *
 * No reference counterpart (the reference's prior is the planar/KNN
 * construction, HPM.cpp:759-1179). The network graph is the one of
 * hpm-mvs_amd/unet_prior.py: encoder 32-64-96-128 (two 3x3 convolutions +
 * ReLU per level, 2x2 average pooling), decoder with nearest upsampling,
 * skip concatenation and one 3x3 convolution + ReLU per level, 1x1 head.
 * Concatenations are channel slices of one buffer per level:
 * cat0 = [up(d1) 64 | s0 32], cat1 = [up(d2) 96 | s1 64],
 * cat2 = [up(b3) 128 | s2 96].
 */
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "hpm_internal.h"
#include "hpm_unet.h"

namespace {

struct Layer {
    int ci, co, ks, relu;
    void* w = nullptr;     /* BF16 [co][ks*ks][ci] */
    float* b = nullptr;    /* [co] */
};

const int kLayers = 12;
const int kShape[kLayers][4] = {
    {32, 32, 3, 1}, {32, 32, 3, 1}, {32, 64, 3, 1}, {64, 64, 3, 1}, {64, 96, 3, 1}, {96, 96, 3, 1},
    {96, 128, 3, 1}, {128, 128, 3, 1}, {224, 96, 3, 1}, {160, 64, 3, 1}, {96, 32, 3, 1}, {32, 32, 1, 0},
};
enum { E0A, E0B, E1A, E1B, E2A, E2B, E3A, E3B, D2, D1, D0, HEAD };

} // namespace

struct hpm_lprior {
    int device = 0;
    int has_conf = 0; /* head channel 4 is a confidence logit (weight file HPMUNET2) */
    Layer layers[kLayers];
    std::string err;
};

extern "C" void hpm_lprior_destroy(hpm_lprior* lp)
{
    if (!lp) {
        return;
    }
    (void)hipSetDevice(lp->device);
    for (Layer& l : lp->layers) {
        if (l.w) (void)hipFree(l.w);
        if (l.b) (void)hipFree(l.b);
    }
    delete lp;
}

extern "C" hpm_lprior* hpm_lprior_load(const char* path, int device)
{
    if (!path) {
        return nullptr;
    }
    FILE* f = fopen(path, "rb");
    if (!f) {
        return nullptr;
    }
    hpm_lprior* lp = new hpm_lprior();
    lp->device = device < 0 ? 0 : device;
    bool ok = hipSetDevice(lp->device) == hipSuccess;
    char magic[8];
    int32_t n = 0;
    ok = ok && fread(magic, 1, 8, f) == 8 &&
         (memcmp(magic, "HPMUNET1", 8) == 0 || memcmp(magic, "HPMUNET2", 8) == 0) && fread(&n, 4, 1, f) == 1 &&
         n == kLayers;
    lp->has_conf = ok && memcmp(magic, "HPMUNET2", 8) == 0;
    for (int i = 0; ok && i < kLayers; ++i) {
        int32_t hdr[4];
        ok = fread(hdr, 4, 4, f) == 4 && hdr[0] == kShape[i][0] && hdr[1] == kShape[i][1] &&
             hdr[2] == kShape[i][2] && hdr[3] == kShape[i][3];
        if (!ok) {
            break;
        }
        Layer& l = lp->layers[i];
        l.ci = hdr[0];
        l.co = hdr[1];
        l.ks = hdr[2];
        l.relu = hdr[3];
        const size_t nw = (size_t)l.co * l.ks * l.ks * l.ci;
        std::vector<uint16_t> w(nw);
        std::vector<float> b((size_t)l.co);
        ok = fread(w.data(), 2, nw, f) == nw && fread(b.data(), 4, b.size(), f) == b.size() &&
             hipMalloc(&l.w, nw * 2) == hipSuccess && hipMalloc((void**)&l.b, b.size() * 4) == hipSuccess &&
             hipMemcpy(l.w, w.data(), nw * 2, hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(l.b, b.data(), b.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
    }
    fclose(f);
    if (!ok) {
        hpm_lprior_destroy(lp);
        return nullptr;
    }
    return lp;
}

namespace {

/* activation buffers of one forward pass at H x W, packed in one area;
 * with io = false the network input, head and judge counter are left out
 * (the caller brings them) */
struct Plan {
    size_t x, t0, cat0, p1, t1, cat1, p2, t2, cat2, p3, t3, b3, d2, d1, d0, head, cnt, total;
};

Plan plan(int W, int H, bool io = true)
{
    const int w1 = (W + 1) / 2, h1 = (H + 1) / 2;
    const int w2 = (w1 + 1) / 2, h2 = (h1 + 1) / 2;
    const int w3 = (w2 + 1) / 2, h3 = (h2 + 1) / 2;
    const size_t p0 = (size_t)W * H, p1 = (size_t)w1 * h1, p2 = (size_t)w2 * h2, p3 = (size_t)w3 * h3;
    Plan pl;
    size_t total = 0;
    auto take = [&](size_t bytes) {
        const size_t off = total;
        total += (bytes + 255) & ~(size_t)255;
        return off;
    };
    pl.x = io ? take(p0 * 32 * 2) : 0;
    pl.t0 = take(p0 * 32 * 2);
    pl.cat0 = take(p0 * 96 * 2);
    pl.p1 = take(p1 * 32 * 2);
    pl.t1 = take(p1 * 64 * 2);
    pl.cat1 = take(p1 * 160 * 2);
    pl.p2 = take(p2 * 64 * 2);
    pl.t2 = take(p2 * 96 * 2);
    pl.cat2 = take(p2 * 224 * 2);
    pl.p3 = take(p3 * 96 * 2);
    pl.t3 = take(p3 * 128 * 2);
    pl.b3 = take(p3 * 128 * 2);
    pl.d2 = take(p2 * 96 * 2);
    pl.d1 = take(p1 * 64 * 2);
    pl.d0 = take(p0 * 32 * 2);
    pl.head = io ? take(p0 * 32 * 4) : 0;
    pl.cnt = io ? take(8) : 0;
    pl.total = total;
    return pl;
}

/* the network on features x (BF16, W*H*32) into head (FP32, W*H*32); base is
 * a workspace laid out by plan() */
hipError_t forward(const hpm_lprior* lp, const void* x, int W, int H, float* head, char* base, const Plan& pl,
                   hipStream_t s)
{
    const int w1 = (W + 1) / 2, h1 = (H + 1) / 2;
    const int w2 = (w1 + 1) / 2, h2 = (h1 + 1) / 2;
    const int w3 = (w2 + 1) / 2, h3 = (h2 + 1) / 2;
    auto at = [&](size_t off, size_t elem_off = 0) { return (void*)(base + off + elem_off * 2); };
    const Layer* L = lp->layers;
    hipError_t e = hipSuccess;
    auto conv = [&](int li, const void* in, int ldi, void* out, int ldo, int hh, int ww, float* out32) {
        if (e == hipSuccess) {
            const Layer& l = L[li];
            e = (hipError_t)hpm_unet_conv_fwd(in, ldi, l.ci, l.w, l.b, out32, out32 ? nullptr : out, ldo, l.co, hh,
                                              ww, l.ks, l.relu, s);
        }
    };
    auto pool = [&](const void* in, int ldi, void* out, int c, int hh, int ww) {
        if (e == hipSuccess) {
            e = (hipError_t)hpm_unet_avgpool2(in, ldi, out, c, c, hh, ww, s);
        }
    };
    /* decoder layers: the upsampled half of the concatenation is read from
     * the coarser map directly (no upsampled copy) */
    auto conv_up = [&](int li, const void* up, int cup, const void* in, int ldi, void* out, int ldo, int hh,
                       int ww) {
        if (e == hipSuccess) {
            const Layer& l = L[li];
            e = (hipError_t)hpm_unet_conv_fwd_up(in, ldi, l.ci, up, cup, cup, l.w, l.b, nullptr, out, ldo, l.co, hh,
                                                 ww, l.relu, s);
        }
    };
    /* encoder; s0 = cat0[64:96], s1 = cat1[96:160], s2 = cat2[128:224] */
    conv(E0A, x, 32, at(pl.t0), 32, H, W, nullptr);
    conv(E0B, at(pl.t0), 32, at(pl.cat0, 64), 96, H, W, nullptr);
    pool(at(pl.cat0, 64), 96, at(pl.p1), 32, H, W);
    conv(E1A, at(pl.p1), 32, at(pl.t1), 64, h1, w1, nullptr);
    conv(E1B, at(pl.t1), 64, at(pl.cat1, 96), 160, h1, w1, nullptr);
    pool(at(pl.cat1, 96), 160, at(pl.p2), 64, h1, w1);
    conv(E2A, at(pl.p2), 64, at(pl.t2), 96, h2, w2, nullptr);
    conv(E2B, at(pl.t2), 96, at(pl.cat2, 128), 224, h2, w2, nullptr);
    pool(at(pl.cat2, 128), 224, at(pl.p3), 96, h2, w2);
    conv(E3A, at(pl.p3), 96, at(pl.t3), 128, h3, w3, nullptr);
    conv(E3B, at(pl.t3), 128, at(pl.b3), 128, h3, w3, nullptr);
    /* decoder */
    conv_up(D2, at(pl.b3), 128, at(pl.cat2), 224, at(pl.d2), 96, h2, w2);
    conv_up(D1, at(pl.d2), 96, at(pl.cat1), 160, at(pl.d1), 64, h1, w1);
    conv_up(D0, at(pl.d1), 64, at(pl.cat0), 96, at(pl.d0), 32, H, W);
    conv(HEAD, at(pl.d0), 32, nullptr, 32, H, W, head);
    return e;
}

} // namespace

size_t hpm_lprior_activation_bytes(int width, int height)
{
    return plan(width, height, false).total;
}

int hpm_lprior_device(const hpm_lprior* lp)
{
    return lp->device;
}

hipError_t hpm_lprior_forward_ws(const hpm_lprior* lp, const void* x, int height, int width, float* head, void* ws,
                                 hipStream_t stream)
{
    return forward(lp, x, width, height, head, (char*)ws, plan(width, height, false), stream);
}

extern "C" int hpm_lprior_forward(hpm_lprior* lp, const void* x, int height, int width, float* head, void* stream)
{
    if (!lp || !x || !head || height < 1 || width < 1) {
        return (int)hipErrorInvalidValue;
    }
    const Plan pl = plan(width, height);
    char* ws = nullptr;
    hipError_t e = hipMalloc((void**)&ws, pl.total);
    if (e != hipSuccess) {
        return (int)e;
    }
    e = forward(lp, x, width, height, head, ws, pl, (hipStream_t)stream);
    const hipError_t e2 = hipStreamSynchronize((hipStream_t)stream);
    (void)hipFree(ws);
    return (int)(e != hipSuccess ? e : e2);
}

extern "C" int hpm_lprior_apply_timed(hpm_lprior* lp, hpm_handle* h, float* accepted, float* ms)
{
    if (!lp || !h) {
        return HPM_EINVAL;
    }
    if (!h->have_device) {
        return hpm_fail(h, HPM_EINVAL, "learned prior: device initialisation must come first");
    }
    HPM_HIP(h, hipSetDevice(h->device));
    hpm_dev_problem pb;
    hpm_fill_problem(h, &pb);
    const int W = pb.W, H = pb.H;
    const Plan pl = plan(W, H);
    void* ws = nullptr;
    int rc = hpm_workspace(h, pl.total, &ws);
    if (rc != HPM_OK) {
        return rc;
    }
    rc = hpm_prior_on_device(h);
    if (rc != HPM_OK) {
        return rc;
    }
    /* events of this call only: several threads may apply one lp */
    hipEvent_t ev[4] = {};
    hipError_t e = hipSuccess;
    if (ms) {
        for (hipEvent_t& x : ev) {
            if (e == hipSuccess) e = hipEventCreate(&x);
        }
    }
    auto mark = [&](int i) { return ms ? hipEventRecord(ev[i], h->stream) : hipSuccess; };
    char* base = (char*)ws;
    hipStream_t s = h->stream;
    float* head = (float*)(base + pl.head);
    unsigned long long* cnt = (unsigned long long*)(base + pl.cnt);
    unsigned long long kept = 0;
    if (e == hipSuccess) e = mark(0);
    if (e == hipSuccess) e = (hipError_t)hpm_launch_lprior_features(&pb, base + pl.x, s);
    if (e == hipSuccess) e = mark(1);
    if (e == hipSuccess) e = forward(lp, base + pl.x, W, H, head, base, pl, s);
    if (e == hipSuccess) e = mark(2);
    /* judge into the handle's prior */
    if (e == hipSuccess) e = (hipError_t)hpm_launch_zero_u64(cnt, 1, s);
    if (e == hipSuccess) e = (hipError_t)hpm_launch_lprior_judge(&pb, base + pl.x, head, h->d_prior, h->d_mask, cnt, lp->has_conf, s);
    if (e == hipSuccess) e = mark(3);
    if (e == hipSuccess) e = hipMemcpyAsync(&kept, cnt, 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (ms) {
        for (int i = 0; i < 3; ++i) {
            if (e == hipSuccess) e = hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]);
        }
        for (hipEvent_t& x : ev) {
            if (x) (void)hipEventDestroy(x);
        }
    }
    HPM_HIP(h, e);
    if (accepted) {
        *accepted = (float)((double)kept / ((double)W * H));
    }
    return HPM_OK;
}

extern "C" int hpm_lprior_apply(hpm_lprior* lp, hpm_handle* h, float* accepted)
{
    return hpm_lprior_apply_timed(lp, h, accepted, nullptr);
}
