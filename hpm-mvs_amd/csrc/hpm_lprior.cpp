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
    ok = ok && fread(magic, 1, 8, f) == 8 && memcmp(magic, "HPMUNET1", 8) == 0 && fread(&n, 4, 1, f) == 1 &&
         n == kLayers;
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

extern "C" int hpm_lprior_apply(hpm_lprior* lp, hpm_handle* h, float* accepted)
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
    const int w1 = (W + 1) / 2, h1 = (H + 1) / 2;
    const int w2 = (w1 + 1) / 2, h2 = (h1 + 1) / 2;
    const int w3 = (w2 + 1) / 2, h3 = (h2 + 1) / 2;
    const size_t p0 = (size_t)W * H, p1 = (size_t)w1 * h1, p2 = (size_t)w2 * h2, p3 = (size_t)w3 * h3;
    /* workspace (BF16 elements unless noted) */
    struct Buf {
        size_t off, bytes;
    };
    size_t total = 0;
    auto take = [&](size_t bytes) {
        Buf b{total, bytes};
        total += (bytes + 255) & ~(size_t)255;
        return b;
    };
    const Buf bx = take(p0 * 32 * 2), bt0 = take(p0 * 32 * 2), bcat0 = take(p0 * 96 * 2);
    const Buf bp1 = take(p1 * 32 * 2), bt1 = take(p1 * 64 * 2), bcat1 = take(p1 * 160 * 2);
    const Buf bp2 = take(p2 * 64 * 2), bt2 = take(p2 * 96 * 2), bcat2 = take(p2 * 224 * 2);
    const Buf bp3 = take(p3 * 96 * 2), bt3 = take(p3 * 128 * 2), bb3 = take(p3 * 128 * 2);
    const Buf bd2 = take(p2 * 96 * 2), bd1 = take(p1 * 64 * 2), bd0 = take(p0 * 32 * 2);
    const Buf bhead = take(p0 * 32 * 4), bcnt = take(8);
    void* ws = nullptr;
    int rc = hpm_workspace(h, total, &ws);
    if (rc != HPM_OK) {
        return rc;
    }
    char* base = (char*)ws;
    auto at = [&](const Buf& b, size_t elem_off = 0) { return (void*)(base + b.off + elem_off * 2); };
    hipStream_t s = h->stream;
    const Layer* L = lp->layers;
    auto conv = [&](int li, const void* in, int ldi, void* out, int ldo, int hh, int ww, float* out32) {
        const Layer& l = L[li];
        return (hipError_t)hpm_unet_conv_fwd(in, ldi, l.ci, l.w, l.b, out32, out32 ? nullptr : out, ldo, l.co, hh,
                                             ww, l.ks, l.relu, s);
    };
    HPM_HIP(h, (hipError_t)hpm_launch_lprior_features(&pb, at(bx), s));
    /* encoder; s0 = cat0[64:96], s1 = cat1[96:160], s2 = cat2[128:224] */
    HPM_HIP(h, conv(E0A, at(bx), 32, at(bt0), 32, H, W, nullptr));
    HPM_HIP(h, conv(E0B, at(bt0), 32, at(bcat0, 64), 96, H, W, nullptr));
    HPM_HIP(h, (hipError_t)hpm_unet_avgpool2(at(bcat0, 64), 96, at(bp1), 32, 32, H, W, s));
    HPM_HIP(h, conv(E1A, at(bp1), 32, at(bt1), 64, h1, w1, nullptr));
    HPM_HIP(h, conv(E1B, at(bt1), 64, at(bcat1, 96), 160, h1, w1, nullptr));
    HPM_HIP(h, (hipError_t)hpm_unet_avgpool2(at(bcat1, 96), 160, at(bp2), 64, 64, h1, w1, s));
    HPM_HIP(h, conv(E2A, at(bp2), 64, at(bt2), 96, h2, w2, nullptr));
    HPM_HIP(h, conv(E2B, at(bt2), 96, at(bcat2, 128), 224, h2, w2, nullptr));
    HPM_HIP(h, (hipError_t)hpm_unet_avgpool2(at(bcat2, 128), 224, at(bp3), 96, 96, h2, w2, s));
    HPM_HIP(h, conv(E3A, at(bp3), 96, at(bt3), 128, h3, w3, nullptr));
    HPM_HIP(h, conv(E3B, at(bt3), 128, at(bb3), 128, h3, w3, nullptr));
    /* decoder */
    HPM_HIP(h, (hipError_t)hpm_unet_upsample2(at(bb3), 128, at(bcat2), 224, 128, h2, w2, s));
    HPM_HIP(h, conv(D2, at(bcat2), 224, at(bd2), 96, h2, w2, nullptr));
    HPM_HIP(h, (hipError_t)hpm_unet_upsample2(at(bd2), 96, at(bcat1), 160, 96, h1, w1, s));
    HPM_HIP(h, conv(D1, at(bcat1), 160, at(bd1), 64, h1, w1, nullptr));
    HPM_HIP(h, (hipError_t)hpm_unet_upsample2(at(bd1), 64, at(bcat0), 96, 64, H, W, s));
    HPM_HIP(h, conv(D0, at(bcat0), 96, at(bd0), 32, H, W, nullptr));
    float* head = (float*)(base + bhead.off);
    HPM_HIP(h, conv(HEAD, at(bd0), 32, nullptr, 32, H, W, head));
    /* judge into the handle's prior */
    rc = hpm_prior_on_device(h);
    if (rc != HPM_OK) {
        return rc;
    }
    unsigned long long* cnt = (unsigned long long*)(base + bcnt.off);
    HPM_HIP(h, (hipError_t)hpm_launch_zero_u64(cnt, 1, s));
    HPM_HIP(h, (hipError_t)hpm_launch_lprior_judge(&pb, at(bx), head, h->d_prior, h->d_mask, cnt, s));
    unsigned long long kept = 0;
    HPM_HIP(h, hipMemcpyAsync(&kept, cnt, 8, hipMemcpyDeviceToHost, s));
    HPM_HIP(h, hipStreamSynchronize(s));
    if (accepted) {
        *accepted = (float)((double)kept / (double)p0);
    }
    return HPM_OK;
}
