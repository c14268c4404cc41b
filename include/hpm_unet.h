/*
 * hpm_unet.h — BF16 MFMA building blocks of the learned depth prior.
 *
 * No counterpart in the reference (its prior is the planar/KNN
 * construction, HPM.cpp:759-1179). Activations are NHWC BF16 per image with
 * a channel stride (ld, a multiple of 8) so concatenations are channel
 * slices of one buffer; weights are [Co][k*k][Ci] BF16 with Ci a multiple of
 * 16 and Co of 32. All pointers are device pointers; every call launches on
 * `stream` and returns a hipError_t value (0 = success).
 */
#ifndef HPM_UNET_H_
#define HPM_UNET_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One 32x32x16 BF16 MFMA of per-lane operands (a, b: 64 x 8 BF16; d: 64 x 16
 * FP32), to check the lane layout the kernels assume. */
int hpm_unet_mfma_probe(const void* a, const void* b, float* d, void* stream);

/* out[p][co] = act(bias[co] + sum_{tap, ci} in[p + tap][ci] w[co][tap][ci]),
 * ks in {1, 3}, zero padding ks/2, relu != 0 clamps at 0; out32 (FP32)
 * and/or out16 (BF16) may be NULL. */
int hpm_unet_conv_fwd(const void* in, int ldi, int ci, const void* w, const float* bias, float* out32, void* out16,
                      int ldo, int co, int height, int width, int ks, int relu, void* stream);

/* The 3x3 convolution of hpm_unet_conv_fwd on an input whose channels
 * [0, cup) are the nearest 2x upsampling of up ((height+1)/2 x (width+1)/2,
 * stride ldu) and whose channels [cup, ci) are in[.. + cup ..]: the
 * decoder's concatenation without its upsampled copy. ci >= 32, ci and cup
 * multiples of 32. */
int hpm_unet_conv_fwd_up(const void* in, int ldi, int ci, const void* up, int ldu, int cup, const void* w,
                         const float* bias, float* out32, void* out16, int ldo, int co, int height, int width,
                         int relu, void* stream);

/* dw[co][tap][ci] += sum_p dy[co][p] x[p + tap][ci]: the weight gradient,
 * dy channel-major ([co][height*width] BF16), x NHWC with stride ldx, dw FP32
 * (accumulated with atomics: the caller zeroes it). ci, co multiples of 32. */
int hpm_unet_conv_wgrad(const void* dy_cm, int co, const void* x, int ldx, int ci, float* dw, int height, int width,
                        int ks, void* stream);

/* 2x2 average pooling to ((h+1)/2, (w+1)/2); nearest 2x upsampling back to
 * (h, w); FP32 -> BF16. */
int hpm_unet_avgpool2(const void* in, int ldi, void* out, int ldo, int c, int height, int width, void* stream);
int hpm_unet_upsample2(const void* in, int ldi, void* out, int ldo, int c, int height, int width, void* stream);
int hpm_unet_to_bf16(const float* in, void* out, int64_t n, void* stream);

/* ------------------------------------------------------ the learned prior
 * Weight file of hpm-mvs_amd/unet_prior.py (train_unet_prior.py writes it);
 * in files marked HPMUNET2 the head's channel 4 is the logit of the
 * proposal's confidence, and proposals below 0.5 are not kept.
 * hpm_lprior_load returns NULL if the file is missing or of another
 * network. hpm_lprior_apply runs on a handle whose last hpm_run_patchmatch
 * was a plain pass: the network proposes a depth and normal per pixel from
 * the pass's maps and the reference image, the judge keeps a proposal where
 * its multi-view cost (top-k NCC mean) is no worse than the plain
 * hypothesis's, and the kept proposals become the handle's planar prior
 * (as hpm_upload_prior would leave it; the planar prior is enabled).
 * accepted receives the fraction of pixels kept. A loaded prior may be
 * applied from several threads to different handles. */
typedef struct hpm_lprior hpm_lprior;
typedef struct hpm_handle hpm_handle;
hpm_lprior* hpm_lprior_load(const char* path, int device);
void hpm_lprior_destroy(hpm_lprior* lp);
int hpm_lprior_apply(hpm_lprior* lp, hpm_handle* h, float* accepted);
/* hpm_lprior_apply, also timing its parts on the handle's stream with
 * events of its own: ms[0] the input features, ms[1] the network, ms[2] the
 * judge (device milliseconds; ms may be NULL). */
int hpm_lprior_apply_timed(hpm_lprior* lp, hpm_handle* h, float* accepted, float* ms);
/* The network alone: head (FP32, height*width*32; channel 0 the depth
 * residual, 1..3 the normal residual) of the features x (BF16,
 * height*width*32, as hpm-mvs_amd/unet_prior.py builds them); synchronous,
 * returns a hipError_t value. */
int hpm_lprior_forward(hpm_lprior* lp, const void* x, int height, int width, float* head, void* stream);
/* hpm_lprior_forward without any allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The activations live in one fixed array per GPU of
 * hpm_lprior_noalloc_capacity() bytes (64 MiB, about 552 bytes per pixel:
 * up to about 120k pixels); a larger input returns HPM_ETOOBIG (-5) and a
 * null pointer or empty size HPM_EINVAL (-1), else a hipError_t value.
 * Asynchronous: the kernels are queued on `stream` and head is complete
 * when the stream reaches them. Every call on one GPU shares that GPU's
 * array (nothing in it needs to start at zero), so calls that overlap on
 * one GPU (other streams, other threads) race. */
int hpm_lprior_forward_noalloc(hpm_lprior* lp, const void* x, int height, int width, float* head, void* stream);
size_t hpm_lprior_noalloc_capacity(void);

#ifdef __cplusplus
}
#endif

#endif /* HPM_UNET_H_ */
