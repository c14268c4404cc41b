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

#ifdef __cplusplus
}
#endif

#endif /* HPM_UNET_H_ */
