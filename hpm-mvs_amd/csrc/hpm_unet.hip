/*
 * hpm_unet.hip — BF16 MFMA kernels of the learned depth prior (a small
 * U-Net): 3x3 / 1x1 convolution forward as an implicit GEMM on the gfx950
 * matrix cores (v_mfma_f32_32x32x16_bf16, FP32 accumulation), its weight
 * gradient, and the memory-bound glue (2x2 average pooling, nearest 2x
 * upsampling, packing of the network input, BF16 conversion).
 *
 * No counterpart in the reference, whose prior is the planar/KNN
 * construction only (HPM.cpp:759-1179); the learned prior is an addition
 * of this build, gated per pixel by the NCC judge and with the planar prior
 * as fallback.
 *
 * Layout: activations NHWC per image (pixel-major, channels contiguous) with
 * a channel stride (ld) so concatenations are slices of one buffer; weights
 * [Co][k*k][Ci] BF16. Channel counts: Ci a multiple of 16, Co of 32.
 *
 * MFMA lane layout (32x32x16, 64 lanes): lane l supplies A[row l%32][k
 * 8(l/32) .. +7] and B[k 8(l/32) .. +7][col l%32] and receives
 * D[8j + 4(l/32) + i][l%32] in accumulator 4j + i (j, i < 4).
 */
#include <hip/hip_runtime.h>

#include <stdint.h>

typedef __bf16 hpm_bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace {

__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 ld8(const hpm_bf16* p)
{
    return *reinterpret_cast<const bf16x8*>(p);
}

} // namespace

/* MFMA lane-layout probe: one 32x32x16 product of the given operands,
 * a and b as each lane supplies them, d as each lane receives it. */
__global__ __launch_bounds__(64) void hpm_k_mfma_probe(const bf16x8* a, const bf16x8* b, f32x16* d)
{
    const int l = threadIdx.x;
    f32x16 acc = {};
    d[l] = mfma16(a[l], b[l], acc);
}

/* out[p][co] = act(bias[co] + sum_{tap, ci} in[p + tap][ci] w[co][tap][ci]),
 * zero padding k/2. One wave: MT tiles of 32 consecutive pixels of a row x
 * NT tiles of 32 output channels, each input fragment feeding NT products
 * and each weight fragment MT (every output takes the same sequence of
 * products, whatever the tiling); a block: 4 waves on consecutive (pixel
 * group, channel group) items, those on one pixel group sharing its input
 * loads. */
template <int MT, int NT>
__global__ __launch_bounds__(256) void hpm_k_conv_fwd(const hpm_bf16* __restrict__ in, int ldi, int ci,
                                                      const hpm_bf16* __restrict__ w, const float* __restrict__ bias,
                                                      float* out32, hpm_bf16* out16, int ldo, int co, int height,
                                                      int width, int ks, int relu)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int groups_n = co / (32 * NT);
    const int g = (int)blockIdx.x * 4 + wave;
    const int x0 = (g / groups_n) * 32 * MT;
    const int y = blockIdx.y;
    const int n0 = (g % groups_n) * 32 * NT;
    if (x0 >= width) {
        return;
    }
    const int r = lane & 31;
    const int kh = lane >> 5;
    const int taps = ks * ks;
    const int half = ks / 2;
    f32x16 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            acc[m][n] = f32x16{};
        }
    }
    const bf16x8 zero = {};
    for (int tap = 0; tap < taps; ++tap) {
        const int sy = y + tap / ks - half;
        const hpm_bf16* arow[MT];
        bool valid[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int px = x0 + 32 * m + r;
            const int sx = px + tap % ks - half;
            valid[m] = px < width && sy >= 0 && sy < height && sx >= 0 && sx < width;
            arow[m] = in + ((int64_t)(valid[m] ? sy : 0) * width + (valid[m] ? sx : 0)) * ldi + kh * 8;
        }
        const hpm_bf16* brow = w + ((int64_t)(n0 + r) * taps + tap) * ci + kh * 8;
        const int64_t bstride = (int64_t)32 * taps * ci; /* 32 output channels */
        for (int c0 = 0; c0 < ci; c0 += 16) {
            bf16x8 a[MT], b[NT];
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                a[m] = valid[m] ? ld8(arow[m] + c0) : zero;
            }
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                b[n] = ld8(brow + n * bstride + c0);
            }
#pragma unroll
            for (int m = 0; m < MT; ++m) {
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    acc[m][n] = mfma16(a[m], b[n], acc[m][n]);
                }
            }
        }
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int ch = n0 + 32 * n + r;
        const float bn = bias ? bias[ch] : 0.0f;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int x = x0 + 32 * m + 8 * j + 4 * kh + i;
                    if (x < width) {
                        float v = acc[m][n][4 * j + i] + bn;
                        if (relu) {
                            v = fmaxf(v, 0.0f);
                        }
                        const int64_t o = ((int64_t)y * width + x) * ldo + ch;
                        if (out32) {
                            out32[o] = v;
                        }
                        if (out16) {
                            out16[o] = (hpm_bf16)v;
                        }
                    }
                }
            }
        }
    }
}

/* The 3x3 convolution with its operands staged in LDS: a block computes a
 * tile of 4 output rows x 32 pixels (one row per wave) x 32 NT output
 * channels (channel group blockIdx.z); for every chunk of 32 input channels
 * the 6 input rows with their 1-pixel halo (6 x 34 x 32, zero outside the
 * image) and the weight slice (32 NT x 9 taps x 32) sit in LDS (rows padded
 * to 40 channels against bank conflicts) while every wave runs the chunk's
 * 9 taps x 2 MFMA steps per channel tile. The square-ish tile reads 1.6
 * input pixels per output (a row of 128 reads 3.05) and wastes at most 31
 * columns per map row (a row of 128: up to 127, 28% of a 200-pixel map).
 * The next chunk's operands are loaded into registers while the current one
 * is multiplied, then stored to LDS:
 * the global loads overlap the matrix work. Products are accumulated chunk
 * by chunk, tap by tap. With up != NULL, input channels [0, cup) are the
 * nearest 2x upsampling of up ((h+1)/2 x (w+1)/2, stride ldu), read at
 * (y/2, x/2) in place of an upsampled copy; cup is a multiple of 32. */
constexpr int kLdsPad = 40;
template <int NT>
__global__ __launch_bounds__(256) void hpm_k_conv3_lds(const hpm_bf16* __restrict__ in, int ldi, int ci,
                                                       const hpm_bf16* __restrict__ up, int ldu, int cup,
                                                       const hpm_bf16* __restrict__ w, const float* __restrict__ bias,
                                                       float* out32, hpm_bf16* out16, int ldo, int co, int height,
                                                       int width, int relu)
{
    constexpr int kTileRows = 4, kCols = 34, kRows = kTileRows + 2;
    constexpr int kInVec = kRows * kCols * 4; /* 16-byte vectors of the input tile */
    constexpr int kWVec = 32 * NT * 9 * 4;   /* and of the weight slice */
    constexpr int kIn = (kInVec + 255) / 256;
    constexpr int kW = (kWVec + 255) / 256;
    __shared__ __align__(16) hpm_bf16 s_in[kRows * kCols * kLdsPad];
    __shared__ __align__(16) hpm_bf16 s_w[32 * NT * 9 * kLdsPad];
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int x_blk = (int)blockIdx.x * 32;
    const int n0 = (int)blockIdx.z * 32 * NT;
    const int y_blk = (int)blockIdx.y * kTileRows;
    const int y = y_blk + wave;
    const int r = lane & 31;
    const int kh = lane >> 5;
    f32x16 acc[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        acc[n] = f32x16{};
    }
    const bf16x8 zero = {};
    bf16x8 rin[kIn], rw[kW];
    /* this thread's vectors of the chunk at input channel c0 */
    const int wu = (width + 1) / 2;
    auto fetch = [&](int c0) {
        const bool from_up = c0 < cup;
#pragma unroll
        for (int k = 0; k < kIn; ++k) {
            const int e = tid + 256 * k;
            const int v = e & 3, pix = e >> 2;
            const int row = pix / kCols, col = pix % kCols;
            const int sy = y_blk + row - 1, sx = x_blk + col - 1;
            rin[k] = zero;
            if (e < kInVec && sy >= 0 && sy < height && sx >= 0 && sx < width) {
                rin[k] = from_up ? ld8(up + ((int64_t)(sy >> 1) * wu + (sx >> 1)) * ldu + c0 + 8 * v)
                                 : ld8(in + ((int64_t)sy * width + sx) * ldi + c0 + 8 * v);
            }
        }
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            const int e = tid + 256 * k;
            const int v = e & 3, rt = e >> 2;
            const int nl = rt / 9, tap = rt % 9;
            rw[k] = zero;
            if (e < kWVec) {
                rw[k] = ld8(w + ((int64_t)(n0 + nl) * 9 + tap) * ci + c0 + 8 * v);
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int k = 0; k < kIn; ++k) {
            const int e = tid + 256 * k;
            if (e < kInVec) {
                const int v = e & 3, pix = e >> 2;
                *reinterpret_cast<bf16x8*>(&s_in[pix * kLdsPad + 8 * v]) = rin[k];
            }
        }
#pragma unroll
        for (int k = 0; k < kW; ++k) {
            const int e = tid + 256 * k;
            if (e < kWVec) {
                const int v = e & 3, rt = e >> 2;
                *reinterpret_cast<bf16x8*>(&s_w[rt * kLdsPad + 8 * v]) = rw[k];
            }
        }
    };
    fetch(0);
    stash();
    __syncthreads();
    for (int c0 = 0; c0 < ci; c0 += 32) {
        const bool more = c0 + 32 < ci;
        if (more) {
            fetch(c0 + 32);
        }
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int ky = tap / 3, kx = tap % 3;
#pragma unroll
            for (int k16 = 0; k16 < 2; ++k16) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(
                    &s_in[((ky + wave) * kCols + r + kx) * kLdsPad + 16 * k16 + 8 * kh]);
#pragma unroll
                for (int n = 0; n < NT; ++n) {
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(
                        &s_w[((32 * n + r) * 9 + tap) * kLdsPad + 16 * k16 + 8 * kh]);
                    acc[n] = mfma16(a, b, acc[n]);
                }
            }
        }
        __syncthreads();
        if (more) {
            stash();
            __syncthreads();
        }
    }
    const int x0 = x_blk;
    if (y >= height) {
        return;
    }
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const int ch = n0 + 32 * n + r;
        const float bn = bias ? bias[ch] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int x = x0 + 8 * j + 4 * kh + i;
                if (x < width) {
                    float v = acc[n][4 * j + i] + bn;
                    if (relu) {
                        v = fmaxf(v, 0.0f);
                    }
                    const int64_t o = ((int64_t)y * width + x) * ldo + ch;
                    if (out32) {
                        out32[o] = v;
                    }
                    if (out16) {
                        out16[o] = (hpm_bf16)v;
                    }
                }
            }
        }
    }
}

/* dw[co][tap][ci] += sum_p dy[co][p] x[p + tap][ci] over one chunk of
 * pixels, FP32 atomics into dw. dy is channel-major ([co][H*W], BF16), x
 * pixel-major with stride ldx. One wave: a 32x32 (co, ci) tile of one tap;
 * A = dy rows (8 consecutive pixels per lane), B = x gathered per pixel. */
__global__ __launch_bounds__(256) void hpm_k_conv_wgrad(const hpm_bf16* __restrict__ dy_cm, int co,
                                                        const hpm_bf16* __restrict__ x, int ldx, int ci,
                                                        float* dw, int height, int width, int ks, int chunk)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ci_tiles = ci / 32;
    const int tile = blockIdx.x * 4 + wave;
    const int taps = ks * ks;
    if (tile >= (co / 32) * ci_tiles * taps) {
        return;
    }
    const int tap = tile % taps;
    const int cit = (tile / taps) % ci_tiles;
    const int cot = tile / (taps * ci_tiles);
    const int r = lane & 31;
    const int kh = lane >> 5;
    const int ty = tap / ks - ks / 2, tx = tap % ks - ks / 2;
    const int64_t npx = (int64_t)height * width;
    const int64_t p_begin = (int64_t)blockIdx.y * chunk;
    const int64_t p_end = p_begin + chunk < npx ? p_begin + chunk : npx;
    const hpm_bf16* arow = dy_cm + (int64_t)(cot * 32 + r) * npx;
    const int cidx = cit * 32 + r;
    f32x16 acc = {};
    for (int64_t p0 = p_begin; p0 < p_end; p0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int64_t p = p0 + kh * 8 + e;
            hpm_bf16 av = (hpm_bf16)0.0f, bv = (hpm_bf16)0.0f;
            if (p < p_end) {
                av = arow[p];
                const int py = (int)(p / width), pxx = (int)(p % width);
                const int sy = py + ty, sx = pxx + tx;
                if (sy >= 0 && sy < height && sx >= 0 && sx < width) {
                    bv = x[((int64_t)sy * width + sx) * ldx + cidx];
                }
            }
            a[e] = av;
            b[e] = bv;
        }
        acc = mfma16(a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 8 * j + 4 * kh + i; /* co row */
            atomicAdd(&dw[((int64_t)(cot * 32 + m) * taps + tap) * ci + cidx], acc[4 * j + i]);
        }
    }
}

/* 2x2 average pooling (BF16 NHWC with strides), odd edges take what exists */
__global__ __launch_bounds__(256) void hpm_k_avgpool2(const hpm_bf16* __restrict__ in, int ldi, hpm_bf16* out,
                                                      int ldo, int c, int height, int width)
{
    const int oh = (height + 1) / 2, ow = (width + 1) / 2;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)oh * ow * c) {
        return;
    }
    const int ch = (int)(i % c);
    const int64_t p = i / c;
    const int oy = (int)(p / ow), ox = (int)(p % ow);
    float s = 0.0f;
    int n = 0;
    for (int dy = 0; dy < 2; ++dy) {
        for (int dx = 0; dx < 2; ++dx) {
            const int y = 2 * oy + dy, x = 2 * ox + dx;
            if (y < height && x < width) {
                s += (float)in[((int64_t)y * width + x) * ldi + ch];
                ++n;
            }
        }
    }
    out[p * ldo + ch] = (hpm_bf16)(s / (float)n);
}

/* nearest 2x upsampling of a (h+1)/2 x (w+1)/2 map to h x w */
__global__ __launch_bounds__(256) void hpm_k_upsample2(const hpm_bf16* __restrict__ in, int ldi, hpm_bf16* out,
                                                       int ldo, int c, int height, int width)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)height * width * c) {
        return;
    }
    const int ch = (int)(i % c);
    const int64_t p = i / c;
    const int y = (int)(p / width), x = (int)(p % width);
    const int iw = (width + 1) / 2;
    out[p * ldo + ch] = in[((int64_t)(y / 2) * iw + x / 2) * ldi + ch];
}

/* BF16 <- FP32 */
__global__ __launch_bounds__(256) void hpm_k_to_bf16(const float* __restrict__ in, hpm_bf16* out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = (hpm_bf16)in[i];
    }
}

/* --------------------------------------------------------------- launchers */

extern "C" int hpm_unet_mfma_probe(const void* a, const void* b, float* d, void* stream)
{
    hipLaunchKernelGGL(hpm_k_mfma_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, (const bf16x8*)a,
                       (const bf16x8*)b, (f32x16*)d);
    return (int)hipGetLastError();
}

namespace {

bool lds_eligible(int ci, int ks)
{
    return ks == 3 && ci >= 32 && ci % 32 == 0;
}

void launch_conv3_lds(const void* in, int ldi, int ci, const void* up, int ldu, int cup, const void* w,
                      const float* bias, float* out32, void* out16, int ldo, int co, int height, int width, int relu,
                      hipStream_t stream)
{
    /* 64 channels per block where co divides by 64 and the map gives four
     * blocks per CU (LDS for two blocks per CU), else 32: the channel groups
     * spread small maps over more blocks */
    const int64_t tiles = (int64_t)((width + 31) / 32) * ((height + 3) / 4);
    const int nt = co % 64 == 0 && tiles * (co / 64) >= 1024 ? 2 : 1;
    const dim3 grid((unsigned)((width + 31) / 32), (unsigned)((height + 3) / 4), (unsigned)(co / (32 * nt)));
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, stream, (const hpm_bf16*)in, ldi, ci, (const hpm_bf16*)up, ldu,
                           cup, (const hpm_bf16*)w, bias, out32, (hpm_bf16*)out16, ldo, co, height, width, relu);
    };
    if (nt == 1) {
        go(hpm_k_conv3_lds<1>);
    } else {
        go(hpm_k_conv3_lds<2>);
    }
}

} // namespace

extern "C" int hpm_unet_conv_fwd_up(const void* in, int ldi, int ci, const void* up, int ldu, int cup, const void* w,
                                    const float* bias, float* out32, void* out16, int ldo, int co, int height,
                                    int width, int relu, void* stream)
{
    if (!lds_eligible(ci, 3) || co % 32 != 0 || cup < 0 || cup % 32 != 0 || cup > ci || ldi % 8 != 0 ||
        ldu % 8 != 0 || height < 1 || width < 1 || (cup > 0 && !up)) {
        return (int)hipErrorInvalidValue;
    }
    launch_conv3_lds(in, ldi, ci, up, ldu, cup, w, bias, out32, out16, ldo, co, height, width, relu,
                     (hipStream_t)stream);
    return (int)hipGetLastError();
}

extern "C" int hpm_unet_conv_fwd(const void* in, int ldi, int ci, const void* w, const float* bias, float* out32,
                                 void* out16, int ldo, int co, int height, int width, int ks, int relu, void* stream)
{
    if (ci < 16 || ci % 16 != 0 || co < 32 || co % 32 != 0 || (ks != 1 && ks != 3) || ldi % 8 != 0 || height < 1 ||
        width < 1) {
        return (int)hipErrorInvalidValue;
    }
    /* tiles per wave: two pixel tiles at 32 output channels, else two or
     * three channel tiles */
    auto launch = [&](auto kernel, int mt, int nt) {
        const int items = ((width + 32 * mt - 1) / (32 * mt)) * (co / (32 * nt));
        const dim3 grid((unsigned)((items + 3) / 4), (unsigned)height, 1);
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, (hipStream_t)stream, (const hpm_bf16*)in, ldi, ci,
                           (const hpm_bf16*)w, bias, out32, (hpm_bf16*)out16, ldo, co, height, width, ks, relu);
    };
    if (lds_eligible(ci, ks)) {
        /* operands staged in LDS with the next chunk prefetched; also ahead
         * of the tiled kernel for a single 32-channel chunk (the staged
         * input tile is read by 9 taps x the block's channel tiles) */
        launch_conv3_lds(in, ldi, ci, nullptr, 0, 0, w, bias, out32, out16, ldo, co, height, width, relu,
                         (hipStream_t)stream);
    } else if (co == 32) {
        launch(hpm_k_conv_fwd<2, 1>, 2, 1);
    } else if (co % 64 == 0) {
        launch(hpm_k_conv_fwd<1, 2>, 1, 2);
    } else if (co % 96 == 0) {
        launch(hpm_k_conv_fwd<1, 3>, 1, 3);
    } else {
        launch(hpm_k_conv_fwd<1, 1>, 1, 1);
    }
    return (int)hipGetLastError();
}

extern "C" int hpm_unet_conv_wgrad(const void* dy_cm, int co, const void* x, int ldx, int ci, float* dw, int height,
                                   int width, int ks, void* stream)
{
    if (ci % 32 != 0 || co % 32 != 0 || (ks != 1 && ks != 3)) {
        return (int)hipErrorInvalidValue;
    }
    const int tiles = (co / 32) * (ci / 32) * ks * ks;
    const int64_t npx = (int64_t)height * width;
    const int chunk = 4096;
    const dim3 grid((tiles + 3) / 4, (unsigned)((npx + chunk - 1) / chunk), 1);
    hipLaunchKernelGGL(hpm_k_conv_wgrad, grid, dim3(256), 0, (hipStream_t)stream, (const hpm_bf16*)dy_cm, co,
                       (const hpm_bf16*)x, ldx, ci, dw, height, width, ks, chunk);
    return (int)hipGetLastError();
}

extern "C" int hpm_unet_avgpool2(const void* in, int ldi, void* out, int ldo, int c, int height, int width,
                                 void* stream)
{
    const int64_t n = (int64_t)((height + 1) / 2) * ((width + 1) / 2) * c;
    hipLaunchKernelGGL(hpm_k_avgpool2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const hpm_bf16*)in, ldi, (hpm_bf16*)out, ldo, c, height, width);
    return (int)hipGetLastError();
}

extern "C" int hpm_unet_upsample2(const void* in, int ldi, void* out, int ldo, int c, int height, int width,
                                  void* stream)
{
    const int64_t n = (int64_t)height * width * c;
    hipLaunchKernelGGL(hpm_k_upsample2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const hpm_bf16*)in, ldi, (hpm_bf16*)out, ldo, c, height, width);
    return (int)hipGetLastError();
}

extern "C" int hpm_unet_to_bf16(const float* in, void* out, int64_t n, void* stream)
{
    hipLaunchKernelGGL(hpm_k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, in,
                       (hpm_bf16*)out, n);
    return (int)hipGetLastError();
}
