// Learned closure: kernels of a small 3x3-convolution network (no
// counterpart in the reference).
//
// Activations are NHWC with the channel stride rounded up to 8; weights and
// activations are BF16, accumulation is FP32 on the matrix cores
// (v_mfma_f32_32x32x8_bf16).  Every convolution is an implicit GEMM: the
// im2col operand is gathered on the fly from the NHWC tensor, never stored.
//   forward       Z[p, co]  = b[co] + sum_{tap, ci} X[p + off(tap), ci] W[tap, ci, co]
//   data grad     dX[p, ci] = sum_{tap, co} dZ[p - off(tap), co] W[tap, ci, co]
//   weight grad   dW[tap, ci, co] = sum_p X[p + off(tap), ci] dZ[p, co]
// One wave computes a 32x32 tile of the GEMM.  MFMA operand layout
// (32x32x8): lane l holds A row l%32 and B column l%32, k = 4*(l/32)+0..3;
// accumulator register r of lane l is row 8*(r/4) + 4*(l/32) + r%4, column
// l%32.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <string>

#include "nemo2d.h"
#include "nemo2d_core.h"
#include "nemo2d_internal.h"

using namespace nemo2d;

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NB = 256;  // threads per block (4 waves)

__device__ __forceinline__ uint16_t f2bf(float f) { return bf16_rne(f); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ bf16x4 load4(const uint16_t* p) {
    // 4 consecutive bf16 (8-byte aligned)
    const uint2 v = *(const uint2*)p;
    bf16x4 r;
    r[0] = (short)(v.x & 0xffff);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffff);
    r[3] = (short)(v.y >> 16);
    return r;
}

struct Img {
    int n, h, w;
    __device__ long pixels() const { return (long)n * h * w; }
};

// Pixel p shifted by (dy, dx) within its image, or -1 outside.
__device__ __forceinline__ long shifted(const Img& im, long p, int dy, int dx) {
    const int x = (int)(p % im.w);
    const long t = p / im.w;
    const int y = (int)(t % im.h);
    const int yy = y + dy, xx = x + dx;
    if (yy < 0 || yy >= im.h || xx < 0 || xx >= im.w) return -1;
    return p + (long)dy * im.w + dx;
}

// ---- forward and data-gradient convolutions -------------------------------
// A block owns a tile of CT_H rows x CT_W = 32 pixels of one image.  The
// input tile with its 1-pixel halo is staged once in LDS (pixel-major, the
// channels of a pixel contiguous, padded to the MFMA K step); each of the 4
// waves computes CT_H/4 row segments, one 32-pixel x 32-channel MFMA tile
// (v_mfma_f32_32x32x16_bf16) each, over the 9 taps.  The weights of every
// tap stay in registers.  Operand layout (32x32x16): lane l holds A row l%32
// and B column l%32 at 8 consecutive k of half l/32; the K order is the same
// for both operands, so any assignment of k to (channel) values works.
typedef short bf16x8 __attribute__((ext_vector_type(8)));
constexpr int CT_W = 32, CT_H = 8;
constexpr int CT_PIX = (CT_H + 2) * (CT_W + 2);

__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) {
    const uint4 v = *(const uint4*)p;
    bf16x8 r;
    r[0] = (short)(v.x & 0xffff); r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xffff); r[3] = (short)(v.y >> 16);
    r[4] = (short)(v.z & 0xffff); r[5] = (short)(v.z >> 16);
    r[6] = (short)(v.w & 0xffff); r[7] = (short)(v.w >> 16);
    return r;
}

// Stage src[n][y0-1 .. y0+CT_H][x0-1 .. x0+CT_W][0..cp) into LDS tile[pixel][KT + 8]
// (channels cp..KT-1 and pixels outside the image zero).
template <int KT>
__device__ __forceinline__ void stage_tile(const uint16_t* __restrict__ src, const Img& im, int cp,
                                           int img, int y0, int x0, uint16_t* tile) {
    constexpr int PS = KT + 8;  // LDS pixel stride (bf16): +16 B against bank conflicts
    constexpr int CH = KT / 8;  // 16-byte chunks per pixel
    for (int e = threadIdx.x; e < CT_PIX * CH; e += NB) {
        const int q = e / CH, c8 = (e % CH) * 8;
        const int yy = y0 - 1 + q / (CT_W + 2), xx = x0 - 1 + q % (CT_W + 2);
        uint4 v = make_uint4(0, 0, 0, 0);
        if (c8 < cp && yy >= 0 && yy < im.h && xx >= 0 && xx < im.w)
            v = *(const uint4*)(src + (((long)img * im.h + yy) * im.w + xx) * cp + c8);
        *(uint4*)(tile + q * PS + c8) = v;
    }
}

// Forward: z = b + sum_tap sum_ci X[p + off(tap)][ci] W[tap][ci][co];
// act = gelu(z) (or z).  wt: [9][32][cinp] (cout-major, from pack).
template <int KT>
__global__ void __launch_bounds__(NB)
k_conv_fwd(const uint16_t* __restrict__ x, Img im, int cinp, const uint16_t* __restrict__ wt,
           const float* __restrict__ bias, int cout, int coutp, float* __restrict__ z,
           uint16_t* __restrict__ zb, uint16_t* __restrict__ act, int use_gelu) {
    constexpr int PS = KT + 8;
    __shared__ __attribute__((aligned(16))) uint16_t tile[CT_PIX * PS];
    const int tiles_x = (im.w + CT_W - 1) / CT_W, tiles_y = (im.h + CT_H - 1) / CT_H;
    const int b = blockIdx.x;
    const int img = b / (tiles_x * tiles_y);
    const int y0 = (b / tiles_x) % tiles_y * CT_H, x0 = b % tiles_x * CT_W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5, c = lane & 31;
    // weights of this lane's column (output channel c), every tap and K step
    bf16x8 wreg[9][KT / 16];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int kk = 0; kk < KT / 16; ++kk) {
            const int ci = kk * 16 + 8 * half;
            bf16x8 w = {0, 0, 0, 0, 0, 0, 0, 0};
            if (ci < cinp) w = lds8(wt + ((long)tap * 32 + c) * cinp + ci);
            wreg[tap][kk] = w;
        }
    stage_tile<KT>(x, im, cinp, img, y0, x0, tile);
    __syncthreads();
    const float bc = c < cout ? bias[c] : 0.f;
    for (int r = wave; r < CT_H; r += NB / 64) {
        const int y = y0 + r;
        if (y >= im.h) break;
        f32x16 acc = {0};
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3 - 1, dx = tap % 3 - 1;
            const uint16_t* a = tile + ((r + 1 + dy) * (CT_W + 2) + c + 1 + dx) * PS + 8 * half;
#pragma unroll
            for (int kk = 0; kk < KT / 16; ++kk)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds8(a + kk * 16), wreg[tap][kk], acc, 0, 0, 0);
        }
        // accumulator register k: pixel 8*(k/4) + 4*half + k%4 of the row
        // segment, output channel c
        if (c >= coutp) continue;
        const long row = ((long)img * im.h + y) * im.w;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int xx = x0 + 8 * (k >> 2) + 4 * half + (k & 3);
            if (xx >= im.w) continue;
            const long q = (row + xx) * coutp + c;
            if (c < cout) {
                const float v = acc[k] + bc;
                if (z) z[q] = v;
                if (zb) zb[q] = f2bf(v);
                if (act) act[q] = f2bf(use_gelu ? gelu(v) : v);
            } else {  // padded channels: zero for the next layer
                if (zb) zb[q] = 0;
                if (act) act[q] = 0;
            }
        }
    }
}

// Data gradient: dX[p][ci] = sum_tap sum_co dZ[p - off(tap)][co] W[tap][ci][co];
// with zp (the layer below's pre-activation) the output is the layer below's
// dz = BF16(dX * GELU'(zp)) (GELU backward fused), else FP32 dX.
// wt: [9][32][coutp] (cin-major, from pack).
template <int KT>
__global__ void __launch_bounds__(NB)
k_conv_bwd_data(const uint16_t* __restrict__ dz, Img im, int coutp, const uint16_t* __restrict__ wt,
                int cin, int cinp, float* __restrict__ dx, const float* __restrict__ zp,
                const uint16_t* __restrict__ zpb, uint16_t* __restrict__ dzp) {
    constexpr int PS = KT + 8;
    __shared__ __attribute__((aligned(16))) uint16_t tile[CT_PIX * PS];
    const int tiles_x = (im.w + CT_W - 1) / CT_W, tiles_y = (im.h + CT_H - 1) / CT_H;
    const int b = blockIdx.x;
    const int img = b / (tiles_x * tiles_y);
    const int y0 = (b / tiles_x) % tiles_y * CT_H, x0 = b % tiles_x * CT_W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5, c = lane & 31;
    bf16x8 wreg[9][KT / 16];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap)
#pragma unroll
        for (int kk = 0; kk < KT / 16; ++kk) {
            const int co = kk * 16 + 8 * half;
            bf16x8 w = {0, 0, 0, 0, 0, 0, 0, 0};
            if (co < coutp) w = lds8(wt + ((long)tap * 32 + c) * coutp + co);
            wreg[tap][kk] = w;
        }
    stage_tile<KT>(dz, im, coutp, img, y0, x0, tile);
    __syncthreads();
    const bool gl = zp || zpb;
    for (int r = wave; r < CT_H; r += NB / 64) {
        const int y = y0 + r;
        if (y >= im.h) break;
        const long row = ((long)img * im.h + y) * im.w;
        // the layer below's pre-activations of this lane's 16 outputs, loaded
        // before the MFMAs so that their latency hides behind them
        float zv[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int xx = x0 + 8 * (k >> 2) + 4 * half + (k & 3);
            const long q = (row + xx) * cinp + c;
            zv[k] = 0.f;
            if (gl && c < cin && xx < im.w) zv[k] = zp ? zp[q] : bf2f(zpb[q]);
        }
        f32x16 acc = {0};
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3 - 1, dxo = tap % 3 - 1;
            const uint16_t* a = tile + ((r + 1 - dy) * (CT_W + 2) + c + 1 - dxo) * PS + 8 * half;
#pragma unroll
            for (int kk = 0; kk < KT / 16; ++kk)
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lds8(a + kk * 16), wreg[tap][kk], acc, 0, 0, 0);
        }
        if (c >= cinp) continue;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int xx = x0 + 8 * (k >> 2) + 4 * half + (k & 3);
            if (xx >= im.w) continue;
            const long q = (row + xx) * cinp + c;
            if (gl) {
                dzp[q] = c < cin ? f2bf(acc[k] * gelu_grad(zv[k])) : (uint16_t)0;
            } else if (c < cin) {
                dx[q] = acc[k];
            }
        }
    }
}

// ---- weight gradient ---------------------------------------------------------
// dW[tap][ci][co] = sum_p X[p + off(tap)][ci] dZ[p][co] and db[co] = sum_p dZ[p][co].
// A block of 9 waves (wave = tap) walks tiles of WG_H rows x WG_W pixels of
// 4 images at once.  Each tile is staged in LDS transposed to
// [channel][y][x][image], so that a lane's 8 K values (2 neighbouring pixels
// x 4 images) of one channel are 16 contiguous bytes for every tap shift;
// the next tile's global loads are issued before the current tile's MFMAs.
// The centre-tap wave also multiplies dZ by a ones operand: every row of
// that product is the bias gradient.  Each block writes its partial sums
// (9 x 32 x 32 + 32 floats); k_wgrad_sum adds the blocks in fixed order.
constexpr int WG_H = 2, WG_W = 32, WG_NT = 9 * 64;
constexpr int WG_XP = (WG_H + 2) * (WG_W + 2);        // halo tile pixels
constexpr int WG_CSX = 290;                           // X channel stride, dwords (= 2 mod 32)
constexpr int WG_CSZ = 132;                           // dZ channel stride, dwords (= 4 mod 32)
constexpr int WG_PART = 9 * 1024 + 32;                // floats per block partial
constexpr int WG_MAXBLK = 512;
static_assert(2 * WG_XP <= WG_CSX && 2 * WG_H * WG_W <= WG_CSZ, "LDS channel planes");

struct WgItems {
    uint2 xv[2][4];  // up to 2 X items per thread, 4 images each
    uint2 zv[4];     // 1 dZ item, 4 images
};

__device__ __forceinline__ uint2 ld_px(const uint16_t* __restrict__ src, const Img& im, int cp, int img,
                                      int y, int x, int c4) {
    if (img >= im.n || y < 0 || y >= im.h || x < 0 || x >= im.w || c4 * 4 >= cp) return make_uint2(0, 0);
    return *(const uint2*)(src + (((long)img * im.h + y) * im.w + x) * cp + c4 * 4);
}

__device__ __forceinline__ void wg_load(const uint16_t* __restrict__ x, int cinp, const uint16_t* __restrict__ dz,
                                        int coutp, const Img& im, int g, int y0, int x0, WgItems& it) {
    const int cq = cinp / 4, zq = coutp / 4;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int e = threadIdx.x + s * WG_NT;
        const int q = e / cq, c4 = e % cq;
        const bool ok = q < WG_XP;
        const int yy = y0 - 1 + q / (WG_W + 2), xx = x0 - 1 + q % (WG_W + 2);
#pragma unroll
        for (int n = 0; n < 4; ++n) it.xv[s][n] = ok ? ld_px(x, im, cinp, 4 * g + n, yy, xx, c4) : make_uint2(0, 0);
    }
    {
        const int e = threadIdx.x;
        const int q = e / zq, c4 = e % zq;
        const bool ok = q < WG_H * WG_W;
        const int yy = y0 + q / WG_W, xx = x0 + q % WG_W;
#pragma unroll
        for (int n = 0; n < 4; ++n) it.zv[n] = ok ? ld_px(dz, im, coutp, 4 * g + n, yy, xx, c4) : make_uint2(0, 0);
    }
}

// 4 images x 4 channels of one pixel -> for each channel the 4 images' bf16
__device__ __forceinline__ void put4(uint32_t* lds, int stride, const uint2* v) {
    const uint32_t c0a = (v[0].x & 0xffffu) | (v[1].x << 16), c0b = (v[2].x & 0xffffu) | (v[3].x << 16);
    const uint32_t c1a = (v[0].x >> 16) | (v[1].x & 0xffff0000u), c1b = (v[2].x >> 16) | (v[3].x & 0xffff0000u);
    const uint32_t c2a = (v[0].y & 0xffffu) | (v[1].y << 16), c2b = (v[2].y & 0xffffu) | (v[3].y << 16);
    const uint32_t c3a = (v[0].y >> 16) | (v[1].y & 0xffff0000u), c3b = (v[2].y >> 16) | (v[3].y & 0xffff0000u);
    *(uint2*)(lds) = make_uint2(c0a, c0b);
    *(uint2*)(lds + stride) = make_uint2(c1a, c1b);
    *(uint2*)(lds + 2 * stride) = make_uint2(c2a, c2b);
    *(uint2*)(lds + 3 * stride) = make_uint2(c3a, c3b);
}

__device__ __forceinline__ void wg_store(int cinp, int coutp, const WgItems& it, uint32_t* xs, uint32_t* zs) {
    const int cq = cinp / 4, zq = coutp / 4;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int e = threadIdx.x + s * WG_NT;
        const int q = e / cq, c4 = e % cq;
        if (q < WG_XP) put4(xs + 4 * c4 * WG_CSX + 2 * q, WG_CSX, it.xv[s]);
    }
    const int q = threadIdx.x / zq, c4 = threadIdx.x % zq;
    if (q < WG_H * WG_W) put4(zs + 4 * c4 * WG_CSZ + 2 * q, WG_CSZ, it.zv);
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint2 a, uint2 b) {
    bf16x8 r;
    r[0] = (short)(a.x & 0xffff); r[1] = (short)(a.x >> 16);
    r[2] = (short)(a.y & 0xffff); r[3] = (short)(a.y >> 16);
    r[4] = (short)(b.x & 0xffff); r[5] = (short)(b.x >> 16);
    r[6] = (short)(b.y & 0xffff); r[7] = (short)(b.y >> 16);
    return r;
}

__global__ void __launch_bounds__(WG_NT)
k_conv_bwd_weight(const uint16_t* __restrict__ x, int cinp, const uint16_t* __restrict__ dz,
                  int coutp, Img im, float* __restrict__ work) {
    __shared__ __attribute__((aligned(16))) uint32_t xs[32 * WG_CSX];
    __shared__ __attribute__((aligned(16))) uint32_t zs[32 * WG_CSZ];
    const int lane = threadIdx.x & 63, tap = threadIdx.x >> 6, half = lane >> 5, c = lane & 31;
    const int dy = tap / 3 - 1, dxo = tap % 3 - 1;
    const int tx = (im.w + WG_W - 1) / WG_W, ty = (im.h + WG_H - 1) / WG_H, ng = (im.n + 3) / 4;
    const long ntiles = (long)tx * ty * ng;
    f32x16 acc = {0}, accb = {0};
    bf16x8 ones;
#pragma unroll
    for (int k = 0; k < 8; ++k) ones[k] = (short)0x3f80;  // BF16 1.0
    WgItems it;
    long t = blockIdx.x;
    if (t < ntiles) {
        const int g = (int)(t / ((long)tx * ty)), r = (int)(t % ((long)tx * ty));
        wg_load(x, cinp, dz, coutp, im, g, r / tx * WG_H, r % tx * WG_W, it);
    }
    for (; t < ntiles; t += gridDim.x) {
        __syncthreads();  // the previous tile's MFMAs have read the LDS
        wg_store(cinp, coutp, it, xs, zs);
        __syncthreads();
        const long tn = t + gridDim.x;
        if (tn < ntiles) {
            const int g = (int)(tn / ((long)tx * ty)), r = (int)(tn % ((long)tx * ty));
            wg_load(x, cinp, dz, coutp, im, g, r / tx * WG_H, r % tx * WG_W, it);
        }
        // K step: pixels xq, xq+1 (xq = x4 + 2*half) of row y, images 0..3
#pragma unroll
        for (int y = 0; y < WG_H; ++y)
#pragma unroll
            for (int x4 = 0; x4 < WG_W; x4 += 4) {
                const int xq = x4 + 2 * half;
                bf16x8 a = {0, 0, 0, 0, 0, 0, 0, 0}, bz = {0, 0, 0, 0, 0, 0, 0, 0};
                if (c < cinp) {
                    const uint32_t* pa = xs + c * WG_CSX + 2 * ((y + 1 + dy) * (WG_W + 2) + xq + 1 + dxo);
                    a = as_bf16x8(*(const uint2*)pa, *(const uint2*)(pa + 2));
                }
                if (c < coutp) {
                    const uint4 v = *(const uint4*)(zs + c * WG_CSZ + 2 * (y * WG_W + xq));
                    bz = as_bf16x8(make_uint2(v.x, v.y), make_uint2(v.z, v.w));
                }
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bz, acc, 0, 0, 0);
                if (tap == 4) accb = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ones, bz, accb, 0, 0, 0);
            }
    }
    // register k of lane l: row ci = 8*(k/4) + 4*half + k%4, column co = c
    float* out = work + (long)blockIdx.x * WG_PART;
#pragma unroll
    for (int k = 0; k < 16; ++k) out[tap * 1024 + (8 * (k >> 2) + 4 * half + (k & 3)) * 32 + c] = acc[k];
    if (tap == 4 && half == 0) out[9 * 1024 + c] = accb[0];
}

// dW[tap][ci][co] and db[co] from the blocks' partials.  A block takes 32
// consecutive entries; its 8 groups of 32 threads add every 8th block partial
// (4 running sums each), then the groups' sums are added in group order: a
// fixed order, so the result does not depend on timing.
constexpr int WS_E = 32, WS_G = 8;
__global__ void __launch_bounds__(WS_E * WS_G)
k_wgrad_sum(const float* __restrict__ work, int nblk, int cin, int cout, float* __restrict__ dw,
            float* __restrict__ db) {
    __shared__ float red[WS_G][WS_E];
    const int el = threadIdx.x % WS_E, grp = threadIdx.x / WS_E;
    const int e = blockIdx.x * WS_E + el;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (e < WG_PART) {
        int b = grp;
        for (; b + 3 * WS_G < nblk; b += 4 * WS_G) {
            s0 += work[(long)b * WG_PART + e];
            s1 += work[(long)(b + WS_G) * WG_PART + e];
            s2 += work[(long)(b + 2 * WS_G) * WG_PART + e];
            s3 += work[(long)(b + 3 * WS_G) * WG_PART + e];
        }
        for (; b < nblk; b += WS_G) s0 += work[(long)b * WG_PART + e];
    }
    red[grp][el] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (grp != 0 || e >= WG_PART) return;
    float s = 0.f;
#pragma unroll
    for (int g = 0; g < WS_G; ++g) s += red[g][el];
    if (e < 9 * 1024) {
        const int tap = e / 1024, ci = (e % 1024) / 32, co = e % 32;
        if (ci < cin && co < cout) dw[((long)tap * cin + ci) * cout + co] = s;
    } else if (e - 9 * 1024 < cout) {
        db[e - 9 * 1024] = s;
    }
}

__global__ void k_gelu_bwd(const float* __restrict__ dact, const float* __restrict__ z, long count,
                           uint16_t* __restrict__ dz) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) dz[k] = f2bf(dact[k] * gelu_grad(z[k]));
}

// Weighted MSE over channels < ch of [P][cp] tensors: loss (one float per
// block into part), gradient dy (BF16).
__global__ void __launch_bounds__(NB)
k_mse(const float* __restrict__ y, const float* __restrict__ t, long P, int cp, int ch,
      const float* __restrict__ wch, float inv_count, uint16_t* __restrict__ dy,
      float* __restrict__ part) {
    __shared__ float red[NB];
    float s = 0.f;
    for (long k = (long)blockIdx.x * NB + threadIdx.x; k < P * cp; k += (long)gridDim.x * NB) {
        const int c = (int)(k % cp);
        float g = 0.f;
        if (c < ch) {
            const float d = y[k] - t[k];
            s += wch[c] * d * d;
            g = 2.f * wch[c] * d * inv_count;
        }
        dy[k] = f2bf(g);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = NB / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] * inv_count;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// W[tap][ci][co] FP32 -> forward layout [tap][co<32][cinp] and data-gradient
// layout [tap][ci<32][coutp], BF16, zero padded.
__global__ void k_pack_w(const float* __restrict__ w, int cin, int cout, int cinp, int coutp,
                         uint16_t* __restrict__ wf, uint16_t* __restrict__ wb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nf = 9 * 32 * cinp, nb = 9 * 32 * coutp;
    if (e < nf) {
        const int tap = e / (32 * cinp), co = (e / cinp) % 32, ci = e % cinp;
        wf[e] = (co < cout && ci < cin) ? f2bf(w[((long)tap * cin + ci) * cout + co]) : 0;
    }
    if (e < nb) {
        const int tap = e / (32 * coutp), ci = (e / coutp) % 32, co = e % coutp;
        wb[e] = (co < cout && ci < cin) ? f2bf(w[((long)tap * cin + ci) * cout + co]) : 0;
    }
}

__global__ void k_f2bf(const float* __restrict__ a, long n, uint16_t* __restrict__ b) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = f2bf(a[k]);
}

int launched(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return NEMO2D_OK;
}

bool bad_ch(int c) { return c < 1 || c > 32; }
int pad8(int c) { return (c + 7) / 8 * 8; }

}  // namespace

extern "C" {

int nemo2d_nn_conv_fwd(const void* x, int n, int h, int w, int cin, const void* wf,
                       const float* bias, int cout, float* z, void* zb, void* act, int use_gelu,
                       void* stream) {
    if (!x || !wf || !bias || !(z || zb || act) || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_fwd: bad argument");
    const Img im{n, h, w};
    const long nb = (long)n * ((h + CT_H - 1) / CT_H) * ((w + CT_W - 1) / CT_W);
    if (nb > 0x7fffffffL) return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_fwd: too many pixels");
    auto* k = pad8(cin) <= 16 ? k_conv_fwd<16> : k_conv_fwd<32>;
    k<<<(unsigned)nb, NB, 0, (hipStream_t)stream>>>((const uint16_t*)x, im, pad8(cin), (const uint16_t*)wf,
                                                  bias, cout, pad8(cout), z, (uint16_t*)zb, (uint16_t*)act,
                                                  use_gelu);
    return launched("conv_fwd");
}

static int bwd_data(const void* dz, int n, int h, int w, int cout, const void* wb, int cin, float* dx,
                    const float* zp, const void* zpb, void* dzp, void* stream, const char* who) {
    const Img im{n, h, w};
    const long nb = (long)n * ((h + CT_H - 1) / CT_H) * ((w + CT_W - 1) / CT_W);
    if (nb > 0x7fffffffL) return set_error(NEMO2D_ERR_ARG, std::string(who) + ": too many pixels");
    auto* k = pad8(cout) <= 16 ? k_conv_bwd_data<16> : k_conv_bwd_data<32>;
    k<<<(unsigned)nb, NB, 0, (hipStream_t)stream>>>((const uint16_t*)dz, im, pad8(cout), (const uint16_t*)wb,
                                                  cin, pad8(cin), dx, zp, (const uint16_t*)zpb,
                                                  (uint16_t*)dzp);
    return launched(who);
}

int nemo2d_nn_conv_bwd_data(const void* dz, int n, int h, int w, int cout, const void* wb, int cin,
                            float* dx, void* stream) {
    if (!dz || !wb || !dx || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_bwd_data: bad argument");
    return bwd_data(dz, n, h, w, cout, wb, cin, dx, nullptr, nullptr, nullptr, stream, "conv_bwd_data");
}

int nemo2d_nn_conv_bwd_data_gelu(const void* dz, int n, int h, int w, int cout, const void* wb, int cin,
                                 const void* zp, int zp_bf16, void* dzp, void* stream) {
    if (!dz || !wb || !zp || !dzp || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_bwd_data_gelu: bad argument");
    return bwd_data(dz, n, h, w, cout, wb, cin, nullptr, zp_bf16 ? nullptr : (const float*)zp,
                    zp_bf16 ? zp : nullptr, dzp, stream, "conv_bwd_data_gelu");
}

// Blocks of the weight gradient for a batch: one per tile up to WG_MAXBLK.
static int wgrad_blocks(long n, long h, long w) {
    const long t = (n + 3) / 4 * ((h + WG_H - 1) / WG_H) * ((w + WG_W - 1) / WG_W);
    return (int)(t < WG_MAXBLK ? t : WG_MAXBLK);
}

// work: at least nemo2d_nn_wgrad_work_floats(n*h*w) floats.
long nemo2d_nn_wgrad_work_floats(long pixels) {
    (void)pixels;
    return (long)WG_MAXBLK * WG_PART;
}

static int bwd_weight(const void* x, const void* dz, int n, int h, int w, int cin, int cout, float* dw,
                      float* db, float* work, void* stream, const char* who) {
    const Img im{n, h, w};
    const int nblk = wgrad_blocks(n, h, w);
    hipStream_t s = (hipStream_t)stream;
    k_conv_bwd_weight<<<nblk, WG_NT, 0, s>>>((const uint16_t*)x, pad8(cin), (const uint16_t*)dz,
                                             pad8(cout), im, work);
    int rc = launched(who);
    if (rc) return rc;
    k_wgrad_sum<<<(WG_PART + WS_E - 1) / WS_E, WS_E * WS_G, 0, s>>>(work, nblk, cin, cout, dw, db);
    return launched(who);
}

int nemo2d_nn_conv_bwd_weight(const void* x, const void* dz, int n, int h, int w, int cin, int cout,
                              float* dw, float* db, float* work, void* stream) {
    if (!x || !dz || !dw || !db || !work || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_bwd_weight: bad argument");
    return bwd_weight(x, dz, n, h, w, cin, cout, dw, db, work, stream, "conv_bwd_weight");
}

int nemo2d_nn_gelu_bwd(const float* dact, const float* z, long count, void* dz, void* stream) {
    if (!dact || !z || !dz || count < 0) return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_gelu_bwd: bad argument");
    if (count == 0) return NEMO2D_OK;
    k_gelu_bwd<<<(unsigned)((count + 255) / 256), 256, 0, (hipStream_t)stream>>>(dact, z, count,
                                                                                 (uint16_t*)dz);
    return launched("gelu_bwd");
}

// part: at least nblocks floats; the host sums them (nblocks returned).
int nemo2d_nn_mse(const float* y, const float* t, long pixels, int ch, const float* wch, void* dy,
                  float* part, int nblocks, void* stream) {
    if (!y || !t || !wch || !dy || !part || pixels < 1 || bad_ch(ch) || nblocks < 1)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_mse: bad argument");
    k_mse<<<nblocks, NB, 0, (hipStream_t)stream>>>(y, t, pixels, pad8(ch), ch, wch,
                                                   1.0f / (float)(pixels * ch), (uint16_t*)dy, part);
    return launched("mse");
}

int nemo2d_nn_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
                   float eps, int step, void* stream) {
    if (!p || !g || !m || !v || n < 1 || step < 1) return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_adam: bad argument");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    k_adam<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return launched("adam");
}

int nemo2d_nn_pack_weights(const float* w, int cin, int cout, void* wf, void* wb, void* stream) {
    if (!w || !wf || !wb || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_pack_weights: bad argument");
    const int n = 9 * 32 * (pad8(cin) > pad8(cout) ? pad8(cin) : pad8(cout));
    k_pack_w<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(w, cin, cout, pad8(cin), pad8(cout),
                                                               (uint16_t*)wf, (uint16_t*)wb);
    return launched("pack_weights");
}

int nemo2d_nn_to_bf16(const float* a, long n, void* b, void* stream) {
    if (!a || !b || n < 0) return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_to_bf16: bad argument");
    if (n == 0) return NEMO2D_OK;
    k_f2bf<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(a, n, (uint16_t*)b);
    return launched("to_bf16");
}

int nemo2d_dev_alloc(long bytes, void** out) {
    if (!out || bytes < 0) return set_error(NEMO2D_ERR_ARG, "nemo2d_dev_alloc: bad argument");
    *out = nullptr;
    if (bytes == 0) return NEMO2D_OK;
    const hipError_t e = hipMalloc(out, (size_t)bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return set_error(NEMO2D_ERR_HIP, std::string("nemo2d_dev_alloc: ") + hipGetErrorString(e));
    }
    return NEMO2D_OK;
}

int nemo2d_dev_free(void* p) {
    if (!p) return NEMO2D_OK;
    const hipError_t e = hipFree(p);
    if (e != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string("nemo2d_dev_free: ") + hipGetErrorString(e));
    return NEMO2D_OK;
}

static int dev_copy(void* dst, const void* src, long bytes, hipMemcpyKind k, const char* what) {
    if (bytes < 0 || (bytes > 0 && (!dst || !src))) return set_error(NEMO2D_ERR_ARG, std::string(what) + ": bad argument");
    if (bytes == 0) return NEMO2D_OK;
    const hipError_t e = hipMemcpy(dst, src, (size_t)bytes, k);
    if (e != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return NEMO2D_OK;
}

int nemo2d_dev_upload(void* dst, const void* src, long bytes) {
    return dev_copy(dst, src, bytes, hipMemcpyHostToDevice, "nemo2d_dev_upload");
}

int nemo2d_dev_download(void* dst, const void* src, long bytes) {
    return dev_copy(dst, src, bytes, hipMemcpyDeviceToHost, "nemo2d_dev_download");
}

int nemo2d_dev_zero(void* p, long bytes) {
    if (bytes < 0 || (bytes > 0 && !p)) return set_error(NEMO2D_ERR_ARG, "nemo2d_dev_zero: bad argument");
    if (bytes == 0) return NEMO2D_OK;
    hipError_t e = hipMemset(p, 0, (size_t)bytes);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string("nemo2d_dev_zero: ") + hipGetErrorString(e));
    return NEMO2D_OK;
}

}  // extern "C"
