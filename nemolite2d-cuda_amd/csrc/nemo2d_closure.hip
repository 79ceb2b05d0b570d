// Learned closure: kernels of a small 3x3-convolution network (no
// counterpart in the reference).
//
// Activations are NHWC with the channel stride rounded up to 8; weights and
// activations are BF16, accumulation is FP32 on the matrix cores
// (v_mfma_f32_32x32x16_bf16).  Every convolution is an implicit GEMM: the
// im2col operand is gathered on the fly from the NHWC tensor, never stored.
//   forward       Z[p, co]  = b[co] + sum_{tap, ci} X[p + off(tap), ci] W[tap, ci, co]
//   data grad     dX[p, ci] = sum_{tap, co} dZ[p - off(tap), co] W[tap, ci, co]
//   weight grad   dW[tap, ci, co] = sum_p X[p + off(tap), ci] dZ[p, co]
// One wave computes a 32x32 tile of the GEMM.  MFMA operand layout
// (32x32x16): lane l holds A row l%32 and B column l%32 at the 8
// consecutive k of half l/32; accumulator register r of lane l is row
// 8*(r/4) + 4*(l/32) + r%4, column l%32.  An image's element offsets fit in
// 32 bits (h w <= 2^24 pixels, checked at launch); the image's base is 64-bit.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "nemo2d.h"
#include "nemo2d_core.h"
#include "nemo2d_internal.h"

using namespace nemo2d;

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int NB = 256;  // threads per block (4 waves)
constexpr int NBC = 512; // threads per block of the convolutions (8 waves, one tile row each)

__device__ __forceinline__ uint16_t f2bf(float f) { return bf16_rne(f); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Two FP32 values as packed BF16 (round to nearest even: the gfx950
// v_cvt_pk_bf16_f32 conversion), a in the low half.
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t){a, b}, bf16x2_t));
}

// GELU(z) = z/2 (1 + erf(z/sqrt2)) and its derivative, with erf by
// Abramowitz & Stegun 7.1.26 (|error| <= 1.5e-7; the values are rounded to
// BF16, whose half ulp is 2e-3 relative): for a = |z|/sqrt2 and t = 1/(1 +
// p a), 1 - erf(a) = t P(t) exp(-a^2).  With E = exp(-z^2/2)/2 and H = t P(t) E
// that gives, without cancellation on either side of 0,
//   GELU(z)  = max(z, 0) - |z| H
//   GELU'(z) = (z >= 0 ? 1 - H : H) + 2 z E / sqrt(2 pi)
// and E is one v_exp_f32: 2^(-z^2 log2(e)/2 - 1).
struct GeluParts {
    float tp;  // t P(t)
    float e;   // exp(-z^2/2) / 2
};
__device__ __forceinline__ GeluParts gelu_parts(float z) {
    const float t = __builtin_amdgcn_rcpf(fmaf(fabsf(z), 0.3275911f * 0.70710678118654752f, 1.0f));
    float p = fmaf(1.061405429f, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    const float e = __builtin_amdgcn_exp2f(fmaf(z * -0.72134752044448170f, z, -1.0f));
    return GeluParts{p * t, e};
}
__device__ __forceinline__ float gelu(float z) {
    const GeluParts g = gelu_parts(z);
    return fmaf(-fabsf(z) * g.e, g.tp, fmaxf(z, 0.0f));
}
__device__ __forceinline__ float gelu_grad(float z) {
    const GeluParts g = gelu_parts(z);
    const float h = g.tp * g.e;
    return fmaf(z * 0.79788456080286536f, g.e, z >= 0.0f ? 1.0f - h : h);
}

struct Img {
    int n, h, w;
    __device__ long pixels() const { return (long)n * h * w; }
};

// ---- forward and data-gradient convolutions -------------------------------
// A tile is CT_H rows x CT_W = 32 pixels of one image.  The kernels are
// persistent: block b takes tiles b, b + G, b + 2G, ... (G blocks, as many as
// stay resident), with the 9 taps' weights staged in LDS once per block and
// the input tiles double-buffered: the next tile's global loads are issued
// before this tile's MFMAs and written to the other LDS buffer after its
// epilogue, so their latency hides behind the arithmetic.  A tile in LDS is
// pixel-major with its 1-pixel halo (the channels of a pixel contiguous,
// padded to the MFMA K step, zero outside the image); each of the 4 waves
// computes CT_H/4 row segments, one 32-channel x 32-pixel MFMA tile
// (v_mfma_f32_32x32x16_bf16) each, over the 9 taps.  Operand layout
// (32x32x16): lane l holds A row l%32 and B column l%32 at 8 consecutive k
// of half l/32; the K order is the same for both operands, so any
// assignment of k to (channel) values works.  The weights are the A operand
// (rows = output channels) and the pixels the B operand (columns), so that
// accumulator register r of lane l holds channel 8*(r/4) + 4*(l/32) + r%4 of
// pixel l%32: each lane owns runs of 4 consecutive channels of one pixel,
// stored 8 (BF16) or 16 (FP32) bytes at a time.
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

// Raw buffer loads within one image (base, bytes <= 2^31): an offset at or
// past `bytes` reads as zero by the hardware's range check (0x00020000: the
// gfx9 raw-buffer word 3), so lanes outside the image or past the channels
// load without a branch, and as many loads are in flight on every path: the
// waits can count them.
constexpr unsigned OOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t img_rsrc(const void* base, long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint4 ld16(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ uint2 ld8(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// The accumulator layout gives lane (c, half) channels 8g + 4 half .. +3
// of pixel c for g = 0..3: four 8-byte pieces in BF16.  One
// v_permlane32_swap_b32 per dword exchanges lane c's pieces of group 2p + 1
// with lane c + 32's of group 2p (it swaps the upper 32 lanes of its first
// operand with the lower 32 lanes of its second), after which lane (c, half)
// holds the 16 contiguous bytes of channels 8 (2p + half) .. +7: stores of
// 16 bytes, half as many.  The same exchange maps loaded 16-byte pieces back
// to the accumulator layout.  All 64 lanes must be active.
__device__ __forceinline__ void pair_swap(uint32_t& a, uint32_t& b) {
    const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = r[0];
    b = r[1];
}
// lo, hi: groups 2p, 2p + 1 in the accumulator layout -> channels 8 (2p + half) .. +7
__device__ __forceinline__ uint4 to_runs(uint2 lo, uint2 hi) {
    pair_swap(lo.x, hi.x);
    pair_swap(lo.y, hi.y);
    return make_uint4(lo.x, lo.y, hi.x, hi.y);
}
// the inverse: channels 8 (2p + half) .. +7 -> groups 2p (lo), 2p + 1 (hi)
__device__ __forceinline__ void from_runs(uint4 v, uint2& lo, uint2& hi) {
    uint32_t a = v.x, b = v.y, c = v.z, d = v.w;
    pair_swap(a, c);
    pair_swap(b, d);
    lo = make_uint2(a, b);
    hi = make_uint2(c, d);
}

// Tile geometry of tile index t.
struct TilePos {
    int img, y0, x0;
};
__device__ __forceinline__ TilePos tile_pos(const Img& im, int t) {
    const int tiles_x = (im.w + CT_W - 1) / CT_W, tiles_y = (im.h + CT_H - 1) / CT_H;
    return TilePos{t / (tiles_x * tiles_y), (t / tiles_x) % tiles_y * CT_H, t % tiles_x * CT_W};
}

// One tile's global data src[img][y0-1 .. y0+CT_H][x0-1 .. x0+CT_W][0..cp)
// in registers (channels cp..KT-1 and pixels outside the image zero), then
// into LDS tile[pixel][KT + 8].
template <int KT>
struct TileRegs {
    static constexpr int CH = KT / 8;  // 16-byte chunks per pixel
    static constexpr int IT = (CT_PIX * CH + NBC - 1) / NBC;
    uint4 v[IT];
};
// (all zero with `any` false: no loads)
template <int KT>
__device__ __forceinline__ void tile_load(const uint16_t* __restrict__ src, const Img& im, int cp,
                                          const TilePos& tp, bool any, TileRegs<KT>& r) {
    constexpr int CH = TileRegs<KT>::CH;
    const long ie = (long)im.h * im.w * cp;  // elements of an image
    const auto rs = img_rsrc(src + tp.img * ie, 2 * ie);
#pragma unroll
    for (int s = 0; s < TileRegs<KT>::IT; ++s) {
        const int e = threadIdx.x + s * NBC;
        const int q = e / CH, c8 = (e % CH) * 8;
        const int yy = tp.y0 - 1 + q / (CT_W + 2), xx = tp.x0 - 1 + q % (CT_W + 2);
        const bool in = any && e < CT_PIX * CH && c8 < cp && yy >= 0 && yy < im.h && xx >= 0 && xx < im.w;
        r.v[s] = ld16(rs, in ? 2u * (unsigned)((yy * im.w + xx) * cp + c8) : OOB);
    }
}
template <int KT>
__device__ __forceinline__ void tile_store(const TileRegs<KT>& r, uint16_t* tile) {
    constexpr int PS = KT + 8;  // LDS pixel stride (bf16): +16 B against bank conflicts
    constexpr int CH = TileRegs<KT>::CH;
#pragma unroll
    for (int s = 0; s < TileRegs<KT>::IT; ++s) {
        const int e = threadIdx.x + s * NBC;
        if (e < CT_PIX * CH) *(uint4*)(tile + (e / CH) * PS + (e % CH) * 8) = r.v[s];
    }
}

// The 9 taps' weights wt[tap][32][kp] into LDS wl[tap][32][KT + 8] (K values
// kp..KT-1 zero).
template <int KT>
__device__ __forceinline__ void weights_store(const uint16_t* __restrict__ wt, int kp, uint16_t* wl) {
    constexpr int RS = KT + 8, CH = KT / 8;
    for (int e = threadIdx.x; e < 9 * 32 * CH; e += NBC) {
        const int row = e / CH, c8 = (e % CH) * 8;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (c8 < kp) v = *(const uint4*)(wt + (long)row * kp + c8);
        *(uint4*)(wl + row * RS + c8) = v;
    }
}

// The operands of one tap are read from LDS while the previous tap's MFMAs
// run (two register sets, alternating), so no MFMA waits for its reads.
template <int KT>
struct TapOps {
    bf16x8 w[KT / 16], b0[KT / 16];
};
template <int KT, int SIGN>
__device__ __forceinline__ void tap_read(const uint16_t* tile, const uint16_t* wl, int wave, int c, int half,
                                         int tap, TapOps<KT>& o) {
    constexpr int PS = KT + 8, RS = KT + 8;
    const int dy = SIGN * (tap / 3 - 1), dx = SIGN * (tap % 3 - 1);
    const uint16_t* wr = wl + (tap * 32 + c) * RS + 8 * half;
    const uint16_t* a0 = tile + ((wave + 1 + dy) * (CT_W + 2) + c + 1 + dx) * PS + 8 * half;
#pragma unroll
    for (int kk = 0; kk < KT / 16; ++kk) {
        o.w[kk] = lds8(wr + kk * 16);
        o.b0[kk] = lds8(a0 + kk * 16);
    }
}
// One 3x3 convolution pass over a tile in LDS: wave w computes row w of the
// tile (one 32 x 32 accumulator).  SIGN = +1: forward (input at p +
// off(tap)); -1: data gradient (dZ at p - off(tap)).
template <int KT, int SIGN>
__device__ __forceinline__ void conv_row(const uint16_t* tile, const uint16_t* wl, int wave, int c,
                                         int half, f32x16& acc0) {
    TapOps<KT> op[2];
    tap_read<KT, SIGN>(tile, wl, wave, c, half, 0, op[0]);
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        if (tap + 1 < 9) tap_read<KT, SIGN>(tile, wl, wave, c, half, tap + 1, op[(tap + 1) & 1]);
        // keep the reads ahead of the MFMAs (the scheduler would sink each
        // read to its use, exposing its latency before every MFMA)
        __builtin_amdgcn_sched_barrier(0);
        const TapOps<KT>& o = op[tap & 1];
#pragma unroll
        for (int kk = 0; kk < KT / 16; ++kk)
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(o.w[kk], o.b0[kk], acc0, 0, 0, 0);
    }
}

static_assert(CT_H * 64 == NBC, "each of the 8 waves owns one row of the tile");

// The persistent tile loop shared by the forward and data-gradient kernels:
// epi.init() loads what every tile's epilogue needs, epi.prefetch(tile)
// issues a tile's epilogue loads before its MFMAs, epi.compute(tile,
// accumulators) forms its outputs in registers and epi.store() writes them.
// Vector memory counters count loads and stores together, in issue order,
// so a wait for a load also waits for every store issued before it.  Each
// iteration therefore issues the previous tile's stores, this tile's
// epilogue loads and the next tile's loads, in that order, and waits only
// twice: in epi.compute for the epilogue loads (the next tile's still in
// flight), and after the MFMAs and the epilogue's arithmetic for the next
// tile, which then goes from registers to the other LDS buffer.  The
// stores' counts are fixed (predicated, no data-dependent loops) and the
// next tile's loads and LDS writes run on every iteration (no loads and
// zeros after the last tile), so no wait is needed before the loads.
template <int KT, int SIGN, class Epi>
__device__ __forceinline__ void conv_tiles(const uint16_t* __restrict__ src, const Img& im, int cp,
                                           const uint16_t* __restrict__ wt, int ntiles, uint16_t* tiles,
                                           uint16_t* wl, Epi& epi) {
    constexpr int TS = CT_PIX * (KT + 8);  // one LDS tile buffer (bf16)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5, c = lane & 31;
    int t = blockIdx.x;
    if (t >= ntiles) return;  // (uniform per block)
    TileRegs<KT> r;
    tile_load<KT>(src, im, cp, tile_pos(im, t), true, r);
    epi.init();
    weights_store<KT>(wt, cp, wl);  // (the weights' K is the input's channels)
    tile_store<KT>(r, tiles);
    __syncthreads();
    int cur = 0;
    for (bool prev = false; t < ntiles; t += gridDim.x, prev = true) {
        const int nxt = t + gridDim.x;
        if (prev) epi.store();           // the previous tile's
        epi.prefetch(tile_pos(im, t));  // this tile's epilogue operands, needed after the MFMAs
        tile_load<KT>(src, im, cp, tile_pos(im, nxt < ntiles ? nxt : t), nxt < ntiles, r);  // in flight meanwhile
        f32x16 acc = {0};
        conv_row<KT, SIGN>(tiles + cur * TS, wl, wave, c, half, acc);
        epi.compute(tile_pos(im, t), acc);
        // (keep the LDS writes, and so the wait for the next tile's loads,
        // after the epilogue's arithmetic)
        __builtin_amdgcn_sched_barrier(0);
        tile_store<KT>(r, tiles + (1 - cur) * TS);  // (that buffer's readers passed the last barrier)
        __syncthreads();  // the next tile is in LDS, and no wave still reads this one
        cur ^= 1;
    }
    epi.store();
}
// LDS of the persistent convolution kernels: two tile buffers and the
// weights, (2 * CT_PIX + 9 * 32) * (KT + 8) bf16 (77,440 B for KT = 32).
template <int KT>
constexpr int conv_lds_elems() { return (2 * CT_PIX + 9 * 32) * (KT + 8); }

// Forward: z = b + sum_tap sum_ci X[p + off(tap)][ci] W[tap][ci][co];
// act = gelu(z) (or z with GELU false).  wt: [9][32][cinp] (cout-major,
// from pack).  The epilogue: lane (c, half) of wave w holds, for row w of
// the tile, channels co0 .. co0+3 (co0 = 8g + 4 half) of pixel c in
// accumulator registers 4g .. 4g+3.  The packed weights and the bias of
// channels co >= cout are zero, so those accumulators, z, and GELU(z) are
// zero too (for finite inputs) and need no masking.
template <bool GELU>
struct FwdEpi {
    Img im;
    const float* bias;
    int cout, coutp;
    float* z;
    uint16_t* zb;
    uint16_t* act;
    int wave, c, half;
    TilePos tp;
    uint4 ob[2], oa[2];  // BF16 z and act, channels 8 (2p + half) .. +7
    float4 oz[4];
    float bv[4][4];  // bias of channels 8g + 4 half + r (0 past cout)
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = 8 * g + 4 * half + r;
                const float b = bias[co < cout ? co : 0];  // (always in bounds)
                bv[g][r] = co < cout ? b : 0.f;
            }
    }
    __device__ __forceinline__ void prefetch(const TilePos&) {}
    __device__ __forceinline__ void compute(const TilePos& t, const f32x16& acc) {
        tp = t;
        uint2 b[4], a[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            float v[4], va[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = acc[4 * g + r] + bv[g][r];
                va[r] = GELU ? gelu(v[r]) : v[r];
            }
            oz[g] = make_float4(v[0], v[1], v[2], v[3]);
            b[g] = make_uint2(pk_bf16(v[0], v[1]), pk_bf16(v[2], v[3]));
            a[g] = make_uint2(pk_bf16(va[0], va[1]), pk_bf16(va[2], va[3]));
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            ob[p] = to_runs(b[2 * p], b[2 * p + 1]);
            oa[p] = to_runs(a[2 * p], a[2 * p + 1]);
        }
        // computed here, ahead of the next tile's LDS writes and the wait for
        // its loads, rather than sunk towards their stores
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            asm volatile("" : "+v"(ob[p].x), "+v"(ob[p].y), "+v"(ob[p].z), "+v"(ob[p].w) :: "memory");
            asm volatile("" : "+v"(oa[p].x), "+v"(oa[p].y), "+v"(oa[p].z), "+v"(oa[p].w) :: "memory");
        }
    }
    __device__ __forceinline__ void store() const {
        const int xx = tp.x0 + c, y = tp.y0 + wave;  // this lane's pixel
        if (xx >= im.w || y >= im.h) return;
        const long ib = (long)tp.img * im.h * im.w * coutp;  // the image (uniform)
        const unsigned q0 = (unsigned)((y * im.w + xx) * coutp);
        if (z) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int co0 = 8 * g + 4 * half;
                if (co0 >= coutp) continue;  // (predicated: the loop is unrolled)
                const unsigned q = q0 + co0;
                if (co0 + 4 <= cout) {
                    *(float4*)(z + ib + q) = oz[g];
                } else {
                    const float v[4] = {oz[g].x, oz[g].y, oz[g].z, oz[g].w};
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (co0 + r < cout) z[ib + q + r] = v[r];
                }
            }
        }
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int c8 = 8 * (2 * p + half);
            if (c8 >= coutp) continue;
            if (zb) *(uint4*)(zb + ib + q0 + c8) = ob[p];
            if (act) *(uint4*)(act + ib + q0 + c8) = oa[p];
        }
    }
};

template <int KT, bool GELU>
__global__ void __launch_bounds__(NBC, 4)  // 4 waves per SIMD: two blocks per CU
k_conv_fwd(const uint16_t* __restrict__ x, Img im, int cinp, const uint16_t* __restrict__ wt,
           const float* __restrict__ bias, int cout, int coutp, float* __restrict__ z,
           uint16_t* __restrict__ zb, uint16_t* __restrict__ act, int ntiles) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[conv_lds_elems<KT>()];
    const int lane = threadIdx.x & 63;
    FwdEpi<GELU> epi;
    epi.im = im;
    epi.bias = bias;
    epi.cout = cout;
    epi.coutp = coutp;
    epi.z = z;
    epi.zb = zb;
    epi.act = act;
    epi.wave = threadIdx.x >> 6;
    epi.c = lane & 31;
    epi.half = lane >> 5;
    conv_tiles<KT, 1>(x, im, cinp, wt, ntiles, lds, lds + 2 * CT_PIX * (KT + 8), epi);
}

// Data gradient: dX[p][ci] = sum_tap sum_co dZ[p - off(tap)][co] W[tap][ci][co];
// with GL, the pre-activation of the layer below (zp FP32, or zpb BF16 with
// ZB) makes the output that layer's dz = BF16(dX * GELU'(z)) (GELU backward
// fused), else FP32 dX.  wt: [9][32][coutp] (cin-major, from pack).  The
// packed weights of channels ci >= cin are zero, so those accumulators are
// zero; their pre-activations are read as zero (GELU'(0) = 1/2), so the
// padded channels of dz are zero whatever the layer below's buffer holds.
template <bool GL, bool ZB>
struct BwdEpi {
    Img im;
    int cin, cinp;
    float* dx;
    const float* zp;
    const uint16_t* zpb;
    uint16_t* dzp;
    int wave, c, half;
    TilePos tp;
    uint4 od[2];  // BF16 dz below, channels 8 (2p + half) .. +7
    float4 ox[4];
    // the layer below's pre-activations at this lane's pixel, as loaded
    // (with ZB BF16 channels 8 (2p + half) .. +7, else FP32 channels 8g + 4
    // half .. +3; converted after the MFMAs, so that the loads' wait comes
    // there)
    uint4 zb8[2];
    float4 zf4[4];
    // this lane's pixel of tile p: false past the image; else its offset
    // within the image (q) and the image's (ib, uniform)
    __device__ __forceinline__ bool at(const TilePos& p, long& ib, unsigned& q) const {
        const int xx = p.x0 + c, y = p.y0 + wave;
        ib = (long)p.img * im.h * im.w * cinp;
        q = (unsigned)((y * im.w + xx) * cinp);
        return xx < im.w && y < im.h;
    }
    __device__ __forceinline__ void init() {}
    __device__ __forceinline__ void prefetch(const TilePos& p) {
        if (!GL) return;
        long ib;
        unsigned q0;
        const bool in = at(p, ib, q0);
        const long ie = (long)im.h * im.w * cinp;  // elements of an image
        const auto rs = ZB ? img_rsrc(zpb + ib, 2 * ie) : img_rsrc(zp + ib, 4 * ie);
        if (ZB) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int c8 = 8 * (2 * p + half);
                zb8[p] = ld16(rs, in && c8 < cinp ? 2u * (q0 + c8) : OOB);
            }
        } else {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int ci0 = 8 * g + 4 * half;
                zf4[g] = __builtin_bit_cast(float4, ld16(rs, in && ci0 < cinp ? 4u * (q0 + ci0) : OOB));
            }
        }
    }
    __device__ __forceinline__ void compute(const TilePos& t, const f32x16& acc) {
        tp = t;
        uint2 zb2[4], od2[4];
        if (GL && ZB) {
#pragma unroll
            for (int p = 0; p < 2; ++p) from_runs(zb8[p], zb2[2 * p], zb2[2 * p + 1]);
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            ox[g] = make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
            if (GL) {
                const int ci0 = 8 * g + 4 * half;
                float zv[4];
                if (ZB) {
                    const uint2 u = zb2[g];
                    zv[0] = bf2f(u.x & 0xffff); zv[1] = bf2f(u.x >> 16);
                    zv[2] = bf2f(u.y & 0xffff); zv[3] = bf2f(u.y >> 16);
                } else {
                    zv[0] = zf4[g].x; zv[1] = zf4[g].y; zv[2] = zf4[g].z; zv[3] = zf4[g].w;
                }
                if (ci0 + 4 > cin) {  // a group across cin (none when cin = cinp)
#pragma unroll
                    for (int r = 0; r < 4; ++r) zv[r] = ci0 + r < cin ? zv[r] : 0.f;
                }
                float o[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) o[r] = acc[4 * g + r] * gelu_grad(zv[r]);
                od2[g] = make_uint2(pk_bf16(o[0], o[1]), pk_bf16(o[2], o[3]));
            }
        }
        // computed here, ahead of the next tile's LDS writes and the wait for
        // its loads, rather than sunk towards their stores
        if (GL) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                od[p] = to_runs(od2[2 * p], od2[2 * p + 1]);
                asm volatile("" : "+v"(od[p].x), "+v"(od[p].y), "+v"(od[p].z), "+v"(od[p].w) :: "memory");
            }
        } else {
#pragma unroll
            for (int g = 0; g < 4; ++g)
                asm volatile("" : "+v"(ox[g].x), "+v"(ox[g].y), "+v"(ox[g].z), "+v"(ox[g].w) :: "memory");
        }
    }
    __device__ __forceinline__ void store() const {
        long ib;
        unsigned q0;
        if (!at(tp, ib, q0)) return;
        if (GL) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int c8 = 8 * (2 * p + half);
                if (c8 < cinp) *(uint4*)(dzp + ib + q0 + c8) = od[p];
            }
            return;
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int ci0 = 8 * g + 4 * half;
            if (ci0 >= cinp) continue;  // (predicated: the loop is unrolled)
            {
                const float v[4] = {ox[g].x, ox[g].y, ox[g].z, ox[g].w};
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (ci0 + r < cin) dx[ib + q0 + ci0 + r] = v[r];
            }
        }
    }
};

template <int KT, bool GL, bool ZB>
__global__ void __launch_bounds__(NBC, 4)  // 4 waves per SIMD: two blocks per CU
k_conv_bwd_data(const uint16_t* __restrict__ dz, Img im, int coutp, const uint16_t* __restrict__ wt,
                int cin, int cinp, float* __restrict__ dx, const float* __restrict__ zp,
                const uint16_t* __restrict__ zpb, uint16_t* __restrict__ dzp, int ntiles) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[conv_lds_elems<KT>()];
    const int lane = threadIdx.x & 63;
    BwdEpi<GL, ZB> epi;
    epi.im = im;
    epi.cin = cin;
    epi.cinp = cinp;
    epi.dx = dx;
    epi.zp = zp;
    epi.zpb = zpb;
    epi.dzp = dzp;
    epi.wave = threadIdx.x >> 6;
    epi.c = lane & 31;
    epi.half = lane >> 5;
    conv_tiles<KT, -1>(dz, im, coutp, wt, ntiles, lds, lds + 2 * CT_PIX * (KT + 8), epi);
}

// ---- weight gradient ---------------------------------------------------------
// dW[tap][ci][co] = sum_p X[p + off(tap)][ci] dZ[p][co] and db[co] = sum_p dZ[p][co].
// A block of 10 waves (waves 0..8: one tap each) walks tiles of WG_H rows x WG_W pixels of
// 4 images at once.  Each tile is staged in LDS transposed to
// [channel][y][x][image], so that a lane's 8 K values (2 neighbouring pixels
// x 4 images) of one channel are 16 contiguous bytes for every tap shift;
// the next tile's global loads are issued before the current tile's MFMAs.
// Wave 9 multiplies dZ by a ones operand: every row of that product is the
// bias gradient.  Each block writes its partial sums
// (9 x 32 x 32 + 32 floats); k_wgrad_sum adds the blocks in fixed order.
constexpr int WG_H = 2, WG_W = 32, WG_NT = 10 * 64;
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

// One item: channels 4 c4 .. 4 c4 + 3 of pixel (y, x) in the 4 images of
// group g (zero outside the image, past the channels or past the batch).
// The pixel's offset within an image is one 32-bit value for all 4 images;
// the images' bases are uniform.
__device__ __forceinline__ void ld_px4(const uint16_t* __restrict__ src, const Img& im, int cp, int g, bool ok,
                                       int y, int x, int c4, uint2* v) {
    const bool in = ok && y >= 0 && y < im.h && x >= 0 && x < im.w && c4 * 4 < cp;
    const long is = (long)im.h * im.w * cp;
    const uint16_t* b = src + 4L * g * is;
    const unsigned q = (unsigned)((y * im.w + x) * cp + c4 * 4);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        v[n] = make_uint2(0, 0);
        if (in && 4 * g + n < im.n) v[n] = *(const uint2*)(b + n * is + q);
    }
}

__device__ __forceinline__ void wg_load(const uint16_t* __restrict__ x, int cinp, const uint16_t* __restrict__ dz,
                                        int coutp, const Img& im, int g, int y0, int x0, WgItems& it) {
    const int cq = cinp / 4, zq = coutp / 4;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int e = threadIdx.x + s * WG_NT;
        const int q = e / cq, c4 = e % cq;
        const int yy = y0 - 1 + q / (WG_W + 2), xx = x0 - 1 + q % (WG_W + 2);
        ld_px4(x, im, cinp, g, q < WG_XP, yy, xx, c4, it.xv[s]);
    }
    {
        const int e = threadIdx.x;
        const int q = e / zq, c4 = e % zq;
        const int yy = y0 + q / WG_W, xx = x0 + q % WG_W;
        ld_px4(dz, im, coutp, g, q < WG_H * WG_W, yy, xx, c4, it.zv);
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

// The MFMAs of one staged tile for one wave: K steps k of pixels xq, xq + 1
// (xq = x4 + 2 half) of row y, images 0..3.  A is X shifted by the wave's
// tap, or ones for the bias wave; B is dZ.  (Channel planes past the
// channels are zero, so the reads need no masks.)  Step k + 1's operands
// are read while step k's MFMA runs.
template <bool BIAS>
__device__ __forceinline__ void wg_operands(const uint32_t* xs, const uint32_t* zs, int c, int half, int dy,
                                            int dxo, int k, bf16x8& a, bf16x8& bz) {
    constexpr int NX = WG_W / 4;
    const int y = k / NX, xq = (k % NX) * 4 + 2 * half;
    if (BIAS) {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = (short)0x3f80;  // BF16 1.0
    } else {
        const uint32_t* pa = xs + c * WG_CSX + 2 * ((y + 1 + dy) * (WG_W + 2) + xq + 1 + dxo);
        a = as_bf16x8(*(const uint2*)pa, *(const uint2*)(pa + 2));
    }
    const uint4 v = *(const uint4*)(zs + c * WG_CSZ + 2 * (y * WG_W + xq));
    bz = as_bf16x8(make_uint2(v.x, v.y), make_uint2(v.z, v.w));
}
template <bool BIAS>
__device__ __forceinline__ void wg_mfmas(const uint32_t* xs, const uint32_t* zs, int c, int half, int dy,
                                         int dxo, f32x16& acc) {
    constexpr int NK = WG_H * (WG_W / 4);
    bf16x8 a[2], bz[2];
    wg_operands<BIAS>(xs, zs, c, half, dy, dxo, 0, a[0], bz[0]);
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        if (k + 1 < NK) wg_operands<BIAS>(xs, zs, c, half, dy, dxo, k + 1, a[(k + 1) & 1], bz[(k + 1) & 1]);
        __builtin_amdgcn_sched_barrier(0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[k & 1], bz[k & 1], acc, 0, 0, 0);
    }
}

__global__ void __launch_bounds__(WG_NT, 5)  // 5 waves per SIMD: 2 blocks per CU
k_conv_bwd_weight(const uint16_t* __restrict__ x, int cinp, const uint16_t* __restrict__ dz,
                  int coutp, Img im, float* __restrict__ work) {
    __shared__ __attribute__((aligned(16))) uint32_t xs[32 * WG_CSX];
    __shared__ __attribute__((aligned(16))) uint32_t zs[32 * WG_CSZ];
    const int lane = threadIdx.x & 63, half = lane >> 5, c = lane & 31;
    const int tap = threadIdx.x >> 6;
    const int dy = tap / 3 - 1, dxo = tap % 3 - 1;
    const int tx = (im.w + WG_W - 1) / WG_W, ty = (im.h + WG_H - 1) / WG_H, ng = (im.n + 3) / 4;
    const int txy = tx * ty, ntiles = txy * ng;  // (< 2^31: checked at launch)
    f32x16 acc = {0};
    const bool bias_wave = tap == 9;
    // the channel planes past cinp / coutp stay zero: wg_store writes only
    // the channels' planes
    for (int e = threadIdx.x; e < 32 * WG_CSX; e += WG_NT) xs[e] = 0;
    for (int e = threadIdx.x; e < 32 * WG_CSZ; e += WG_NT) zs[e] = 0;
    WgItems it;
    int t = blockIdx.x;
    if (t < ntiles) {
        const int g = t / txy, r = t % txy;
        wg_load(x, cinp, dz, coutp, im, g, r / tx * WG_H, r % tx * WG_W, it);
    }
    for (; t < ntiles; t += gridDim.x) {
        __syncthreads();  // the previous tile's MFMAs have read the LDS (and the zeroing is done)
        wg_store(cinp, coutp, it, xs, zs);
        __syncthreads();
        const int tn = t + gridDim.x;
        if (tn < ntiles) {
            const int g = tn / txy, r = tn % txy;
            wg_load(x, cinp, dz, coutp, im, g, r / tx * WG_H, r % tx * WG_W, it);
        }
        if (bias_wave)
            wg_mfmas<true>(xs, zs, c, half, dy, dxo, acc);
        else
            wg_mfmas<false>(xs, zs, c, half, dy, dxo, acc);
    }
    // register k of lane l: row ci = 8*(k/4) + 4*half + k%4, column co = c
    float* out = work + (long)blockIdx.x * WG_PART;
    if (bias_wave) {
        if (half == 0) out[9 * 1024 + c] = acc[0];
        return;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k) out[tap * 1024 + (8 * (k >> 2) + 4 * half + (k & 3)) * 32 + c] = acc[k];
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

// Weighted MSE over channels < ch (<= 8) of [P][cp] tensors, cp = 8: loss
// (one float per block into part), gradient dy (BF16).  A thread takes whole
// pixels: 32 bytes of y and t in, 16 bytes of dy out.
__global__ void __launch_bounds__(NB)
k_mse(const float* __restrict__ y, const float* __restrict__ t, long P, int ch,
      const float* __restrict__ wch, float inv_count, uint16_t* __restrict__ dy,
      float* __restrict__ part) {
    __shared__ float red[NB];
    float w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = c < ch ? wch[c] : 0.f;
    float s = 0.f;
    for (long p = (long)blockIdx.x * NB + threadIdx.x; p < P; p += (long)gridDim.x * NB) {
        const float4* yp = (const float4*)(y + 8 * p);
        const float4* tp = (const float4*)(t + 8 * p);
        const float4 y0 = yp[0], y1 = yp[1], t0 = tp[0], t1 = tp[1];
        const float d[8] = {y0.x - t0.x, y0.y - t0.y, y0.z - t0.z, y0.w - t0.w,
                            y1.x - t1.x, y1.y - t1.y, y1.z - t1.z, y1.w - t1.w};
        uint32_t g[4];
#pragma unroll
        for (int c = 0; c < 8; c += 2) {
            const float d0 = c < ch ? d[c] : 0.f, d1 = c + 1 < ch ? d[c + 1] : 0.f;
            s += w[c] * d0 * d0;
            s += w[c + 1] * d1 * d1;
            g[c / 2] = (uint32_t)f2bf(2.f * w[c] * d0 * inv_count) |
                       (uint32_t)f2bf(2.f * w[c + 1] * d1 * inv_count) << 16;
        }
        *(uint4*)(dy + 8 * p) = make_uint4(g[0], g[1], g[2], g[3]);
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

// Blocks of a persistent convolution kernel: as many as stay resident on the
// device (the occupancy of each kernel, ID, is asked once), at most one per
// tile.
template <int ID, class K>
int conv_grid(K kern, long ntiles) {
    static int resident = 0;  // per ID
    if (!resident) {
        int dev = 0, per_cu = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NBC, 0) != hipSuccess || per_cu < 1)
            return 0;
        resident = per_cu * cus;
    }
    return (int)(ntiles < resident ? ntiles : resident);
}

bool bad_ch(int c) { return c < 1 || c > 32; }
int pad8(int c) { return (c + 7) / 8 * 8; }
// An image's bytes (32 channels at most, FP32 at most) within the 2^31 of
// a buffer load's range: h w <= 2^24 pixels.
bool too_big(int h, int w) { return (long)h * w > (1L << 24); }

}  // namespace

extern "C" {

int nemo2d_nn_conv_fwd(const void* x, int n, int h, int w, int cin, const void* wf,
                       const float* bias, int cout, float* z, void* zb, void* act, int use_gelu,
                       void* stream) {
    if (!x || !wf || !bias || !(z || zb || act) || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_fwd: bad argument");
    const Img im{n, h, w};
    const long nt = (long)n * ((h + CT_H - 1) / CT_H) * ((w + CT_W - 1) / CT_W);
    if (nt > 0x7fffffffL || too_big(h, w))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_fwd: too many pixels");
    const bool k16 = pad8(cin) <= 16;
    int g = 0;
    auto* k = k_conv_fwd<32, true>;
    if (use_gelu) {
        k = k16 ? k_conv_fwd<16, true> : k_conv_fwd<32, true>;
        g = k16 ? conv_grid<0>(k, nt) : conv_grid<1>(k, nt);
    } else {
        k = k16 ? k_conv_fwd<16, false> : k_conv_fwd<32, false>;
        g = k16 ? conv_grid<2>(k, nt) : conv_grid<3>(k, nt);
    }
    if (g < 1) return set_error(NEMO2D_ERR_HIP, "nemo2d_nn_conv_fwd: occupancy query failed");
    k<<<g, NBC, 0, (hipStream_t)stream>>>((const uint16_t*)x, im, pad8(cin), (const uint16_t*)wf, bias, cout,
                                         pad8(cout), z, (uint16_t*)zb, (uint16_t*)act, (int)nt);
    return launched("conv_fwd");
}

static int bwd_data(const void* dz, int n, int h, int w, int cout, const void* wb, int cin, float* dx,
                    const float* zp, const void* zpb, void* dzp, void* stream, const char* who) {
    const Img im{n, h, w};
    const long nt = (long)n * ((h + CT_H - 1) / CT_H) * ((w + CT_W - 1) / CT_W);
    if (nt > 0x7fffffffL || too_big(h, w))
        return set_error(NEMO2D_ERR_ARG, std::string(who) + ": too many pixels");
    const bool k16 = pad8(cout) <= 16;
    int g = 0;
    auto* k = k_conv_bwd_data<32, false, false>;
    if (zpb) {
        k = k16 ? k_conv_bwd_data<16, true, true> : k_conv_bwd_data<32, true, true>;
        g = k16 ? conv_grid<4>(k, nt) : conv_grid<5>(k, nt);
    } else if (zp) {
        k = k16 ? k_conv_bwd_data<16, true, false> : k_conv_bwd_data<32, true, false>;
        g = k16 ? conv_grid<6>(k, nt) : conv_grid<7>(k, nt);
    } else {
        k = k16 ? k_conv_bwd_data<16, false, false> : k_conv_bwd_data<32, false, false>;
        g = k16 ? conv_grid<8>(k, nt) : conv_grid<9>(k, nt);
    }
    if (g < 1) return set_error(NEMO2D_ERR_HIP, std::string(who) + ": occupancy query failed");
    k<<<g, NBC, 0, (hipStream_t)stream>>>((const uint16_t*)dz, im, pad8(cout), (const uint16_t*)wb, cin,
                                         pad8(cin), dx, zp, (const uint16_t*)zpb, (uint16_t*)dzp, (int)nt);
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
    if (too_big(h, w) || (long)n * h * w > 0x7fffffffL)
        return set_error(NEMO2D_ERR_ARG, std::string(who) + ": too many pixels");
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

}  // extern "C"

namespace {

// Block partials of the allocation-free weight gradient: every call on a
// device shares that device's copy (its address comes from
// hipGetSymbolAddress on the first call there).  Each call writes every
// entry it later reads (no zeroing needed); calls that overlap on one device
// race on it.
constexpr long WG_SCRATCH_FLOATS = (long)WG_MAXBLK * WG_PART;  // 18.9 MB
__device__ __align__(256) float scratch[WG_SCRATCH_FLOATS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float* g_scratch[MAX_DEVICES] = {};

int scratch_of_current_device(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string("no HIP device: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(NEMO2D_ERR_HIP, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(NEMO2D_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = (float*)p;
    }
    *out = g_scratch[dev];
    return NEMO2D_OK;
}

}  // namespace

extern "C" {

int nemo2d_nn_conv_bwd_weight_noalloc(const void* x, const void* dz, int n, int h, int w, int cin, int cout,
                                      float* dw, float* db, void* stream) {
    if (!x || !dz || !dw || !db || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_bwd_weight_noalloc: bad argument");
    if ((long)n * h * w > 0x7fffffffL)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_bwd_weight_noalloc: " + std::to_string((long)n * h * w) +
                                             " pixels exceed the maximum of 2147483647");
    static_assert((long)WG_MAXBLK * WG_PART <= WG_SCRATCH_FLOATS, "scratch holds every block's partial");
    float* scr = nullptr;
    const int rc = scratch_of_current_device(&scr);
    if (rc) return rc;
    return bwd_weight(x, dz, n, h, w, cin, cout, dw, db, scr, stream, "conv_bwd_weight_noalloc");
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
    if (!y || !t || !wch || !dy || !part || pixels < 1 || ch < 1 || ch > 8 || nblocks < 1)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_mse: bad argument");
    k_mse<<<nblocks, NB, 0, (hipStream_t)stream>>>(y, t, pixels, ch, wch, 1.0f / (float)(pixels * ch),
                                                   (uint16_t*)dy, part);
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

int nemo2d_dev_synchronize(void) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string("nemo2d_dev_synchronize: ") + hipGetErrorString(e));
    return NEMO2D_OK;
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
