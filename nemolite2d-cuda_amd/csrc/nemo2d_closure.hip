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

// Forward: tile of 32 pixels x 32 output channels per wave.
// wt: [9][32][cinp] (cout-major, zero-padded), x: [P][cinp], out z: [P][coutp] FP32,
// act: [P][coutp] BF16 (gelu(z) or z).
__global__ void __launch_bounds__(NB)
k_conv_fwd(const uint16_t* __restrict__ x, Img im, int cinp, const uint16_t* __restrict__ wt,
           const float* __restrict__ bias, int cout, int coutp, float* __restrict__ z,
           uint16_t* __restrict__ act, int use_gelu) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long tile0 = ((long)blockIdx.x * (NB / 64) + wave) * 32;
    const long P = im.pixels();
    if (tile0 >= P) return;  // whole wave leaves together
    const long p = tile0 + (lane & 31);
    const int half = lane >> 5;
    f32x16 acc = {0};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
        const long ps = p < P ? shifted(im, p, dy, dx) : -1;
        for (int c0 = 0; c0 < cinp; c0 += 8) {
            bf16x4 a = {0, 0, 0, 0};
            if (ps >= 0) a = load4(x + ps * cinp + c0 + 4 * half);
            const bf16x4 b = load4(wt + ((long)tap * 32 + (lane & 31)) * cinp + c0 + 4 * half);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    const int co = lane & 31;
    if (co >= cout) return;
    const float bc = bias[co];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const long q = tile0 + 8 * (r >> 2) + 4 * half + (r & 3);
        if (q >= P) continue;
        const float v = acc[r] + bc;
        z[q * coutp + co] = v;
        act[q * coutp + co] = f2bf(use_gelu ? gelu(v) : v);
    }
}

// Data gradient: tile of 32 pixels x 32 input channels per wave.
// wt: [9][32][coutp] (cin-major), dz: [P][coutp] BF16, dx: [P][cinp] FP32.
__global__ void __launch_bounds__(NB)
k_conv_bwd_data(const uint16_t* __restrict__ dz, Img im, int coutp, const uint16_t* __restrict__ wt,
                int cin, int cinp, float* __restrict__ dx) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long tile0 = ((long)blockIdx.x * (NB / 64) + wave) * 32;
    const long P = im.pixels();
    if (tile0 >= P) return;
    const long p = tile0 + (lane & 31);
    const int half = lane >> 5;
    f32x16 acc = {0};
    for (int tap = 0; tap < 9; ++tap) {
        const int dy = tap / 3 - 1, dxo = tap % 3 - 1;
        const long ps = p < P ? shifted(im, p, -dy, -dxo) : -1;
        for (int c0 = 0; c0 < coutp; c0 += 8) {
            bf16x4 a = {0, 0, 0, 0};
            if (ps >= 0) a = load4(dz + ps * coutp + c0 + 4 * half);
            const bf16x4 b = load4(wt + ((long)tap * 32 + (lane & 31)) * coutp + c0 + 4 * half);
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    const int ci = lane & 31;
    if (ci >= cin) return;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const long q = tile0 + 8 * (r >> 2) + 4 * half + (r & 3);
        if (q < P) dx[q * cinp + ci] = acc[r];
    }
}

// Weight gradient, split over pixel chunks: block (tap, chunk) writes the
// partial dW[tap] (32x32) of its chunk to work[chunk][tap][ci][co].
constexpr int WCHUNK = 4096;  // pixels per chunk (per wave: WCHUNK/4)
__global__ void __launch_bounds__(NB)
k_conv_bwd_weight(const uint16_t* __restrict__ x, int cinp, const uint16_t* __restrict__ dz,
                  int coutp, Img im, float* __restrict__ work) {
    const int tap = blockIdx.y, chunk = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int half = lane >> 5, rc = lane & 31;
    const int dy = tap / 3 - 1, dxo = tap % 3 - 1;
    const long P = im.pixels();
    const long p0 = (long)chunk * WCHUNK + (long)wave * (WCHUNK / 4);
    f32x16 acc = {0};
    for (long k0 = 0; k0 < WCHUNK / 4; k0 += 8) {
        bf16x4 a, b;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long p = p0 + k0 + 4 * half + v;
            short av = 0, bv = 0;
            if (p < P) {
                const long ps = shifted(im, p, dy, dxo);
                if (ps >= 0 && rc < cinp) av = (short)x[ps * cinp + rc];
                if (rc < coutp) bv = (short)dz[p * coutp + rc];
            }
            a[v] = av;
            b[v] = bv;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    // reduce the 4 waves' tiles in LDS, fixed order
    __shared__ float red[4][32 * 32];
#pragma unroll
    for (int r = 0; r < 16; ++r) red[wave][(8 * (r >> 2) + 4 * half + (r & 3)) * 32 + rc] = acc[r];
    __syncthreads();
    for (int e = threadIdx.x; e < 1024; e += NB)
        work[((long)chunk * 9 + tap) * 1024 + e] = red[0][e] + red[1][e] + red[2][e] + red[3][e];
}

// dW[tap][ci][co] = sum over chunks (fixed order); db[co] = sum_p dz[p][co].
__global__ void k_conv_wsum(const float* __restrict__ work, int nchunks, int cin, int cout,
                            float* __restrict__ dw) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;  // over 9*1024
    if (e >= 9 * 1024) return;
    const int tap = e / 1024, ci = (e % 1024) / 32, co = e % 32;
    if (ci >= cin || co >= cout) return;
    float s = 0.f;
    for (int c = 0; c < nchunks; ++c) s += work[(long)c * 9 * 1024 + e];
    dw[((long)tap * cin + ci) * cout + co] = s;
}

// Bias gradient, pass 1: block b sums each channel (< 32) over pixels
// [b*WCHUNK, (b+1)*WCHUNK); thread (g, c) takes pixels g, g+8, ... of the
// chunk, so a row of the block reads 8 consecutive pixels' channels at once.
__global__ void __launch_bounds__(NB)
k_bias_grad_part(const uint16_t* __restrict__ dz, long P, int coutp, int cout,
                 float* __restrict__ part) {
    __shared__ float red[NB];
    const int co = threadIdx.x & 31, g = threadIdx.x >> 5;  // NB/32 pixel groups
    const long p0 = (long)blockIdx.x * WCHUNK;
    const long p1 = p0 + WCHUNK < P ? p0 + WCHUNK : P;
    float s = 0.f;
    if (co < cout)
        for (long p = p0 + g; p < p1; p += NB / 32) s += bf2f(dz[p * coutp + co]);
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.f;
        for (int q = 0; q < NB / 32; ++q) t += red[q * 32 + co];
        part[(long)blockIdx.x * 32 + co] = t;
    }
}
// pass 2: the chunks' sums in chunk order
__global__ void k_bias_grad_sum(const float* __restrict__ part, int nchunks, int cout,
                                float* __restrict__ db) {
    const int co = threadIdx.x;
    if (co >= cout) return;
    float t = 0.f;
    for (int b = 0; b < nchunks; ++b) t += part[(long)b * 32 + co];
    db[co] = t;
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
                       const float* bias, int cout, float* z, void* act, int use_gelu,
                       void* stream) {
    if (!x || !wf || !bias || !z || !act || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_fwd: bad argument");
    const Img im{n, h, w};
    const long P = (long)n * h * w;
    const long tiles = (P + 31) / 32;
    k_conv_fwd<<<(unsigned)((tiles + 3) / 4), NB, 0, (hipStream_t)stream>>>(
        (const uint16_t*)x, im, pad8(cin), (const uint16_t*)wf, bias, cout, pad8(cout), z,
        (uint16_t*)act, use_gelu);
    return launched("conv_fwd");
}

int nemo2d_nn_conv_bwd_data(const void* dz, int n, int h, int w, int cout, const void* wb, int cin,
                            float* dx, void* stream) {
    if (!dz || !wb || !dx || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_bwd_data: bad argument");
    const Img im{n, h, w};
    const long P = (long)n * h * w;
    const long tiles = (P + 31) / 32;
    k_conv_bwd_data<<<(unsigned)((tiles + 3) / 4), NB, 0, (hipStream_t)stream>>>(
        (const uint16_t*)dz, im, pad8(cout), (const uint16_t*)wb, cin, pad8(cin), dx);
    return launched("conv_bwd_data");
}

// work: at least nemo2d_nn_wgrad_work_floats(n*h*w) floats.
long nemo2d_nn_wgrad_work_floats(long pixels) { return ((pixels + WCHUNK - 1) / WCHUNK) * 9 * 1024; }

int nemo2d_nn_conv_bwd_weight(const void* x, const void* dz, int n, int h, int w, int cin, int cout,
                              float* dw, float* db, float* work, void* stream) {
    if (!x || !dz || !dw || !db || !work || n < 1 || h < 1 || w < 1 || bad_ch(cin) || bad_ch(cout))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_nn_conv_bwd_weight: bad argument");
    const Img im{n, h, w};
    const long P = (long)n * h * w;
    const int nchunks = (int)((P + WCHUNK - 1) / WCHUNK);
    hipStream_t s = (hipStream_t)stream;
    k_conv_bwd_weight<<<dim3(nchunks, 9), NB, 0, s>>>((const uint16_t*)x, pad8(cin),
                                                        (const uint16_t*)dz, pad8(cout), im, work);
    int rc = launched("conv_bwd_weight");
    if (rc) return rc;
    k_conv_wsum<<<(9 * 1024 + 255) / 256, 256, 0, s>>>(work, nchunks, cin, cout, dw);
    if ((rc = launched("conv_wsum"))) return rc;
    // the weight partials are consumed; the work buffer takes the bias
    // partials (nchunks x 32 floats, within its nchunks x 9 x 1024)
    k_bias_grad_part<<<nchunks, NB, 0, s>>>((const uint16_t*)dz, P, pad8(cout), cout, work);
    if ((rc = launched("bias_grad_part"))) return rc;
    k_bias_grad_sum<<<1, 32, 0, s>>>(work, nchunks, cout, db);
    return launched("bias_grad_sum");
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
