// Device-side building blocks of the FP64 NEMOLite2D step for gfx950.
//
// Storage: every field (T, U, V, the mask flags) lives in one padded index
// space.  Point (i, j) -- Fortran indices of the serial model, i in
// [-1, jpi+2], j in [-1, jpj+2] -- sits at offset base + j*pitch + i, with
// base = pitch + COL0.  Column i = 0 is aligned to 128 bytes; one guard
// column/row beyond the mask's ring on every side is kept zero, so a stencil
// centred anywhere on the ring reads inside the allocation.
//
// The point functions below are the only place the step's arithmetic is
// written; the stage kernels and the fused kernel both call them, so the two
// paths agree bit for bit.  Each keeps the operation order of
// original/nemolite2d.f90 (cited per function); the library is compiled with
// -ffp-contract=off, so no product is fused into an FMA.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nemo2d {

constexpr double GRAV = 9.80665;         // original/nemolite2d.f90:16
constexpr double OMEGA = 7.292116e-05;   // :17
constexpr double PI = 3.1415926535897932; // :15
constexpr int COL0 = 16;                 // storage column of i = 0 (128 B)

struct Layout {
    int jpi, jpj;
    long pitch;  // elements per storage row
    long base;   // offset of (0, 0)
    long rows;   // storage rows (jpj + 4)
    __host__ __device__ long at(int i, int j) const { return base + (long)j * pitch + i; }
    __host__ __device__ long size() const { return pitch * rows; }
};

// Per-point classification, computed once on the host from the mask pt.
enum : unsigned {
    U_CLS_SHIFT = 0,          // 2 bits: class of the U point (i, j)
    U_WALL_S = 1u << 2,       // momentum_u: free-slip south (:459, :481)
    U_WALL_N = 1u << 3,       // momentum_u: free-slip north (:466, :488)
    U_FL_EAST = 1u << 4,      // Flather u: inner point is i+1 (else i-1)
    V_CLS_SHIFT = 5,          // 2 bits: class of the V point (i, j)
    V_WALL_W = 1u << 7,       // momentum_v: free-slip west (:547, :571)
    V_WALL_E = 1u << 8,       // momentum_v: free-slip east (:554, :578)
    V_FL_NORTH = 1u << 9,     // Flather v: inner point is j+1 (else j-1)
    T_CLAMP = 1u << 10,       // bc_ssh clamps ssha here (:633-642)
    SU_SHIFT = 11,            // 2 bits: how next sets sshn_u (:738-746)
    SV_SHIFT = 13,            // 2 bits: how next sets sshn_v (:754-762)
};
enum : unsigned { CLS_NONE = 0, CLS_COMP = 1, CLS_SOLID = 2, CLS_FLATHER = 3 };
enum : unsigned { NX_SKIP = 0, NX_AVG = 1, NX_HI = 2, NX_LO = 3 };

__host__ __device__ inline unsigned ucls(unsigned f) { return (f >> U_CLS_SHIFT) & 3u; }
__host__ __device__ inline unsigned vcls(unsigned f) { return (f >> V_CLS_SHIFT) & 3u; }
__host__ __device__ inline unsigned sucls(unsigned f) { return (f >> SU_SHIFT) & 3u; }
__host__ __device__ inline unsigned svcls(unsigned f) { return (f >> SV_SHIFT) & 3u; }

struct Params {
    double rdt, cbfr, visc;
    double fric;  // 1 + cbfr*rdt, as the serial model forms it (:514)
};

// Scalar statics, used when every element of a metric/depth/Coriolis array
// holds the same value (the serial model's jphgr_msh=1 grid).
struct Scal {
    double e1t, e2t, e1u, e2u, e1v, e2v, e12t, e12u, e12v, ht, hu, hv, fcu, fcv;
};

struct StUniform {
    Scal s;
    __device__ double e1t(long) const { return s.e1t; }
    __device__ double e2t(long) const { return s.e2t; }
    __device__ double e1u(long) const { return s.e1u; }
    __device__ double e2u(long) const { return s.e2u; }
    __device__ double e1v(long) const { return s.e1v; }
    __device__ double e2v(long) const { return s.e2v; }
    __device__ double e12t(long) const { return s.e12t; }
    __device__ double e12u(long) const { return s.e12u; }
    __device__ double e12v(long) const { return s.e12v; }
    __device__ double ht(long) const { return s.ht; }
    __device__ double hu(long) const { return s.hu; }
    __device__ double hv(long) const { return s.hv; }
    __device__ double fcu(long) const { return s.fcu; }
    __device__ double fcv(long) const { return s.fcv; }
};

// Per-point statics (any grid).  fcu/fcv hold 2*omega*sin(gphi*d2r).
struct StArrays {
    const double *a_e1t, *a_e2t, *a_e1u, *a_e2u, *a_e1v, *a_e2v;
    const double *a_e12t, *a_e12u, *a_e12v, *a_ht, *a_hu, *a_hv, *a_fcu, *a_fcv;
    __device__ double e1t(long o) const { return a_e1t[o]; }
    __device__ double e2t(long o) const { return a_e2t[o]; }
    __device__ double e1u(long o) const { return a_e1u[o]; }
    __device__ double e2u(long o) const { return a_e2u[o]; }
    __device__ double e1v(long o) const { return a_e1v[o]; }
    __device__ double e2v(long o) const { return a_e2v[o]; }
    __device__ double e12t(long o) const { return a_e12t[o]; }
    __device__ double e12u(long o) const { return a_e12u[o]; }
    __device__ double e12v(long o) const { return a_e12v[o]; }
    __device__ double ht(long o) const { return a_ht[o]; }
    __device__ double hu(long o) const { return a_hu[o]; }
    __device__ double hv(long o) const { return a_hv[o]; }
    __device__ double fcu(long o) const { return a_fcu[o]; }
    __device__ double fcv(long o) const { return a_fcv[o]; }
};

// The now-state a step reads.
struct NState {
    const double *__restrict__ sshn, *__restrict__ sshn_u, *__restrict__ sshn_v;
    const double *__restrict__ un, *__restrict__ vn;
};

// continuity at T point o, original/nemolite2d.f90:399-403.
template <class S>
__device__ __forceinline__ double continuity_at(const S& st, const NState& n, long o, long P,
                                                double rdt) {
    const double r1 = (n.sshn_u[o] + st.hu(o)) * n.un[o];
    const double r2 = (n.sshn_u[o - 1] + st.hu(o - 1)) * n.un[o - 1];
    const double r3 = (n.sshn_v[o] + st.hv(o)) * n.vn[o];
    const double r4 = (n.sshn_v[o - P] + st.hv(o - P)) * n.vn[o - P];
    return n.sshn[o] + (r2 - r1 + r4 - r3) * rdt / st.e12t(o);
}

// Upwind weight SIGN(0.5, x) with Fortran's signed-zero rule.
__device__ __forceinline__ double sgn_half(double x) { return copysign(0.5, x); }

// momentum_u at U point o (flags f), original/nemolite2d.f90:439-514.
template <class S>
__device__ __forceinline__ double momentum_u_at(const S& st, const NState& n, long o, long P,
                                                unsigned f, const Params& p) {
    const double un_c = n.un[o], un_e = n.un[o + 1], un_w = n.un[o - 1];
    const double u_e = 0.5 * (un_c + un_e) * st.e2t(o + 1);
    const double depe = st.ht(o + 1) + n.sshn[o + 1];
    const double u_w = 0.5 * (un_c + un_w) * st.e2t(o);
    const double depw = st.ht(o) + n.sshn[o];
    const double v_sc = 0.5 * (n.vn[o - P] + n.vn[o - P + 1]);
    const double v_s = 0.5 * v_sc * (st.e1v(o - P) + st.e1v(o - P + 1));
    const double deps = 0.5 * (st.hv(o - P) + n.sshn_v[o - P] + st.hv(o - P + 1) + n.sshn_v[o - P + 1]);
    const double v_nc = 0.5 * (n.vn[o] + n.vn[o + 1]);
    const double v_n = 0.5 * v_nc * (st.e1v(o) + st.e1v(o + 1));
    const double depn = 0.5 * (st.hv(o) + n.sshn_v[o] + st.hv(o + 1) + n.sshn_v[o + 1]);

    const double uu_w = (0.5 - sgn_half(u_w)) * un_c + (0.5 + sgn_half(u_w)) * un_w;
    const double uu_e = (0.5 + sgn_half(u_e)) * un_c + (0.5 - sgn_half(u_e)) * un_e;
    const bool ws = (f & U_WALL_S) != 0, wn = (f & U_WALL_N) != 0;
    const double un_s = n.un[o - P], un_n = n.un[o + P];
    double uu_s = (0.5 - sgn_half(v_s)) * un_c;
    if (!ws) uu_s = uu_s + (0.5 + sgn_half(v_s)) * un_s;
    double uu_n = (0.5 + sgn_half(v_n)) * un_c;
    if (!wn) uu_n = uu_n + (0.5 - sgn_half(v_n)) * un_n;
    const double adv = uu_w * u_w * depw - uu_e * u_e * depe + uu_s * v_s * deps - uu_n * v_n * depn;

    const double dudx_e = (un_e - un_c) / st.e1t(o + 1) * (st.ht(o + 1) + n.sshn[o + 1]);
    const double dudx_w = (un_c - un_w) / st.e1t(o) * (st.ht(o) + n.sshn[o]);
    const double hsu = st.hu(o) + n.sshn_u[o];
    double dudy_s = 0.0, dudy_n = 0.0;
    if (!ws)
        dudy_s = (un_c - un_s) / (st.e2u(o) + st.e2u(o - P)) *
                 (st.hu(o) + n.sshn_u[o] + st.hu(o - P) + n.sshn_u[o - P]);
    if (!wn)
        dudy_n = (un_n - un_c) / (st.e2u(o) + st.e2u(o + P)) *
                 (st.hu(o) + n.sshn_u[o] + st.hu(o + P) + n.sshn_u[o + P]);
    double vis = (dudx_e - dudx_w) * st.e2u(o) + (dudy_n - dudy_s) * st.e1u(o) * 0.5;
    vis = p.visc * vis;

    const double cor = 0.5 * (st.fcu(o) * (v_sc + v_nc)) * st.e12u(o) * hsu;
    const double hpg = -GRAV * hsu * st.e2u(o) * (n.sshn[o + 1] - n.sshn[o]);
    // ssha_u is never written by the serial model (:302, :315): it is 0.
    return (un_c * hsu + p.rdt * (adv + vis + cor + hpg) / st.e12u(o)) / (st.hu(o) + 0.0) / p.fric;
}

// momentum_v at V point o (flags f), original/nemolite2d.f90:527-606.
template <class S>
__device__ __forceinline__ double momentum_v_at(const S& st, const NState& n, long o, long P,
                                                unsigned f, const Params& p) {
    const double vn_c = n.vn[o], vn_n = n.vn[o + P], vn_s = n.vn[o - P];
    const double v_n = 0.5 * (vn_c + vn_n) * st.e1t(o + P);
    const double depn = st.ht(o + P) + n.sshn[o + P];
    const double v_s = 0.5 * (vn_c + vn_s) * st.e1t(o);
    const double deps = st.ht(o) + n.sshn[o];
    const double u_wc = 0.5 * (n.un[o - 1] + n.un[o - 1 + P]);
    const double u_w = 0.5 * u_wc * (st.e2u(o - 1) + st.e2u(o - 1 + P));
    const double depw = 0.50 * (st.hu(o - 1) + n.sshn_u[o - 1] + st.hu(o - 1 + P) + n.sshn_u[o - 1 + P]);
    const double u_ec = 0.5 * (n.un[o] + n.un[o + P]);
    const double u_e = 0.5 * u_ec * (st.e2u(o) + st.e2u(o + P));
    const double depe = 0.50 * (st.hu(o) + n.sshn_u[o] + st.hu(o + P) + n.sshn_u[o + P]);

    const double vv_s = (0.5 - sgn_half(v_s)) * vn_c + (0.5 + sgn_half(v_s)) * vn_s;
    const double vv_n = (0.5 + sgn_half(v_n)) * vn_c + (0.5 - sgn_half(v_n)) * vn_n;
    const bool ww = (f & V_WALL_W) != 0, we = (f & V_WALL_E) != 0;
    const double vn_w = n.vn[o - 1], vn_e = n.vn[o + 1];
    double vv_w = (0.5 - sgn_half(u_w)) * vn_c;
    if (!ww) vv_w = vv_w + (0.5 + sgn_half(u_w)) * vn_w;
    double vv_e = (0.5 + sgn_half(u_e)) * vn_c;
    if (!we) vv_e = vv_e + (0.5 - sgn_half(u_e)) * vn_e;
    const double adv = vv_w * u_w * depw - vv_e * u_e * depe + vv_s * v_s * deps - vv_n * v_n * depn;

    const double dvdy_n = (vn_n - vn_c) / st.e2t(o + P) * (st.ht(o + P) + n.sshn[o + P]);
    const double dvdy_s = (vn_c - vn_s) / st.e2t(o) * (st.ht(o) + n.sshn[o]);
    const double hsv = st.hv(o) + n.sshn_v[o];
    double dvdx_w = 0.0, dvdx_e = 0.0;
    if (!ww)
        dvdx_w = (vn_c - vn_w) / (st.e1v(o) + st.e1v(o - 1)) *
                 (st.hv(o) + n.sshn_v[o] + st.hv(o - 1) + n.sshn_v[o - 1]);
    if (!we)
        dvdx_e = (vn_e - vn_c) / (st.e1v(o) + st.e1v(o + 1)) *
                 (st.hv(o) + n.sshn_v[o] + st.hv(o + 1) + n.sshn_v[o + 1]);
    double vis = (dvdy_n - dvdy_s) * st.e1v(o) + (dvdx_e - dvdx_w) * st.e2v(o) * 0.5;
    vis = p.visc * vis;

    const double cor = -0.5 * (st.fcv(o) * (u_ec + u_wc)) * st.e12v(o) * hsv;
    const double hpg = -GRAV * hsv * st.e1v(o) * (n.sshn[o + P] - n.sshn[o]);
    return (vn_c * hsv + p.rdt * (adv + vis + cor + hpg) / st.e12v(o)) / (st.hv(o) + 0.0) / p.fric;
}

// next: sshn_u from the new ssh at T points o (i) and o+1 (i+1),
// original/nemolite2d.f90:738-746.  `keep` is the old value (skip class).
template <class S>
__device__ __forceinline__ double next_su(const S& st, unsigned cls, long o, double s_lo,
                                          double s_hi, double keep) {
    switch (cls) {
        case NX_AVG: {
            const double r = st.e12t(o) * s_lo + st.e12t(o + 1) * s_hi;
            return 0.5 * r / st.e12u(o);
        }
        case NX_HI: return s_hi;
        case NX_LO: return s_lo;
        default: return keep;
    }
}

// next: sshn_v from the new ssh at T points o (j) and o+P (j+1), :754-762.
template <class S>
__device__ __forceinline__ double next_sv(const S& st, unsigned cls, long o, long P, double s_lo,
                                          double s_hi, double keep) {
    switch (cls) {
        case NX_AVG: {
            const double r = st.e12t(o) * s_lo + st.e12t(o + P) * s_hi;
            return 0.5 * r / st.e12v(o);
        }
        case NX_HI: return s_hi;
        case NX_LO: return s_lo;
        default: return keep;
    }
}

// ---- compensated, order-fixed reductions (checksums) -----------------------
// Double-double accumulator built on Knuth's TwoSum; with FMA contraction off
// the error term is exact, so block and grid reductions in a fixed tree give
// run-to-run identical sums that are within about 1e-30 relative of the
// exact sum for the sizes used here.
struct DD {
    double s, c;
};
__device__ __forceinline__ void dd_add(DD& a, double x) {
    const double t = a.s + x;
    const double bp = t - a.s;
    const double e = (a.s - (t - bp)) + (x - bp);
    a.s = t;
    a.c += e;
}
__device__ __forceinline__ DD dd_merge(DD a, DD b) {
    const double t = a.s + b.s;
    const double bp = t - a.s;
    const double e = (a.s - (t - bp)) + (b.s - bp);
    DD r;
    r.s = t;
    r.c = (a.c + b.c) + e;
    return r;
}
__device__ __forceinline__ DD dd_shfl_xor(DD a, int m) {
    DD r;
    r.s = __shfl_xor(a.s, m, 64);
    r.c = __shfl_xor(a.c, m, 64);
    return r;
}

// Reduce three DD accumulators over a block of nthreads (multiple of 64,
// at most 1024) and write them (s, c pairs) to out[0..5] from thread 0.
__device__ __forceinline__ void block_reduce3(DD a0, DD a1, DD a2, double* out) {
    __shared__ DD red[3][16];
    for (int m = 32; m >= 1; m >>= 1) {
        a0 = dd_merge(a0, dd_shfl_xor(a0, m));
        a1 = dd_merge(a1, dd_shfl_xor(a1, m));
        a2 = dd_merge(a2, dd_shfl_xor(a2, m));
    }
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int nthr = blockDim.x * blockDim.y;
    const int wave = tid >> 6, lane = tid & 63, nw = nthr >> 6;
    if (lane == 0) {
        red[0][wave] = a0;
        red[1][wave] = a1;
        red[2][wave] = a2;
    }
    __syncthreads();
    if (tid == 0) {
        DD r0 = red[0][0], r1 = red[1][0], r2 = red[2][0];
        for (int w = 1; w < nw; ++w) {
            r0 = dd_merge(r0, red[0][w]);
            r1 = dd_merge(r1, red[1][w]);
            r2 = dd_merge(r2, red[2][w]);
        }
        out[0] = r0.s; out[1] = r0.c;
        out[2] = r1.s; out[3] = r1.c;
        out[4] = r2.s; out[5] = r2.c;
    }
}

}  // namespace nemo2d
