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
    // columns: guard (i = -1) .. i = jpi+2, with i = 0 at COL0; pitch a
    // multiple of 16 doubles (128 B); rows: guard (j = -1) .. j = jpj+2
    static Layout of(int jpi, int jpj) {
        Layout L{};
        L.jpi = jpi;
        L.jpj = jpj;
        L.pitch = ((long)COL0 + jpi + 3 + 15) / 16 * 16;
        L.rows = (long)jpj + 4;
        L.base = L.pitch + COL0;
        return L;
    }
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

template <class T>
struct ParamsT {
    T rdt, cbfr, visc;
    T fric;  // 1 + cbfr*rdt, as the serial model forms it (:514)
};
using Params = ParamsT<double>;

// Divisor tags of the point functions: a window with uniform statics divides
// by precomputed reciprocals, any other window divides.
enum {
    QD_E12T = 0, QD_E1T, QD_E2T, QD_E2U2, QD_E1V2, QD_E12U, QD_E12V, QD_HU, QD_HV, QD_FRIC, QD_N
};

// Scalar statics, used when every element of a metric/depth/Coriolis array
// holds the same value (the serial model's jphgr_msh=1 grid).
struct Scal {
    double e1t, e2t, e1u, e2u, e1v, e2v, e12t, e12u, e12v, ht, hu, hv, fcu, fcv;
    double rcp[QD_N];  // RN(1/divisor) for each QD_ tag (uniform statics)
    int fast_div;      // divisors in range: divide by reciprocals
    int one_corr;      // and every |divisor * rcp - 1| <= 2^-54: one correction suffices
    int one_corr_met;  // the same for the metric divisors alone (all but hu, hv)
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

// Stencil windows.  The point functions read every operand through a
// window, w.F<DI,DJ>() = field F at (i+DI, j+DJ), so the same arithmetic runs
// on values loaded from memory (MemWin) or held in registers by a marching
// kernel (a register window with the same interface).
template <class S>
struct MemWin {
    using real = double;
    const S& st;
    const NState& n;
    long o, P;
    template <int DI, int DJ> __device__ long at() const { return o + DI + (long)DJ * P; }
#define NEMO2D_MEMF(F) \
    template <int DI, int DJ> __device__ double F() const { return n.F[at<DI, DJ>()]; }
    NEMO2D_MEMF(un) NEMO2D_MEMF(vn) NEMO2D_MEMF(sshn) NEMO2D_MEMF(sshn_u) NEMO2D_MEMF(sshn_v)
#undef NEMO2D_MEMF
#define NEMO2D_MEMS(F) \
    template <int DI, int DJ> __device__ double F() const { return st.F(at<DI, DJ>()); }
    NEMO2D_MEMS(e1t) NEMO2D_MEMS(e2t) NEMO2D_MEMS(e1u) NEMO2D_MEMS(e2u) NEMO2D_MEMS(e1v)
    NEMO2D_MEMS(e2v) NEMO2D_MEMS(e12t) NEMO2D_MEMS(e12u) NEMO2D_MEMS(e12v) NEMO2D_MEMS(ht)
    NEMO2D_MEMS(hu) NEMO2D_MEMS(hv) NEMO2D_MEMS(fcu) NEMO2D_MEMS(fcv)
#undef NEMO2D_MEMS
    template <int TAG> __device__ double qd(double x, double d) const { return x / d; }
};

#define W_(F, DI, DJ) w.template F<DI, DJ>()
#define QD_(TAG, X, D) w.template qd<TAG>((X), (D))

// Division used by the point functions: IEEE for double (bitwise parity);
// the single-precision form multiplies by the hardware reciprocal
// (v_rcp_f32, 1 ulp), a quotient within ~2 ulp of the FP32 one: one
// multiply per division where the divisor is a uniform metric (the
// reciprocal is computed once), a reciprocal and a multiply where it is a
// per-point depth.  (A correctly rounded FP32 division is ~9 instructions.)
__device__ __forceinline__ double dv(double a, double b) { return a / b; }
__device__ __forceinline__ float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Correctly rounded x / d for a divisor d > 0 in [2^-60, 2^60] with y =
// RN(1/d) precomputed: q0 = RN(x*y) and two FMA remainder corrections
// (Markstein: y within half an ulp of 1/d and a faithful quotient make
// RN(q + (x - d*q)*y) the correctly rounded quotient; the first correction
// makes the quotient faithful, the remainder x - d*q is exact by FMA while
// nothing underflows).  The quotient takes x's sign, which keeps the sign of
// a zero x.  |x| outside [2^-960, 2^960) other than zero (where a remainder
// could underflow), and non-finite x, take the IEEE division; the test is on
// the biased exponent, 63 <= e <= 1982.  Same bits as x / d for every x.
__device__ __forceinline__ double div_rcp(double x, double d, double y) {
    double q = x * y;
    double r = fma(-d, q, x);
    q = fma(r, y, q);
    r = fma(-d, q, x);
    q = copysign(fma(r, y, q), x);
    const unsigned e = ((unsigned)__double2hiint(x) >> 20) & 0x7ffu;
    if (__builtin_expect(e - 63u >= 1920u && x != 0.0, 0)) q = x / d;
    return q;
}


// The same quotient without the per-division guard, for code that
// accumulates the guard over many divisions and redoes them all with the
// IEEE division when it fails: the corrections are written q - (d*q - x)*y
// (the same rounded values as q + (x - d*q)*y), which keeps the sign of a
// zero x without a copysign.  DivCheck gathers the frexp exponents of the
// numerators and whether one was NaN, infinite or subnormal; the quotients
// are x / d's bits when div_check_ok holds (every numerator zero or in
// [2^-960, 2^960), exponents -959..960).
struct DivCheck {
    int emin, emax;
    bool odd;
};
__device__ __forceinline__ void div_check_reset(DivCheck& c) {
    c.emin = 0;
    c.emax = 0;
    c.odd = false;
}
__device__ __forceinline__ double div_rcp_acc(double x, double d, double y, DivCheck& c) {
    double q = x * y;
    double r = fma(d, q, -x);
    q = fma(-r, y, q);
    r = fma(d, q, -x);
    q = fma(-r, y, q);
    const int e = __builtin_amdgcn_frexp_exp(x);
    c.emin = e < c.emin ? e : c.emin;
    c.emax = e > c.emax ? e : c.emax;
    c.odd = c.odd || __builtin_amdgcn_class(x, 0x297);  // NaN, infinity, subnormal
    return q;
}
// The one-correction form: RN(x*y) is already a faithful quotient when
// |d*y - 1| <= 2^-54 (then |x*y - x/d| < ulp(x/d)/2, and rounding adds at
// most another half ulp), so one remainder correction gives the correctly
// rounded quotient (Markstein), with the same guard on x.  The host allows
// it only for grids whose every divisor passes that test (Scal::one_corr).
__device__ __forceinline__ double div_rcp1_acc(double x, double d, double y, DivCheck& c) {
    double q = x * y;
    const double r = fma(d, q, -x);
    q = fma(-r, y, q);
    const int e = __builtin_amdgcn_frexp_exp(x);
    c.emin = e < c.emin ? e : c.emin;
    c.emax = e > c.emax ? e : c.emax;
    c.odd = c.odd || __builtin_amdgcn_class(x, 0x297);  // NaN, infinity, subnormal
    return q;
}
__device__ __forceinline__ bool div_check_ok(const DivCheck& c) {
    return !c.odd && c.emin >= -959 && c.emax <= 960;
}

// Value of x held by lane l+1 of the wave, by a DPP wavefront shift (one
// v_mov per 32-bit half, no LDS round trip).  Lane 63 receives 0 (bound
// control on, every lane written, so no initial value is needed).  Every lane
// of the wave must be active.
__device__ __forceinline__ float lane_next(float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ double lane_next(double x) {
    int2 v = __builtin_bit_cast(int2, x);
    v.x = __builtin_amdgcn_update_dpp(0, v.x, 0x130, 0xf, 0xf, true);
    v.y = __builtin_amdgcn_update_dpp(0, v.y, 0x130, 0xf, 0xf, true);
    return __builtin_bit_cast(double, v);
}
// Value of x held by lane l-1 (DPP wavefront shift the other way); lane 0
// receives 0.  Every lane of the wave must be active.
__device__ __forceinline__ float lane_prev(float x) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ double lane_prev(double x) {
    int2 v = __builtin_bit_cast(int2, x);
    v.x = __builtin_amdgcn_update_dpp(0, v.x, 0x138, 0xf, 0xf, true);
    v.y = __builtin_amdgcn_update_dpp(0, v.y, 0x138, 0xf, 0xf, true);
    return __builtin_bit_cast(double, v);
}

// Upwind weight SIGN(0.5, x) with Fortran's signed-zero rule.
__device__ __forceinline__ double sgn_half(double x) { return copysign(0.5, x); }
__device__ __forceinline__ float sgn_half(float x) { return copysignf(0.5f, x); }

// continuity increment at T point (i, j): ssha - sshn, original/nemolite2d.f90:399-403
// (ssha = sshn + incr is the same expression tree as :403).
template <class W, class T = typename W::real>
__device__ __forceinline__ T continuity_incr_w(const W& w, T rdt) {
    const T r1 = (W_(sshn_u, 0, 0) + W_(hu, 0, 0)) * W_(un, 0, 0);
    const T r2 = (W_(sshn_u, -1, 0) + W_(hu, -1, 0)) * W_(un, -1, 0);
    const T r3 = (W_(sshn_v, 0, 0) + W_(hv, 0, 0)) * W_(vn, 0, 0);
    const T r4 = (W_(sshn_v, 0, -1) + W_(hv, 0, -1)) * W_(vn, 0, -1);
    return QD_(QD_E12T, (r2 - r1 + r4 - r3) * rdt, W_(e12t, 0, 0));
}
template <class W>
__device__ __forceinline__ double continuity_w(const W& w, double rdt) {
    return W_(sshn, 0, 0) + continuity_incr_w(w, rdt);
}

// momentum_u at U point (i, j) with flags f, original/nemolite2d.f90:439-514.
template <class W, class T = typename W::real>
__device__ __forceinline__ T momentum_u_w(const W& w, unsigned f, const ParamsT<T>& p) {
    const T h = T(0.5);
    const T un_c = W_(un, 0, 0), un_e = W_(un, 1, 0), un_w = W_(un, -1, 0);
    const T u_e = h * (un_c + un_e) * W_(e2t, 1, 0);
    const T depe = W_(ht, 1, 0) + W_(sshn, 1, 0);
    const T u_w = h * (un_c + un_w) * W_(e2t, 0, 0);
    const T depw = W_(ht, 0, 0) + W_(sshn, 0, 0);
    const T v_sc = h * (W_(vn, 0, -1) + W_(vn, 1, -1));
    const T v_s = h * v_sc * (W_(e1v, 0, -1) + W_(e1v, 1, -1));
    const T deps = h * (W_(hv, 0, -1) + W_(sshn_v, 0, -1) + W_(hv, 1, -1) + W_(sshn_v, 1, -1));
    const T v_nc = h * (W_(vn, 0, 0) + W_(vn, 1, 0));
    const T v_n = h * v_nc * (W_(e1v, 0, 0) + W_(e1v, 1, 0));
    const T depn = h * (W_(hv, 0, 0) + W_(sshn_v, 0, 0) + W_(hv, 1, 0) + W_(sshn_v, 1, 0));

    const T uu_w = (h - sgn_half(u_w)) * un_c + (h + sgn_half(u_w)) * un_w;
    const T uu_e = (h + sgn_half(u_e)) * un_c + (h - sgn_half(u_e)) * un_e;
    const bool ws = (f & U_WALL_S) != 0, wn = (f & U_WALL_N) != 0;
    const T un_s = W_(un, 0, -1), un_n = W_(un, 0, 1);
    T uu_s = (h - sgn_half(v_s)) * un_c;
    if (!ws) uu_s = uu_s + (h + sgn_half(v_s)) * un_s;
    T uu_n = (h + sgn_half(v_n)) * un_c;
    if (!wn) uu_n = uu_n + (h - sgn_half(v_n)) * un_n;
    const T adv = uu_w * u_w * depw - uu_e * u_e * depe + uu_s * v_s * deps - uu_n * v_n * depn;

    const T dudx_e = QD_(QD_E1T, un_e - un_c, W_(e1t, 1, 0)) * (W_(ht, 1, 0) + W_(sshn, 1, 0));
    const T dudx_w = QD_(QD_E1T, un_c - un_w, W_(e1t, 0, 0)) * (W_(ht, 0, 0) + W_(sshn, 0, 0));
    const T hsu = W_(hu, 0, 0) + W_(sshn_u, 0, 0);
    T dudy_s = T(0.0), dudy_n = T(0.0);
    if (!ws)
        dudy_s = QD_(QD_E2U2, un_c - un_s, W_(e2u, 0, 0) + W_(e2u, 0, -1)) *
                 (W_(hu, 0, 0) + W_(sshn_u, 0, 0) + W_(hu, 0, -1) + W_(sshn_u, 0, -1));
    if (!wn)
        dudy_n = QD_(QD_E2U2, un_n - un_c, W_(e2u, 0, 0) + W_(e2u, 0, 1)) *
                 (W_(hu, 0, 0) + W_(sshn_u, 0, 0) + W_(hu, 0, 1) + W_(sshn_u, 0, 1));
    T vis = (dudx_e - dudx_w) * W_(e2u, 0, 0) + (dudy_n - dudy_s) * W_(e1u, 0, 0) * h;
    vis = p.visc * vis;

    const T cor = h * (W_(fcu, 0, 0) * (v_sc + v_nc)) * W_(e12u, 0, 0) * hsu;
    const T hpg = -T(GRAV) * hsu * W_(e2u, 0, 0) * (W_(sshn, 1, 0) - W_(sshn, 0, 0));
    // ssha_u is never written by the serial model (:302, :315): it is 0.
    return QD_(QD_FRIC,
               QD_(QD_HU, un_c * hsu + QD_(QD_E12U, p.rdt * (adv + vis + cor + hpg), W_(e12u, 0, 0)),
                   W_(hu, 0, 0) + T(0.0)),
               p.fric);
}

// momentum_v at V point (i, j) with flags f, original/nemolite2d.f90:527-606.
template <class W, class T = typename W::real>
__device__ __forceinline__ T momentum_v_w(const W& w, unsigned f, const ParamsT<T>& p) {
    const T h = T(0.5);
    const T vn_c = W_(vn, 0, 0), vn_n = W_(vn, 0, 1), vn_s = W_(vn, 0, -1);
    const T v_n = h * (vn_c + vn_n) * W_(e1t, 0, 1);
    const T depn = W_(ht, 0, 1) + W_(sshn, 0, 1);
    const T v_s = h * (vn_c + vn_s) * W_(e1t, 0, 0);
    const T deps = W_(ht, 0, 0) + W_(sshn, 0, 0);
    const T u_wc = h * (W_(un, -1, 0) + W_(un, -1, 1));
    const T u_w = h * u_wc * (W_(e2u, -1, 0) + W_(e2u, -1, 1));
    const T depw = h * (W_(hu, -1, 0) + W_(sshn_u, -1, 0) + W_(hu, -1, 1) + W_(sshn_u, -1, 1));
    const T u_ec = h * (W_(un, 0, 0) + W_(un, 0, 1));
    const T u_e = h * u_ec * (W_(e2u, 0, 0) + W_(e2u, 0, 1));
    const T depe = h * (W_(hu, 0, 0) + W_(sshn_u, 0, 0) + W_(hu, 0, 1) + W_(sshn_u, 0, 1));

    const T vv_s = (h - sgn_half(v_s)) * vn_c + (h + sgn_half(v_s)) * vn_s;
    const T vv_n = (h + sgn_half(v_n)) * vn_c + (h - sgn_half(v_n)) * vn_n;
    const bool ww = (f & V_WALL_W) != 0, we = (f & V_WALL_E) != 0;
    const T vn_w = W_(vn, -1, 0), vn_e = W_(vn, 1, 0);
    T vv_w = (h - sgn_half(u_w)) * vn_c;
    if (!ww) vv_w = vv_w + (h + sgn_half(u_w)) * vn_w;
    T vv_e = (h + sgn_half(u_e)) * vn_c;
    if (!we) vv_e = vv_e + (h - sgn_half(u_e)) * vn_e;
    const T adv = vv_w * u_w * depw - vv_e * u_e * depe + vv_s * v_s * deps - vv_n * v_n * depn;

    const T dvdy_n = QD_(QD_E2T, vn_n - vn_c, W_(e2t, 0, 1)) * (W_(ht, 0, 1) + W_(sshn, 0, 1));
    const T dvdy_s = QD_(QD_E2T, vn_c - vn_s, W_(e2t, 0, 0)) * (W_(ht, 0, 0) + W_(sshn, 0, 0));
    const T hsv = W_(hv, 0, 0) + W_(sshn_v, 0, 0);
    T dvdx_w = T(0.0), dvdx_e = T(0.0);
    if (!ww)
        dvdx_w = QD_(QD_E1V2, vn_c - vn_w, W_(e1v, 0, 0) + W_(e1v, -1, 0)) *
                 (W_(hv, 0, 0) + W_(sshn_v, 0, 0) + W_(hv, -1, 0) + W_(sshn_v, -1, 0));
    if (!we)
        dvdx_e = QD_(QD_E1V2, vn_e - vn_c, W_(e1v, 0, 0) + W_(e1v, 1, 0)) *
                 (W_(hv, 0, 0) + W_(sshn_v, 0, 0) + W_(hv, 1, 0) + W_(sshn_v, 1, 0));
    T vis = (dvdy_n - dvdy_s) * W_(e1v, 0, 0) + (dvdx_e - dvdx_w) * W_(e2v, 0, 0) * h;
    vis = p.visc * vis;

    const T cor = -h * (W_(fcv, 0, 0) * (u_ec + u_wc)) * W_(e12v, 0, 0) * hsv;
    const T hpg = -T(GRAV) * hsv * W_(e1v, 0, 0) * (W_(sshn, 0, 1) - W_(sshn, 0, 0));
    return QD_(QD_FRIC,
               QD_(QD_HV, vn_c * hsv + QD_(QD_E12V, p.rdt * (adv + vis + cor + hpg), W_(e12v, 0, 0)),
                   W_(hv, 0, 0) + T(0.0)),
               p.fric);
}

// Memory-window wrappers at offset o.
template <class S>
__device__ __forceinline__ double continuity_at(const S& st, const NState& n, long o, long P,
                                                double rdt) {
    return continuity_w(MemWin<S>{st, n, o, P}, rdt);
}
template <class S>
__device__ __forceinline__ double momentum_u_at(const S& st, const NState& n, long o, long P,
                                                unsigned f, const Params& p) {
    return momentum_u_w(MemWin<S>{st, n, o, P}, f, p);
}
template <class S>
__device__ __forceinline__ double momentum_v_at(const S& st, const NState& n, long o, long P,
                                                unsigned f, const Params& p) {
    return momentum_v_w(MemWin<S>{st, n, o, P}, f, p);
}

// next: sshn_u from the new ssh at T points (i, j) and (i+1, j),
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

// next: sshn_v from the new ssh at T points (i, j) and (i, j+1), :754-762.
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

// FP32 -> BF16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

}  // namespace nemo2d
