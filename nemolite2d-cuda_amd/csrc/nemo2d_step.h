// Per-cell bodies of the fused step (FP64 and single-precision forms),
// shared by the single-model kernels (nemo2d.hip) and the ensemble kernel
// (nemo2d_ensemble.hip).
#pragma once

#include <type_traits>

#include "nemo2d_core.h"

namespace nemo2d {

constexpr int TBX = 64, TBY = 4;   // thread-per-cell tiles: one wave per row segment
constexpr int RED_THREADS = 1024;  // one-block reductions

// The after-state a fused step writes (the other half of the ping-pong).
struct AOut {
    double *sshn, *sshn_u, *sshn_v, *un, *vn;
};

__device__ __forceinline__ bool in_rect(int i, int j, int i0, int i1, int j0, int j1) {
    return i >= i0 && i <= i1 && j >= j0 && j <= j1;
}

// Pre-Flather value of the after-velocity at U point o, for a point that is
// the inner neighbour of a Flather point (never itself a Flather point when
// the fused path is used): momentum if computed, else 0 (solid points are
// zeroed, untouched points keep the zero of the allocation).
template <class S>
__device__ __forceinline__ double u_pre_flather(const S& st, const NState& n, const uint16_t* fl,
                                                long o, long P, const Params& p) {
    const unsigned f = fl[o];
    return ucls(f) == CLS_COMP ? momentum_u_at(st, n, o, P, f, p) : 0.0;
}
template <class S>
__device__ __forceinline__ double v_pre_flather(const S& st, const NState& n, const uint16_t* fl,
                                                long o, long P, const Params& p) {
    const unsigned f = fl[o];
    return vcls(f) == CLS_COMP ? momentum_v_at(st, n, o, P, f, p) : 0.0;
}

// ssha at T point o after bc_ssh: the tide value where clamped.
template <class S>
__device__ __forceinline__ double ssha_at(const S& st, const NState& n, const uint16_t* fl, long o,
                                          long P, double rdt, double tide) {
    return (fl[o] & T_CLAMP) ? tide : continuity_at(st, n, o, P, rdt);
}

// ---- fused step, one thread per cell ------------------------------------------
// One thread per cell (i, j), i in [0, jpi], j in [0, jpj]: it produces the
// T point (i, j), the U point (i, j) (east face), the V point (i, j) (north
// face) and sshn_u/sshn_v there.  The new ssh east and north of the cell,
// needed by the interpolation of next, is recomputed from the now-state.
// Flather points are left to k_epilogue unless FLIN (then the inner
// neighbour's momentum is recomputed here).
template <class S, bool FLIN>
__device__ __forceinline__ void cell_step(const S& st, const NState& n, const AOut& a,
                                          const uint16_t* __restrict__ fl,
                                          const double* __restrict__ flu,
                                          const double* __restrict__ flv, const Layout& L,
                                          const Params& p, double tide, int i, int j, DD& ct,
                                          DD& cu, DD& cv) {
    const long P = L.pitch;
    if (i <= L.jpi && j <= L.jpj) {
        const long o = L.at(i, j);
        const unsigned f = fl[o];
        double s_c = 0.0;
        if (i >= 1 && j >= 1) {
            s_c = (f & T_CLAMP) ? tide : continuity_at(st, n, o, P, p.rdt);
            a.sshn[o] = s_c;
            dd_add(ct, fabs(s_c));
        }
        if (j >= 1) {
            // U point (i, j)
            const unsigned c = ucls(f);
            if (c != CLS_FLATHER || FLIN) {
                double u = 0.0;
                if (c == CLS_COMP) {
                    u = momentum_u_at(st, n, o, P, f, p);
                } else if (c == CLS_FLATHER) {
                    const long oi = (f & U_FL_EAST) ? o + 1 : o - 1;
                    u = u_pre_flather(st, n, fl, oi, P, p) + flu[o] * (n.sshn_u[o] - n.sshn_u[oi]);
                }
                a.un[o] = u;
                if (i >= 1) dd_add(cu, fabs(u));
            }
            const unsigned su = sucls(f);
            double s_hi = 0.0;
            if (su == NX_AVG || su == NX_HI) s_hi = ssha_at(st, n, fl, o + 1, P, p.rdt, tide);
            a.sshn_u[o] = next_su(st, su, o, s_c, s_hi, n.sshn_u[o]);
        }
        if (i >= 1) {
            // V point (i, j)
            const unsigned c = vcls(f);
            if (c != CLS_FLATHER || FLIN) {
                double v = 0.0;
                if (c == CLS_COMP) {
                    v = momentum_v_at(st, n, o, P, f, p);
                } else if (c == CLS_FLATHER) {
                    const long oi = (f & V_FL_NORTH) ? o + P : o - P;
                    v = v_pre_flather(st, n, fl, oi, P, p) + flv[o] * (n.sshn_v[o] - n.sshn_v[oi]);
                }
                a.vn[o] = v;
                if (j >= 1) dd_add(cv, fabs(v));
            }
            const unsigned sv = svcls(f);
            double s_hi = 0.0;
            if (sv == NX_AVG || sv == NX_HI) s_hi = ssha_at(st, n, fl, o + P, P, p.rdt, tide);
            a.sshn_v[o] = next_sv(st, sv, o, P, s_c, s_hi, n.sshn_v[o]);
        }
    }
}

// ---- single-precision form ------------------------------------------------------
// FP32 state and stencil arithmetic; the new ssh is accumulated as
// sshn + incr in FP64 when `COMP` (sshn kept in double), in FP32 otherwise
// (the ablation).  Uniform metrics and Coriolis parameter, per-point depths.
struct ScalF {
    float e1t, e2t, e1u, e2u, e1v, e2v, e12t, e12u, e12v, fcu, fcv;
    float ht, hu, hv;  // used when the depths are uniform (UNI)
};

template <bool COMP>
struct MixedState {
    using ssh_t = typename std::conditional<COMP, double, float>::type;
    const ssh_t* __restrict__ sshn;
    const float *__restrict__ sshn_u, *__restrict__ sshn_v, *__restrict__ un, *__restrict__ vn;
};
template <bool COMP>
struct MixedOut {
    typename MixedState<COMP>::ssh_t* sshn;
    float *sshn_u, *sshn_v, *un, *vn;
};

template <bool COMP, bool UNI>
struct MemWinF {
    using real = float;
    const MixedState<COMP>& n;
    const ScalF& s;
    const float *ht_, *hu_, *hv_;
    long o, P;
    template <int DI, int DJ> __device__ long at() const { return o + DI + (long)DJ * P; }
    template <int DI, int DJ> __device__ float un() const { return n.un[at<DI, DJ>()]; }
    template <int DI, int DJ> __device__ float vn() const { return n.vn[at<DI, DJ>()]; }
    template <int DI, int DJ> __device__ float sshn() const { return (float)n.sshn[at<DI, DJ>()]; }
    template <int DI, int DJ> __device__ float sshn_u() const { return n.sshn_u[at<DI, DJ>()]; }
    template <int DI, int DJ> __device__ float sshn_v() const { return n.sshn_v[at<DI, DJ>()]; }
    template <int DI, int DJ> __device__ float ht() const { return UNI ? s.ht : ht_[at<DI, DJ>()]; }
    template <int DI, int DJ> __device__ float hu() const { return UNI ? s.hu : hu_[at<DI, DJ>()]; }
    template <int DI, int DJ> __device__ float hv() const { return UNI ? s.hv : hv_[at<DI, DJ>()]; }
#define NEMO2D_SF(F) \
    template <int DI, int DJ> __device__ float F() const { return s.F; }
    NEMO2D_SF(e1t) NEMO2D_SF(e2t) NEMO2D_SF(e1u) NEMO2D_SF(e2u) NEMO2D_SF(e1v) NEMO2D_SF(e2v)
    NEMO2D_SF(e12t) NEMO2D_SF(e12u) NEMO2D_SF(e12v) NEMO2D_SF(fcu) NEMO2D_SF(fcv)
#undef NEMO2D_SF
    template <int TAG> __device__ float qd(float x, float d) const { return dv(x, d); }
};

template <bool COMP, bool UNI>
__device__ __forceinline__ typename MixedState<COMP>::ssh_t mixed_ssha(
    const MemWinF<COMP, UNI>& w, unsigned f, float rdt, double tide) {
    using ssh_t = typename MixedState<COMP>::ssh_t;
    if (f & T_CLAMP) return (ssh_t)tide;
    return w.n.sshn[w.o] + (ssh_t)continuity_incr_w(w, rdt);
}

template <bool COMP, bool UNI>
__device__ __forceinline__ void cell_step_mixed(const ScalF& s, const MixedState<COMP>& n,
                                                const MixedOut<COMP>& a, const float* __restrict__ ht,
                                                const float* __restrict__ hu,
                                                const float* __restrict__ hv,
                                                const uint16_t* __restrict__ fl,
                                                const double* __restrict__ flu,
                                                const double* __restrict__ flv, const Layout& L,
                                                const ParamsT<float>& p, double tide, int i, int j,
                                                DD& ct, DD& cu, DD& cv) {
    const long P = L.pitch;
    if (i <= L.jpi && j <= L.jpj) {
        const long o = L.at(i, j);
        const unsigned f = fl[o];
        MemWinF<COMP, UNI> w{n, s, ht, hu, hv, o, P};
        double s_c = 0.0;
        if (i >= 1 && j >= 1) {
            const auto sa = mixed_ssha(w, f, p.rdt, tide);
            a.sshn[o] = sa;
            s_c = (double)sa;
            dd_add(ct, fabs(s_c));
        }
        if (j >= 1) {
            float u;
            const unsigned c = ucls(f);
            if (c == CLS_COMP) {
                u = momentum_u_w(w, f, p);
            } else if (c == CLS_FLATHER) {
                const long oi = (f & U_FL_EAST) ? o + 1 : o - 1;
                const unsigned fi = fl[oi];
                MemWinF<COMP, UNI> wi{n, s, ht, hu, hv, oi, P};
                const float ui = ucls(fi) == CLS_COMP ? momentum_u_w(wi, fi, p) : 0.0f;
                u = ui + (float)flu[o] * (n.sshn_u[o] - n.sshn_u[oi]);
            } else {
                u = 0.0f;
            }
            a.un[o] = u;
            if (i >= 1) dd_add(cu, fabs((double)u));
            const unsigned su = sucls(f);
            float r = n.sshn_u[o];
            if (su != NX_SKIP) {
                float s_hi = 0.0f;
                if (su == NX_AVG || su == NX_HI) {
                    const long oe = o + 1;
                    MemWinF<COMP, UNI> we_{n, s, ht, hu, hv, oe, P};
                    s_hi = (float)mixed_ssha(we_, fl[oe], p.rdt, tide);
                }
                const float s_lo = (float)s_c;
                r = su == NX_AVG ? dv(0.5f * (s.e12t * s_lo + s.e12t * s_hi), s.e12u)
                                 : (su == NX_HI ? s_hi : s_lo);
            }
            a.sshn_u[o] = r;
        }
        if (i >= 1) {
            float v;
            const unsigned c = vcls(f);
            if (c == CLS_COMP) {
                v = momentum_v_w(w, f, p);
            } else if (c == CLS_FLATHER) {
                const long oi = (f & V_FL_NORTH) ? o + P : o - P;
                const unsigned fi = fl[oi];
                MemWinF<COMP, UNI> wi{n, s, ht, hu, hv, oi, P};
                const float vi = vcls(fi) == CLS_COMP ? momentum_v_w(wi, fi, p) : 0.0f;
                v = vi + (float)flv[o] * (n.sshn_v[o] - n.sshn_v[oi]);
            } else {
                v = 0.0f;
            }
            a.vn[o] = v;
            if (j >= 1) dd_add(cv, fabs((double)v));
            const unsigned sv = svcls(f);
            float r = n.sshn_v[o];
            if (sv != NX_SKIP) {
                float s_hi = 0.0f;
                if (sv == NX_AVG || sv == NX_HI) {
                    const long on = o + P;
                    MemWinF<COMP, UNI> wn_{n, s, ht, hu, hv, on, P};
                    s_hi = (float)mixed_ssha(wn_, fl[on], p.rdt, tide);
                }
                const float s_lo = (float)s_c;
                r = sv == NX_AVG ? dv(0.5f * (s.e12t * s_lo + s.e12t * s_hi), s.e12v)
                                 : (sv == NX_HI ? s_hi : s_lo);
            }
            a.sshn_v[o] = r;
        }
    }
}

// Fixed-order reduction of nparts block partials into out[0..2] (one block).
__device__ __forceinline__ void finish_partials(const double* __restrict__ partials, int nparts,
                                                double* __restrict__ out) {
    DD a0 = {0.0, 0.0}, a1 = {0.0, 0.0}, a2 = {0.0, 0.0};
    // four partials per thread per pass, all loads issued before the merges
    const int nt = blockDim.x;
    for (int k0 = threadIdx.x; k0 < nparts; k0 += 4 * nt) {
        double q[4][6];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * nt;
#pragma unroll
            for (int c = 0; c < 6; ++c) q[u][c] = k < nparts ? partials[6 * (long)k + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            a0 = dd_merge(a0, DD{q[u][0], q[u][1]});
            a1 = dd_merge(a1, DD{q[u][2], q[u][3]});
            a2 = dd_merge(a2, DD{q[u][4], q[u][5]});
        }
    }
    __shared__ double r[6];
    block_reduce3(a0, a1, a2, r);
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = r[0] + r[1];
        out[1] = r[2] + r[3];
        out[2] = r[4] + r[5];
    }
}


}  // namespace nemo2d
