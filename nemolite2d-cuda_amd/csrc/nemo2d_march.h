// Column-marching fused step with shared face terms, generic over the
// arithmetic type R (double for the drop-in, float for the single-precision
// form), the storage type SS of sshn (double, or float for the
// uncompensated single-precision form) and the source of the depths
// (uniform scalars or per-point arrays); metrics and Coriolis parameters
// are uniform.  The arithmetic of every point is the reference's
// (original/nemolite2d.f90:399-403 continuity, :439-514 momentum_u,
// :527-606 momentum_v, :735-765 next, :676-690 bc_flather_v), evaluated as
// the per-point kernels of this library evaluate it, so the results equal
// theirs bit for bit:
//  * a term on a face shared by two points of a column is computed once at
//    the lower point and carried up a row (continuity's v flux, the
//    momentum_u south advective flux, v average and du/dy quotient,
//    momentum_v's south advective flux and dv/dy);
//  * a term on a face shared by two points of a row is computed by the
//    lane of the east point and shuffled to the west lane (momentum_u's
//    east advective flux and du/dx, momentum_v's east advective flux, u
//    average and dv/dx quotient);
//  * every shared term is the same expression tree as the point's own (an
//    IEEE sum a+b equals b+a); a walled point recomputes its own advective
//    term;
//  * rows are held in a ring of three whose roles rotate, so no row is
//    copied, and the arrays are addressed by 32-bit offsets from the kernel
//    arguments.
#pragma once

#include <hip/hip_runtime.h>

#include "nemo2d_core.h"
#include "nemo2d_step.h"

namespace nemo2d {

constexpr int GW = 4;      // waves per block (column strips side by side)
constexpr int GOUT = 63;   // output columns per wave (lane 63 is the east halo)

// uniform metrics, Coriolis parameters and (when not per point) depths
template <class R>
struct MetG {
    R e1t, e2t, e1u, e2u, e1v, e2v, e12t, e12u, e12v, fcu, fcv, ht, hu, hv;
};

template <class R, class SS, bool DARR>
struct StateG {
    const SS* sshn;
    const R *sshn_u, *sshn_v, *un, *vn;
    const R *ht, *hu, *hv;  // per-point depths (DARR)
};
template <class R, class SS>
struct OutG {
    SS* sshn;
    R *sshn_u, *sshn_v, *un, *vn;
};

// element k of an array addressed by a 32-bit element offset
template <class T>
__device__ __forceinline__ T ldg_(const T* b, unsigned k) {
    return *(const T*)((const char*)b + k * (unsigned)sizeof(T));
}
template <class T>
__device__ __forceinline__ void stg_(T* b, unsigned k, T v) {
    *(T*)((char*)b + k * (unsigned)sizeof(T)) = v;
}

// One row of the now-state at columns i-1, i, i+1.
template <class R, class SS, bool DARR>
struct RowG {
    R un[3], vn[3], su[3], sv[3];
    SS ss[3];
    R ht_[DARR ? 3 : 1], hu_[DARR ? 3 : 1], hv_[DARR ? 3 : 1];
};

template <class R, class SS, bool DARR>
__device__ __forceinline__ void load_rowg(RowG<R, SS, DARR>& r, const StateG<R, SS, DARR>& n,
                                          unsigned k) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.un[c] = ldg_(n.un, k + c);
        r.vn[c] = ldg_(n.vn, k + c);
        r.su[c] = ldg_(n.sshn_u, k + c);
        r.sv[c] = ldg_(n.sshn_v, k + c);
        r.ss[c] = ldg_(n.sshn, k + c);
        if constexpr (DARR) {
            r.ht_[c] = ldg_(n.ht, k + c);
            r.hu_[c] = ldg_(n.hu, k + c);
            r.hv_[c] = ldg_(n.hv, k + c);
        }
    }
}

template <class R, class SS, bool DARR>
struct DepG {
    const MetG<R>& m;
    __device__ R ht(const RowG<R, SS, DARR>& r, int c) const { return DARR ? r.ht_[DARR ? c : 0] : m.ht; }
    __device__ R hu(const RowG<R, SS, DARR>& r, int c) const { return DARR ? r.hu_[DARR ? c : 0] : m.hu; }
    __device__ R hv(const RowG<R, SS, DARR>& r, int c) const { return DARR ? r.hv_[DARR ? c : 0] : m.hv; }
};

// Terms on the north faces of the points of row c (row nr above it).
template <class R>
struct NorthG {
    R r3;     // continuity: (sshn_v + hv) * vn
    R kn;     // momentum_u: uu_n * v_n * depn, no wall
    R vnc;    // momentum_u: v_nc
    R vnu;    // momentum_u: v_n
    R depnu;  // momentum_u: depn
    R hn;     // momentum_v: vv_n * v_n * depn
    R dvdyn;  // momentum_v: dvdy_n
    R qdudy;  // momentum_u: dudy_n's quotient (un_n - un_c) / (e2u + e2u)
};

template <class R, class SS, bool DARR>
__device__ __forceinline__ NorthG<R> north_g(const MetG<R>& m, const DepG<R, SS, DARR>& d,
                                             const RowG<R, SS, DARR>& c,
                                             const RowG<R, SS, DARR>& nr) {
    const R h = R(0.5);
    NorthG<R> t;
    t.r3 = (c.sv[1] + d.hv(c, 1)) * c.vn[1];
    t.vnc = h * (c.vn[1] + c.vn[2]);
    t.vnu = h * t.vnc * (m.e1v + m.e1v);
    t.depnu = h * (d.hv(c, 1) + c.sv[1] + d.hv(c, 2) + c.sv[2]);
    const R sn = sgn_half(t.vnu);
    t.kn = ((h + sn) * c.un[1] + (h - sn) * nr.un[1]) * t.vnu * t.depnu;
    const R vn_n = h * (c.vn[1] + nr.vn[1]) * m.e1t;
    const R depn = d.ht(nr, 1) + R(nr.ss[1]);
    const R sv = sgn_half(vn_n);
    t.hn = ((h + sv) * c.vn[1] + (h - sv) * nr.vn[1]) * vn_n * depn;
    t.dvdyn = dv(nr.vn[1] - c.vn[1], m.e2t) * (d.ht(nr, 1) + R(nr.ss[1]));
    t.qdudy = dv(nr.un[1] - c.un[1], m.e2u + m.e2u);
    return t;
}

// next (original/nemolite2d.f90:735-765) for uniform areas
template <class R>
__device__ __forceinline__ R next_g(unsigned cls, R lo, R hi, R keep, R e12t, R e12x) {
    switch (cls) {
        case NX_AVG: return dv(R(0.5) * (e12t * lo + e12t * hi), e12x);  // (IEEE for double)
        case NX_HI: return hi;
        case NX_LO: return lo;
        default: return keep;
    }
}

// Reciprocals RN(1/d) of the uniform-metric divisors (QD_ tags) for the
// one-correction division of the FP64 sshn-only kernel (DM 3: every metric
// divisor d of the grid has |d*RN(1/d) - 1| <= 2^-54; the per-point depths
// hu, hv divide by IEEE division).
struct RcpG {
    double r[QD_N];
};
// x / d for the uniform-metric divisor d of tag T: DM 0 dv (IEEE for
// double), DM 3 div_rcp1_acc with its guard gathered in *ck (double only).
template <int DM, int T, class R>
__device__ __forceinline__ R gq(const RcpG& rc, R x, R d, DivCheck* ck) {
    if constexpr (DM == 3) return div_rcp1_acc(x, d, rc.r[T], *ck);
    else return dv(x, d);
}
// next_g with the average's division by gq
template <int DM, int T, class R>
__device__ __forceinline__ R next_gq(const RcpG& rc, unsigned cls, R lo, R hi, R keep, R e12t, R e12x,
                                     DivCheck* ck) {
    switch (cls) {
        case NX_AVG: return gq<DM, T>(rc, R(0.5) * (e12t * lo + e12t * hi), e12x, ck);
        case NX_HI: return hi;
        case NX_LO: return lo;
        default: return keep;
    }
}

template <class R, class SS, bool DARR>
struct CtxG {
    MetG<R> m;
    ParamsT<R> p;
    StateG<R, SS, DARR> n;
    OutG<R, SS> a;
    const uint16_t* fl;
    const double* flv;  // Flather sqrt(g/hv) (row-0 points)
    Layout L;
    double tide;
    int i, il;
    unsigned lk;  // element offset of the window's west column within a row
    bool own, tcol;
    __device__ unsigned rowk(int j) const { return (unsigned)L.at(0, j); }
};

template <class R, class SS>
struct CarryG {
    NorthG<R> south;  // the row below's north terms
    SS s_prev;        // new ssh at (i, j-1)
    R sv_prev;        // old sshn_v at (i, j-1)
    unsigned f_prev, f_nxt;
    double ct, cu, cv;  // per-lane checksum sums
};

// One row j: A, B hold rows j-1, j; C is loaded with row j+1 here.
template <class R, class SS, bool DARR>
__device__ __forceinline__ void row_g(const CtxG<R, SS, DARR>& x, CarryG<R, SS>& k,
                                      const RowG<R, SS, DARR>& A, const RowG<R, SS, DARR>& B,
                                      RowG<R, SS, DARR>& C, int j) {
    const Layout& L = x.L;
    const MetG<R>& m = x.m;
    const DepG<R, SS, DARR> d{m};
    const R h = R(0.5);
    const unsigned f = k.f_nxt;
    load_rowg(C, x.n, x.rowk(j + 1) + x.lk);
    {
        const int jf = j + 1 <= L.jpj + 1 ? j + 1 : L.jpj + 1;
        k.f_nxt = ldg_(x.fl, x.rowk(jf) + x.lk + 1);
    }
    const unsigned ko = x.rowk(j) + x.lk + 1;  // element offset of (i, j)
    const unsigned kP = (unsigned)L.pitch;
    const NorthG<R> nt = north_g(m, d, B, C);
    const NorthG<R>& st = k.south;

    // continuity
    SS s_c = SS(0);
    if (j >= 1 && x.tcol) {
        if (f & T_CLAMP) {
            s_c = SS(x.tide);
        } else {
            const R r1 = (B.su[1] + d.hu(B, 1)) * B.un[1];
            const R r2 = (B.su[0] + d.hu(B, 0)) * B.un[0];
            s_c = B.ss[1] + SS(dv((r2 - r1 + st.r3 - nt.r3) * x.p.rdt, m.e12t));
        }
    }
    const SS s_e = lane_next(s_c);

    // west-face terms; the east ones come from lane+1
    const R un_c = B.un[1];
    const R u_w = h * (un_c + B.un[0]) * m.e2t;
    const R depw_u = d.ht(B, 1) + R(B.ss[1]);
    const R sw_u = sgn_half(u_w);
    const R fw = ((h - sw_u) * un_c + (h + sw_u) * B.un[0]) * u_w * depw_u;
    const R dudx_w = dv(un_c - B.un[0], m.e1t) * (d.ht(B, 1) + R(B.ss[1]));
    const R vn_c = B.vn[1];
    const R u_wc = h * (B.un[0] + C.un[0]);
    const R u_wv = h * u_wc * (m.e2u + m.e2u);
    const R depw_v = h * (d.hu(B, 0) + B.su[0] + d.hu(C, 0) + C.su[0]);
    const R sw_v = sgn_half(u_wv);
    const R gw = ((h - sw_v) * vn_c + (h + sw_v) * B.vn[0]) * u_wv * depw_v;
    const R qdvdx_w = dv(vn_c - B.vn[0], m.e1v + m.e1v);
    const R fe = lane_next(fw);
    const R dudx_e = lane_next(dudx_w);
    const R ge = lane_next(gw);
    const R u_ec = lane_next(u_wc);
    const R qdvdx_e = lane_next(qdvdx_w);

    const R su_old = B.su[1], sv_old = B.sv[1];
    if (x.own) {
        const int i = x.i;
        R u = R(0), v = R(0);
        const unsigned cu = ucls(f), cv = vcls(f);
        if (j >= 1 && cu == CLS_COMP) {
            const bool ws = (f & U_WALL_S) != 0, wn = (f & U_WALL_N) != 0;
            R ts = st.kn, tn = nt.kn;
            if (ws) ts = ((h - sgn_half(st.vnu)) * un_c) * st.vnu * st.depnu;
            if (wn) tn = ((h + sgn_half(nt.vnu)) * un_c) * nt.vnu * nt.depnu;
            const R adv = fw - fe + ts - tn;
            R dudy_s = R(0), dudy_n = R(0);
            if (!ws) dudy_s = st.qdudy * (d.hu(B, 1) + B.su[1] + d.hu(A, 1) + A.su[1]);
            if (!wn) dudy_n = nt.qdudy * (d.hu(B, 1) + B.su[1] + d.hu(C, 1) + C.su[1]);
            R vis = (dudx_e - dudx_w) * m.e2u + (dudy_n - dudy_s) * m.e1u * h;
            vis = x.p.visc * vis;
            const R hsu = d.hu(B, 1) + B.su[1];
            const R cor = h * (m.fcu * (st.vnc + nt.vnc)) * m.e12u * hsu;
            const R hpg = -R(GRAV) * hsu * m.e2u * (R(B.ss[2]) - R(B.ss[1]));
            u = dv(dv(un_c * hsu + dv(x.p.rdt * (adv + vis + cor + hpg), m.e12u), d.hu(B, 1) + R(0)),
                   x.p.fric);
        }
        if (i >= 1 && cv == CLS_COMP) {
            const bool ww = (f & V_WALL_W) != 0, we = (f & V_WALL_E) != 0;
            R tw = gw, te = ge;
            if (ww) tw = ((h - sw_v) * vn_c) * u_wv * depw_v;
            if (we) {
                const R u_e = h * u_ec * (m.e2u + m.e2u);
                const R depe = h * (d.hu(B, 1) + B.su[1] + d.hu(C, 1) + C.su[1]);
                te = ((h + sgn_half(u_e)) * vn_c) * u_e * depe;
            }
            const R adv = tw - te + st.hn - nt.hn;
            R dvdx_w = R(0), dvdx_e = R(0);
            if (!ww) dvdx_w = qdvdx_w * (d.hv(B, 1) + B.sv[1] + d.hv(B, 0) + B.sv[0]);
            if (!we) dvdx_e = qdvdx_e * (d.hv(B, 1) + B.sv[1] + d.hv(B, 2) + B.sv[2]);
            R vis = (nt.dvdyn - st.dvdyn) * m.e1v + (dvdx_e - dvdx_w) * m.e2v * h;
            vis = x.p.visc * vis;
            const R hsv = d.hv(B, 1) + B.sv[1];
            const R cor = -h * (m.fcv * (u_ec + u_wc)) * m.e12v * hsv;
            const R hpg = -R(GRAV) * hsv * m.e1v * (R(C.ss[1]) - R(B.ss[1]));
            v = dv(dv(vn_c * hsv + dv(x.p.rdt * (adv + vis + cor + hpg), m.e12v), d.hv(B, 1) + R(0)),
                   x.p.fric);
        }
        if (j >= 1) {
            if (i >= 1) {
                stg_(x.a.sshn, ko, s_c);
                k.ct += fabs((double)s_c);
            }
            if (cu != CLS_FLATHER) {
                stg_(x.a.un, ko, u);
                if (i >= 1) k.cu += fabs((double)u);
            }
            stg_(x.a.sshn_u, ko, next_g<R>(sucls(f), R(s_c), R(s_e), su_old, m.e12t, m.e12u));
        }
        if (i >= 1) {
            if (cv != CLS_FLATHER) {
                stg_(x.a.vn, ko, v);
                if (j >= 1) k.cv += fabs((double)v);
            }
            // bc_flather_v at a row-0 point whose inner point is this one
            // (row 0 is outside the checksums)
            if (j == 1 && vcls(k.f_prev) == CLS_FLATHER && (k.f_prev & V_FL_NORTH))
                stg_(x.a.vn, ko - kP, v + R(ldg_(x.flv, ko - kP)) * (A.sv[1] - B.sv[1]));
            if (k.f_prev != 0xffffffffu)
                stg_(x.a.sshn_v, ko - kP,
                     next_g<R>(svcls(k.f_prev), R(k.s_prev), R(s_c), k.sv_prev, m.e12t, m.e12v));
        }
    }
    k.south = nt;
    k.s_prev = s_c;
    k.sv_prev = sv_old;
    k.f_prev = f;
}

// continuity at (i, j) from memory, for the segment's last sshn_v
template <class R, class SS, bool DARR>
__device__ __forceinline__ SS ssha_at_g(const CtxG<R, SS, DARR>& x, unsigned k) {
    const StateG<R, SS, DARR>& n = x.n;
    const unsigned kP = (unsigned)x.L.pitch;
    const R hu_c = DARR ? ldg_(n.hu, k) : x.m.hu, hu_w = DARR ? ldg_(n.hu, k - 1) : x.m.hu;
    const R hv_c = DARR ? ldg_(n.hv, k) : x.m.hv, hv_s = DARR ? ldg_(n.hv, k - kP) : x.m.hv;
    const R r1 = (ldg_(n.sshn_u, k) + hu_c) * ldg_(n.un, k);
    const R r2 = (ldg_(n.sshn_u, k - 1) + hu_w) * ldg_(n.un, k - 1);
    const R r3 = (ldg_(n.sshn_v, k) + hv_c) * ldg_(n.vn, k);
    const R r4 = (ldg_(n.sshn_v, k - kP) + hv_s) * ldg_(n.vn, k - kP);
    return ldg_(n.sshn, k) + SS(dv((r2 - r1 + r4 - r3) * x.p.rdt, x.m.e12t));
}

// One wave: column strip `strip` (columns strip*GOUT .. +GOUT-1; lane 63
// is the east halo) over the NJ-row segment `seg`, adding its checksum
// terms to k.ct/cu/cv.
template <class R, class SS, bool DARR, int NJ>
__device__ __forceinline__ void march_wave_g(const MetG<R>& m, const ParamsT<R>& p,
                                             const StateG<R, SS, DARR>& n, const OutG<R, SS>& a,
                                             const uint16_t* __restrict__ fl,
                                             const double* __restrict__ flv, const Layout& L,
                                             double tide, int strip, int seg, int lane,
                                             CarryG<R, SS>& k) {
    CtxG<R, SS, DARR> x;
    x.m = m;
    x.p = p;
    x.n = n;
    x.a = a;
    x.fl = fl;
    x.flv = flv;
    x.L = L;
    x.tide = tide;
    x.i = strip * GOUT + lane;
    x.il = x.i < L.jpi + 1 ? x.i : L.jpi + 1;  // clamped column for loads
    x.lk = (unsigned)(x.il - 1);
    x.own = lane < GOUT && x.i <= L.jpi;
    x.tcol = x.i >= 1 && x.i <= L.jpi;
    const int jb = seg * NJ;
    const int je = jb + NJ - 1 < L.jpj ? jb + NJ - 1 : L.jpj;
    const DepG<R, SS, DARR> d{m};
    RowG<R, SS, DARR> R0, R1, R2;
    load_rowg(R0, n, x.rowk(jb - 1) + x.lk);
    load_rowg(R1, n, x.rowk(jb) + x.lk);
    k.f_nxt = ldg_(fl, x.rowk(jb) + x.lk + 1);
    k.f_prev = 0xffffffffu;  // no row below inside this segment
    k.s_prev = SS(0);
    k.sv_prev = R(0);
    k.south = north_g(m, d, R0, R1);
    for (int j = jb; j <= je; j += 3) {
        row_g(x, k, R0, R1, R2, j);
        if (j + 1 > je) break;
        row_g(x, k, R1, R2, R0, j + 1);
        if (j + 2 > je) break;
        row_g(x, k, R2, R0, R1, j + 2);
    }
    // the last row's sshn_v needs the new ssh one row up
    if (x.own && x.i >= 1) {
        const unsigned ko = x.rowk(je) + x.lk + 1, kP = (unsigned)L.pitch;
        const unsigned c = svcls(k.f_prev);
        SS s_n = SS(0);
        if (je + 1 <= L.jpj && (c == NX_AVG || c == NX_HI))
            s_n = (ldg_(fl, ko + kP) & T_CLAMP) ? SS(tide) : ssha_at_g(x, ko + kP);
        stg_(a.sshn_v, ko, next_g<R>(c, R(k.s_prev), R(s_n), k.sv_prev, m.e12t, m.e12v));
    }
}

// Block (GW waves) = GW column strips of GOUT columns, one NJ-row segment;
// tiles are handed out XCD by XCD (block b runs on XCD b % 8).
template <class R, class SS, bool DARR, int NJ, int WPE = 0>
__global__ void __launch_bounds__(64 * GW) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
k_step_march_g(MetG<R> m, ParamsT<R> p, StateG<R, SS, DARR> n, OutG<R, SS> a,
               const uint16_t* __restrict__ fl, const double* __restrict__ flv, Layout L,
               double tide, double* __restrict__ partials, int nbx, int ntiles) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int per = gridDim.x >> 3;  // gridDim.x is a multiple of 8
    const int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    CarryG<R, SS> k;
    k.ct = k.cu = k.cv = 0.0;
    if (t < ntiles)
        march_wave_g<R, SS, DARR, NJ>(m, p, n, a, fl, flv, L, tide, (t % nbx) * GW + wave, t / nbx,
                                      lane, k);
    block_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)blockIdx.x);
}

// ---- the same step reading sshn, un, vn (and the depths) only ----------------
// As k_step_march_rc of the FP64 drop-in, generic like the kernel above:
// sshn_u and sshn_v of the now-state are formed from sshn with next's own
// expression in R arithmetic (next_g of R(sshn) values, as next formed the
// stored ones), the stored value is read only where next leaves a point
// alone, and only sshn, un and vn are written; the library fills the new
// state's sshn_u/sshn_v in (the same interpolation) before anything reads
// them.  Each lane loads its own column; neighbours come by DPP lane
// shifts; lanes 1..GROUT own the columns (lane 0 and lanes GROUT+1..63 are
// halos); the ring of LA+3 rows keeps the next LA rows' loads in flight;
// with PLAIN (a tile whose every point has the plain interior class word)
// the class tests drop out.
constexpr int GROUT = 61;
constexpr unsigned G_PLAIN = (CLS_COMP << U_CLS_SHIFT) | (CLS_COMP << V_CLS_SHIFT) |
                             (NX_AVG << SU_SHIFT) | (NX_AVG << SV_SHIFT);

template <class R, class SS>
struct ColG {
    R un, vn, su, sv;
    SS ss;
    R ht, hu, hv;  // per-point depths (DARR)
    unsigned f;
};

// DV (with DARR): on plain tiles hu and hv are not loaded but formed from
// ht as the means of the adjacent T depths (derived_depths)
template <class R, class SS, bool DARR, bool PLAIN, bool DV = false>
__device__ __forceinline__ void load_colg(ColG<R, SS>& r, const StateG<R, SS, DARR>& n,
                                          const uint16_t* fl, unsigned k) {
    r.un = ldg_(n.un, k);
    r.vn = ldg_(n.vn, k);
    r.ss = ldg_(n.sshn, k);
    if constexpr (DARR) {
        r.ht = ldg_(n.ht, k);
        if constexpr (!(PLAIN && DV)) {
            r.hu = ldg_(n.hu, k);
            r.hv = ldg_(n.hv, k);
        }
    }
    if constexpr (PLAIN) r.f = G_PLAIN;
    else r.f = ldg_(fl, k);
}

// the stored value where next leaves a point alone (wait inside the branch)
template <class R>
__device__ __forceinline__ R keep_g(R v, bool keep, const R* arr, unsigned k) {
    if (keep) {
        v = ldg_(arr, k);
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0), other counters untouched
    }
    return v;
}

template <class R, class SS, bool DARR>
struct CtxGR : CtxG<R, SS, DARR> {
    bool ucol, vcol;
    RcpG rc;  // (DM 3)
    __device__ unsigned ucls_at(unsigned f, int r) const {
        return (ucol && r >= 1 && r <= this->L.jpj) ? sucls(f) : NX_SKIP;
    }
    __device__ unsigned vcls_at(unsigned f, int r) const {
        return (vcol && r >= 0 && r <= this->L.jpj) ? svcls(f) : NX_SKIP;
    }
};

template <class R>
struct CarryGR {
    NorthG<R> south;
    R un_w, su_w, hu_w;  // un, sshn_u, hu one column west, this row
    double ct, cu, cv;
};

template <class R, class SS, bool DARR, bool PLAIN, int DM = 0>
__device__ __forceinline__ R su_at_g(const CtxGR<R, SS, DARR>& x, const ColG<R, SS>& c, SS ss_e,
                                     int r, unsigned k, DivCheck* ck = nullptr) {
    const unsigned cls = PLAIN ? NX_AVG : x.ucls_at(c.f, r);
    const R v = next_gq<DM, QD_E12U, R>(x.rc, cls == NX_SKIP ? NX_AVG : cls, R(c.ss), R(ss_e), R(0), x.m.e12t,
                                        x.m.e12u, ck);
    return PLAIN ? v : keep_g(v, cls == NX_SKIP, x.n.sshn_u, k);
}
template <class R, class SS, bool DARR, bool PLAIN, int DM = 0>
__device__ __forceinline__ R sv_at_g(const CtxGR<R, SS, DARR>& x, const ColG<R, SS>& c, SS ss_n,
                                     int r, unsigned k, DivCheck* ck = nullptr) {
    const unsigned cls = PLAIN ? NX_AVG : x.vcls_at(c.f, r);
    const R v = next_gq<DM, QD_E12V, R>(x.rc, cls == NX_SKIP ? NX_AVG : cls, R(c.ss), R(ss_n), R(0), x.m.e12t,
                                        x.m.e12v, ck);
    return PLAIN ? v : keep_g(v, cls == NX_SKIP, x.n.sshn_v, k);
}

// north terms of row c (own column; vn, sshn_v and hv one column east) with
// row nr above
template <class R, class SS, bool DARR, int DM = 0>
__device__ __forceinline__ NorthG<R> north_gc(const MetG<R>& m, const ColG<R, SS>& c, R vn_e,
                                              R sv_e, R hv_e, const ColG<R, SS>& nr,
                                              const RcpG* rc = nullptr, DivCheck* ck = nullptr) {
    const R h = R(0.5);
    const R hv_c = DARR ? c.hv : m.hv, ht_n = DARR ? nr.ht : m.ht;
    NorthG<R> t;
    t.r3 = (c.sv + hv_c) * c.vn;
    t.vnc = h * (c.vn + vn_e);
    t.vnu = h * t.vnc * (m.e1v + m.e1v);
    t.depnu = h * (hv_c + c.sv + (DARR ? hv_e : m.hv) + sv_e);
    const R sn = sgn_half(t.vnu);
    t.kn = ((h + sn) * c.un + (h - sn) * nr.un) * t.vnu * t.depnu;
    const R vn_n = h * (c.vn + nr.vn) * m.e1t;
    const R depn = ht_n + R(nr.ss);
    const R sv = sgn_half(vn_n);
    t.hn = ((h + sv) * c.vn + (h - sv) * nr.vn) * vn_n * depn;
    if constexpr (DM == 3) {
        t.dvdyn = gq<DM, QD_E2T>(*rc, nr.vn - c.vn, m.e2t, ck) * (ht_n + R(nr.ss));
        t.qdudy = gq<DM, QD_E2U2>(*rc, nr.un - c.un, m.e2u + m.e2u, ck);
    } else {
        t.dvdyn = dv(nr.vn - c.vn, m.e2t) * (ht_n + R(nr.ss));
        t.qdudy = dv(nr.un - c.un, m.e2u + m.e2u);
    }
    return t;
}

// Depths at the U point (i, j) and V point (i, j) of a plain tile whose hu
// and hv are the means of adjacent T depths: 0.5 * (ht(i) + ht(i+1)) and
// 0.5 * (ht(j) + ht(j+1)), the expressions the grid was checked against.
template <class R, class SS>
__device__ __forceinline__ R hu_of(const ColG<R, SS>& c) {
    return R(0.5) * (c.ht + lane_next(c.ht));
}
template <class R, class SS>
__device__ __forceinline__ R hv_of(const ColG<R, SS>& c, const ColG<R, SS>& nr) {
    return R(0.5) * (c.ht + nr.ht);
}

// What row j computes: C's sshn_u and B's sshn_v (the now-state's
// interpolations first needed here), the row's north terms, the west values
// carried to the next row, and the new sshn, u and v.
template <class R, class SS>
struct RcgOut {
    NorthG<R> nt;
    R csu, bsv, cun_w, csu_w, chu_w;
    SS s_c;
    R u, v;
};

// Row j's arithmetic, no stores: A, B, C hold rows j-1, j, j+1 (C just
// arrived; with DARR, PLAIN and DV its hu and B's hv are already formed).
// DM 3: the uniform-metric divisions by reciprocal, their guard gathered in
// *ck; DM 0: dv.
template <class R, class SS, bool DARR, bool PLAIN, int DM>
__device__ __forceinline__ RcgOut<R, SS> rcg_compute(const CtxGR<R, SS, DARR>& x, const CarryGR<R>& k,
                                                     const ColG<R, SS>& A, const ColG<R, SS>& B,
                                                     const ColG<R, SS>& C, int j, DivCheck* ck) {
    const Layout& L = x.L;
    const MetG<R>& m = x.m;
    const R h = R(0.5);
    const unsigned kP = (unsigned)L.pitch;
    const unsigned f = B.f;
    const unsigned ko = x.rowk(j) + x.lk;  // element offset of (i, j)
    RcgOut<R, SS> o;
    o.csu = su_at_g<R, SS, DARR, PLAIN, DM>(x, C, lane_next(C.ss), j + 1, ko + kP, ck);
    o.bsv = sv_at_g<R, SS, DARR, PLAIN, DM>(x, B, C.ss, j, ko, ck);
    const R csu = o.csu, bsv = o.bsv;
    // neighbours by lane shifts (every lane active here)
    const R bun_w = k.un_w, bsu_w = k.su_w;
    const R bvn_w = lane_prev(B.vn), bsv_w = lane_prev(bsv);
    const R bvn_e = lane_next(B.vn), bsv_e = lane_next(bsv);
    const SS bss_e = lane_next(B.ss);
    const R cun_w = lane_prev(C.un), csu_w = lane_prev(csu);
    const R hu_b = DARR ? B.hu : m.hu, hv_b = DARR ? B.hv : m.hv, ht_b = DARR ? B.ht : m.ht;
    const R hu_a = DARR ? A.hu : m.hu, hu_c = DARR ? C.hu : m.hu;
    const R bhu_w = DARR ? k.hu_w : m.hu;
    const R chu_w = DARR ? lane_prev(C.hu) : m.hu;
    const R bhv_w = DARR ? lane_prev(B.hv) : m.hv, bhv_e = DARR ? lane_next(B.hv) : m.hv;
    o.cun_w = cun_w;
    o.csu_w = csu_w;
    o.chu_w = chu_w;
    ColG<R, SS> Bn = B;  // B with its interpolation of this row
    Bn.sv = bsv;
    o.nt = north_gc<R, SS, DARR, DM>(m, Bn, bvn_e, bsv_e, bhv_e, C, &x.rc, ck);
    const NorthG<R>& nt = o.nt;
    const NorthG<R>& st = k.south;

    // continuity
    SS s_c = SS(0);
    if (PLAIN || (j >= 1 && x.tcol)) {
        if (!PLAIN && (f & T_CLAMP)) {
            s_c = SS(x.tide);
        } else {
            const R r1 = (B.su + hu_b) * B.un;
            const R r2 = (bsu_w + bhu_w) * bun_w;
            s_c = B.ss + SS(gq<DM, QD_E12T>(x.rc, (r2 - r1 + st.r3 - nt.r3) * x.p.rdt, m.e12t, ck));
        }
    }
    o.s_c = s_c;

    // west-face terms; the east ones come from lane+1
    const R un_c = B.un;
    const R u_w = h * (un_c + bun_w) * m.e2t;
    const R depw_u = ht_b + R(B.ss);
    const R sw_u = sgn_half(u_w);
    const R fw = ((h - sw_u) * un_c + (h + sw_u) * bun_w) * u_w * depw_u;
    const R dudx_w = gq<DM, QD_E1T>(x.rc, un_c - bun_w, m.e1t, ck) * (ht_b + R(B.ss));
    const R vn_c = B.vn;
    const R u_wc = h * (bun_w + cun_w);
    const R u_wv = h * u_wc * (m.e2u + m.e2u);
    const R depw_v = h * (bhu_w + bsu_w + chu_w + csu_w);
    const R sw_v = sgn_half(u_wv);
    const R gw = ((h - sw_v) * vn_c + (h + sw_v) * bvn_w) * u_wv * depw_v;
    const R qdvdx_w = gq<DM, QD_E1V2>(x.rc, vn_c - bvn_w, m.e1v + m.e1v, ck);
    const R fe = lane_next(fw);
    const R dudx_e = lane_next(dudx_w);
    const R ge = lane_next(gw);
    const R u_ec = lane_next(u_wc);
    const R qdvdx_e = lane_next(qdvdx_w);

    R u = R(0), v = R(0);
    if (x.own) {
        const int i = x.i;
        const unsigned cu = ucls(f), cv = vcls(f);
        if (PLAIN || (j >= 1 && cu == CLS_COMP)) {
            const bool ws = !PLAIN && (f & U_WALL_S) != 0, wn = !PLAIN && (f & U_WALL_N) != 0;
            R ts = st.kn, tn = nt.kn;
            if (ws) ts = ((h - sgn_half(st.vnu)) * un_c) * st.vnu * st.depnu;
            if (wn) tn = ((h + sgn_half(nt.vnu)) * un_c) * nt.vnu * nt.depnu;
            const R adv = fw - fe + ts - tn;
            R dudy_s = R(0), dudy_n = R(0);
            if (!ws) dudy_s = st.qdudy * (hu_b + B.su + hu_a + A.su);
            if (!wn) dudy_n = nt.qdudy * (hu_b + B.su + hu_c + csu);
            R vis = (dudx_e - dudx_w) * m.e2u + (dudy_n - dudy_s) * m.e1u * h;
            vis = x.p.visc * vis;
            const R hsu = hu_b + B.su;
            const R cor = h * (m.fcu * (st.vnc + nt.vnc)) * m.e12u * hsu;
            const R hpg = -R(GRAV) * hsu * m.e2u * (R(bss_e) - R(B.ss));
            u = gq<DM, QD_FRIC>(x.rc,
                                dv(un_c * hsu + gq<DM, QD_E12U>(x.rc, x.p.rdt * (adv + vis + cor + hpg), m.e12u, ck),
                                   hu_b + R(0)),
                                x.p.fric, ck);
        }
        if (PLAIN || (i >= 1 && cv == CLS_COMP)) {
            const bool ww = !PLAIN && (f & V_WALL_W) != 0, we = !PLAIN && (f & V_WALL_E) != 0;
            R tw = gw, te = ge;
            if (ww) tw = ((h - sw_v) * vn_c) * u_wv * depw_v;
            if (we) {
                const R u_e = h * u_ec * (m.e2u + m.e2u);
                const R depe = h * (hu_b + B.su + hu_c + csu);
                te = ((h + sgn_half(u_e)) * vn_c) * u_e * depe;
            }
            const R adv = tw - te + st.hn - nt.hn;
            R dvdx_w = R(0), dvdx_e = R(0);
            if (!ww) dvdx_w = qdvdx_w * (hv_b + bsv + bhv_w + bsv_w);
            if (!we) dvdx_e = qdvdx_e * (hv_b + bsv + bhv_e + bsv_e);
            R vis = (nt.dvdyn - st.dvdyn) * m.e1v + (dvdx_e - dvdx_w) * m.e2v * h;
            vis = x.p.visc * vis;
            const R hsv = hv_b + bsv;
            const R cor = -h * (m.fcv * (u_ec + u_wc)) * m.e12v * hsv;
            const R hpg = -R(GRAV) * hsv * m.e1v * (R(C.ss) - R(B.ss));
            v = gq<DM, QD_FRIC>(x.rc,
                                dv(vn_c * hsv + gq<DM, QD_E12V>(x.rc, x.p.rdt * (adv + vis + cor + hpg), m.e12v, ck),
                                   hv_b + R(0)),
                                x.p.fric, ck);
        }
    }
    o.u = u;
    o.v = v;
    return o;
}

// Row j's stores and carries.
template <class R, class SS, bool DARR, bool PLAIN>
__device__ __forceinline__ void rcg_commit(const CtxGR<R, SS, DARR>& x, CarryGR<R>& k, const ColG<R, SS>& A,
                                           ColG<R, SS>& B, ColG<R, SS>& C, const RcgOut<R, SS>& o, int j) {
    const unsigned kP = (unsigned)x.L.pitch;
    const unsigned f = B.f;
    const unsigned ko = x.rowk(j) + x.lk;
    C.su = o.csu;
    B.sv = o.bsv;
    if (x.own) {
        const int i = x.i;
        const unsigned cu = ucls(f), cv = vcls(f);
        if (PLAIN || j >= 1) {
            if (PLAIN || i >= 1) {
                stg_(x.a.sshn, ko, o.s_c);
                k.ct += fabs((double)o.s_c);
            }
            if (PLAIN || cu != CLS_FLATHER) {
                stg_(x.a.un, ko, o.u);
                if (PLAIN || i >= 1) k.cu += fabs((double)o.u);
            }
        }
        if (PLAIN || i >= 1) {
            if (PLAIN || cv != CLS_FLATHER) {
                stg_(x.a.vn, ko, o.v);
                if (PLAIN || j >= 1) k.cv += fabs((double)o.v);
            }
            // bc_flather_v at a row-0 point whose inner point is this one
            // (row 0 is outside the checksums)
            if (!PLAIN && j == 1 && vcls(A.f) == CLS_FLATHER && (A.f & V_FL_NORTH))
                stg_(x.a.vn, ko - kP, o.v + R(ldg_(x.flv, ko - kP)) * (A.sv - B.sv));
        }
    }
    k.south = o.nt;
    k.un_w = o.cun_w;
    k.su_w = o.csu_w;
    k.hu_w = o.chu_w;
}

// One row j: A, B, C hold rows j-1, j, j+1 (C just arrived: its sshn_u is
// formed here, and B's sshn_v); D receives row j+1+LA.  DM 3 divides by the
// metrics' reciprocals with one guard for the row; if it fails on any lane,
// the wave does the row again with IEEE divisions.
template <class R, class SS, bool DARR, int LA, bool PLAIN, bool DV = false, int DM = 0>
__device__ __forceinline__ void rcg_row(const CtxGR<R, SS, DARR>& x, CarryGR<R>& k,
                                        const ColG<R, SS>& A, ColG<R, SS>& B, ColG<R, SS>& C,
                                        ColG<R, SS>& D, int j) {
    const Layout& L = x.L;
    {
        const int jn = j + 1 + LA <= L.jpj + 2 ? j + 1 + LA : L.jpj + 2;
        load_colg<R, SS, DARR, PLAIN, DV>(D, x.n, x.fl, x.rowk(jn) + x.lk);
    }
    if constexpr (DARR && PLAIN && DV) {
        C.hu = hu_of(C);
        B.hv = hv_of(B, C);
    }
    if constexpr (DM == 3) {
        DivCheck ck;
        div_check_reset(ck);
        RcgOut<R, SS> o = rcg_compute<R, SS, DARR, PLAIN, 3>(x, k, A, B, C, j, &ck);
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(!div_check_ok(ck)) != 0, 0))
            o = rcg_compute<R, SS, DARR, PLAIN, 0>(x, k, A, B, C, j, nullptr);
        rcg_commit<R, SS, DARR, PLAIN>(x, k, A, B, C, o, j);
    } else {
        const RcgOut<R, SS> o = rcg_compute<R, SS, DARR, PLAIN, 0>(x, k, A, B, C, j, nullptr);
        rcg_commit<R, SS, DARR, PLAIN>(x, k, A, B, C, o, j);
    }
}

template <class R, class SS, bool DARR, int LA, bool PLAIN, bool DV = false, int DM = 0>
__device__ __forceinline__ void rcg_march(const CtxGR<R, SS, DARR>& x, CarryGR<R>& k, int jb,
                                          int je) {
    const Layout& L = x.L;
    ColG<R, SS> R0, R1, R2, R3, R4;
    load_colg<R, SS, DARR, PLAIN, DV>(R0, x.n, x.fl, x.rowk(jb - 1) + x.lk);
    load_colg<R, SS, DARR, PLAIN, DV>(R1, x.n, x.fl, x.rowk(jb) + x.lk);
    load_colg<R, SS, DARR, PLAIN, DV>(R2, x.n, x.fl, x.rowk(jb + 1) + x.lk);
    if constexpr (LA == 2)
        load_colg<R, SS, DARR, PLAIN, DV>(R3, x.n, x.fl,
                                          x.rowk(jb + 2 <= L.jpj + 2 ? jb + 2 : L.jpj + 2) + x.lk);
    if constexpr (DARR && PLAIN && DV) {
        R0.hu = hu_of(R0);
        R1.hu = hu_of(R1);
        R0.hv = hv_of(R0, R1);
    }
    R0.su = su_at_g<R, SS, DARR, PLAIN>(x, R0, lane_next(R0.ss), jb - 1, x.rowk(jb - 1) + x.lk);
    R1.su = su_at_g<R, SS, DARR, PLAIN>(x, R1, lane_next(R1.ss), jb, x.rowk(jb) + x.lk);
    R0.sv = sv_at_g<R, SS, DARR, PLAIN>(x, R0, R1.ss, jb - 1, x.rowk(jb - 1) + x.lk);
    k.south = north_gc<R, SS, DARR>(x.m, R0, lane_next(R0.vn), lane_next(R0.sv),
                                    DARR ? lane_next(R0.hv) : x.m.hv, R1);
    k.un_w = lane_prev(R1.un);
    k.su_w = lane_prev(R1.su);
    k.hu_w = DARR ? lane_prev(R1.hu) : x.m.hu;
    if constexpr (LA == 1) {
        for (int j = jb; j <= je; j += 4) {
            rcg_row<R, SS, DARR, 1, PLAIN, DV, DM>(x, k, R0, R1, R2, R3, j);
            if (j + 1 > je) break;
            rcg_row<R, SS, DARR, 1, PLAIN, DV, DM>(x, k, R1, R2, R3, R0, j + 1);
            if (j + 2 > je) break;
            rcg_row<R, SS, DARR, 1, PLAIN, DV, DM>(x, k, R2, R3, R0, R1, j + 2);
            if (j + 3 > je) break;
            rcg_row<R, SS, DARR, 1, PLAIN, DV, DM>(x, k, R3, R0, R1, R2, j + 3);
        }
    } else {
        for (int j = jb; j <= je; j += 5) {
            rcg_row<R, SS, DARR, 2, PLAIN, DV, DM>(x, k, R0, R1, R2, R4, j);
            if (j + 1 > je) break;
            rcg_row<R, SS, DARR, 2, PLAIN, DV, DM>(x, k, R1, R2, R3, R0, j + 1);
            if (j + 2 > je) break;
            rcg_row<R, SS, DARR, 2, PLAIN, DV, DM>(x, k, R2, R3, R4, R1, j + 2);
            if (j + 3 > je) break;
            rcg_row<R, SS, DARR, 2, PLAIN, DV, DM>(x, k, R3, R4, R0, R2, j + 3);
            if (j + 4 > je) break;
            rcg_row<R, SS, DARR, 2, PLAIN, DV, DM>(x, k, R4, R0, R1, R3, j + 4);
        }
    }
}

// plain: per (wave strip of GROUT columns, NJ-row segment) 1 where every
// class word the wave touches (with LA rows loaded ahead) is G_PLAIN; null
// for none
// DV: the grid's hu and hv are the means of adjacent T depths (plain tiles
// form them from ht instead of loading them).
// DM 3 (double): plain tiles divide by the metrics' reciprocals rc.
template <class R, class SS, bool DARR, int NJ, int LA, bool DV = false, int DM = 0>
__global__ void __launch_bounds__(64 * GW)
k_step_march_gr(MetG<R> m, ParamsT<R> p, StateG<R, SS, DARR> n, OutG<R, SS> a,
                const uint16_t* __restrict__ fl, const double* __restrict__ flv, Layout L,
                double tide, double* __restrict__ partials, int nbx, int ntiles,
                const uint8_t* __restrict__ plain, RcpG rc) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int per = gridDim.x >> 3;  // gridDim.x is a multiple of 8
    const int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    CarryGR<R> k;
    k.ct = k.cu = k.cv = 0.0;
    if (t < ntiles) {
        const int strip = (t % nbx) * GW + wave, seg = t / nbx;
        CtxGR<R, SS, DARR> x;
        x.m = m;
        x.p = p;
        x.n = n;
        x.a = a;
        x.fl = fl;
        x.flv = flv;
        x.L = L;
        x.tide = tide;
        x.rc = rc;
        // lane 0 is column i0 - 1 (>= -1, inside the left padding)
        x.i = strip * GROUT + lane - 1;
        x.il = x.i < L.jpi + 1 ? x.i : L.jpi + 1;  // clamped column for loads
        x.lk = (unsigned)x.il;                     // wraps for -1; sums stay in range
        x.own = lane >= 1 && lane <= GROUT && x.i <= L.jpi;
        x.tcol = x.i >= 1 && x.i <= L.jpi;
        x.ucol = x.i >= 0 && x.i <= L.jpi;
        x.vcol = x.i >= 1 && x.i <= L.jpi;
        const int jb = seg * NJ;
        const int je = jb + NJ - 1 < L.jpj ? jb + NJ - 1 : L.jpj;
        if (plain && plain[(long)strip * (ntiles / nbx) + seg])
            rcg_march<R, SS, DARR, LA, true, DV, DM>(x, k, jb, je);
        else
            rcg_march<R, SS, DARR, LA, false>(x, k, jb, je);
    }
    block_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)blockIdx.x);
}

// Block b of a launch goes to XCD b % 8; hand each XCD a contiguous range of
// tiles so that tiles sharing halo rows share an L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int per = nb >> 3;  // nb is a multiple of 8
    return (b & 7) * per + (b >> 3);
}

// Sums of the wave's lanes (fixed order), written by lane 0 to out[0..5].
__device__ __forceinline__ void wave_reduce3(DD a0, DD a1, DD a2, double* out) {
    for (int mk = 32; mk >= 1; mk >>= 1) {
        a0 = dd_merge(a0, dd_shfl_xor(a0, mk));
        a1 = dd_merge(a1, dd_shfl_xor(a1, mk));
        a2 = dd_merge(a2, dd_shfl_xor(a2, mk));
    }
    if ((threadIdx.x & 63) == 0) {
        out[0] = a0.s; out[1] = a0.c;
        out[2] = a1.s; out[3] = a1.c;
        out[4] = a2.s; out[5] = a2.c;
    }
}

}  // namespace nemo2d
