// The sshn-only column-marching step with uniform statics (the body of
// k_step_march_rc, shared with the ensemble's FP64 step): row loads, the
// shared north-face terms, next's interpolations recomputed from sshn, the
// guarded division by reciprocals and the segment march.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nemo2d_core.h"
#include "nemo2d_march.h"

namespace nemo2d {

// Accesses by 32-bit byte offset from a kernel-argument base (the whole
// padded array is below 4 GiB), so the address is base + offset in the
// load/store itself.
template <class T>
__device__ __forceinline__ T ldb(const T* b, unsigned boff) {
    return *(const T*)((const char*)b + boff);
}
template <class T>
__device__ __forceinline__ void stb(T* b, unsigned boff, T v) {
    *(T*)((char*)b + boff) = v;
}

// Terms on the north faces of the points of row `c` (row `nr` above it).
struct NorthTerms {
    double r3;     // continuity: (sshn_v + hv) * vn
    double kn;     // momentum_u: uu_n * v_n * depn, no wall
    double vnc;    // momentum_u: v_nc
    double vnu;    // momentum_u: v_n
    double depnu;  // momentum_u: depn
    double hn;     // momentum_v: vv_n * v_n * depn
    double dvdyn;  // momentum_v: dvdy_n
    double qdudy;  // momentum_u: (un_n - un_c) / (e2u + e2u), dudy_n's quotient
};
// Own-column row values of the five now-state arrays.
struct Col5 {
    double un, vn, ss, su, sv;
};
// c: this row (own column), c_vn_e / c_sv_e: its vn and sshn_v one column
// east, nr: the row above (own column)
// x / d for the uniform divisor d of tag T (QD_*).  DM 0: the IEEE
// division; 1: the guarded division by the precomputed reciprocal (div_rcp);
// 2: the unguarded one, its guard gathered in *ck (div_rcp_acc); 3: as 2
// with one correction (div_rcp1_acc; grids with Scal::one_corr only).  All
// give x / d's bits (2 and 3 when the gathered guard holds).
template <int DM, int T>
__device__ __forceinline__ double udiv(const Scal& s, double x, double d, DivCheck* ck = nullptr) {
    if constexpr (DM == 3) return div_rcp1_acc(x, d, s.rcp[T], *ck);
    else if constexpr (DM == 2) return div_rcp_acc(x, d, s.rcp[T], *ck);
    else if constexpr (DM == 1) return div_rcp(x, d, s.rcp[T]);
    else return x / d;
}

template <int FD = 0>
__device__ __forceinline__ NorthTerms north_terms1(const Scal& s, const Col5& c, double c_vn_e,
                                                   double c_sv_e, const Col5& nr,
                                                   DivCheck* ck = nullptr) {
    const double h = 0.5;
    NorthTerms t;
    t.r3 = (c.sv + s.hv) * c.vn;
    // momentum_u (original/nemolite2d.f90:439-514) at this row's U point
    t.vnc = h * (c.vn + c_vn_e);
    t.vnu = h * t.vnc * (s.e1v + s.e1v);
    t.depnu = h * (s.hv + c.sv + s.hv + c_sv_e);
    const double sn = sgn_half(t.vnu);
    const double uun = (h + sn) * c.un + (h - sn) * nr.un;
    t.kn = uun * t.vnu * t.depnu;
    // momentum_v (:527-606) at this row's V point
    const double vn_n = h * (c.vn + nr.vn) * s.e1t;
    const double depn = s.ht + nr.ss;
    const double sv = sgn_half(vn_n);
    const double vvn = (h + sv) * c.vn + (h - sv) * nr.vn;
    t.hn = vvn * vn_n * depn;
    t.dvdyn = udiv<FD, QD_E2T>(s, nr.vn - c.vn, s.e2t, ck) * (s.ht + nr.ss);
    t.qdudy = udiv<FD, QD_E2U2>(s, nr.un - c.un, s.e2u + s.e2u, ck);
    return t;
}

struct ShCtx {
    Scal s;
    Params p;
    NState n;
    AOut a;
    const uint16_t* fl;
    const double* flv;  // Flather sqrt(g/hv) (row-0 points)
    Layout L;
    double tide;
    int i, il;
    unsigned lb8;  // 8 * (il - 1): byte offset of the window's west column in a row
    bool own, tcol;
    __device__ unsigned rowb(int j) const { return (unsigned)(8 * L.at(0, j)); }
};

// A row's own-column state and class word, loaded together: a wait for
// any of them (vector loads complete in issue order) then never waits for
// loads issued later for rows further ahead.
struct Col5F : Col5 {
    unsigned f;
};

// ---- fused step reading sshn, un, vn only (uniform statics) -----------------------
// sshn_u and sshn_v of a state are next's interpolations of its sshn
// (original/nemolite2d.f90:735-762) except at the points next leaves alone,
// which keep their values.  This kernel reads sshn, un and vn (and the class
// word), recomputes the now-state's sshn_u and sshn_v from sshn with next's
// own expressions (using the stored value only where next leaves a point
// alone; tiles that are not plain read it with the row), and writes sshn, un
// and vn: 48 of the 80 bytes per cell of the five-array step (64 on tiles
// that are not plain).  The new state's sshn_u/sshn_v arrays are not written; the
// library fills them with the same interpolation before anything reads them.
// The ring, the shared face terms and the lane shifts are those of
// k_step_march_dp; since a lane's sshn_u needs the sshn one column east, the
// east halo is two lanes wide (lanes 1..61 own the columns).
constexpr int ROUT = 61;
// The class word of an interior point that nothing special touches: U and V
// computed, no walls, no clamp, next averaging both interpolations.
constexpr unsigned RC_PLAIN = (CLS_COMP << U_CLS_SHIFT) | (CLS_COMP << V_CLS_SHIFT) |
                              (NX_AVG << SU_SHIFT) | (NX_AVG << SV_SHIFT);
constexpr int RC_NJ = 12, RC_LA = 2;  // the tiling the plain-tile table is made for
constexpr int RC_NJ_SHORT = 8;       // the shorter segments of mid-size grids (a second table)

// sshn_u / sshn_v at a point from next's class of it (NX_SKIP where next does
// not set it): next_su / next_sv's expressions, or the stored value (loaded
// with the row on tiles that are not plain) where next leaves the point alone
template <int DM, bool PLAIN = false>
__device__ __forceinline__ double rc_su(const Scal& s, unsigned cls, double s_lo, double s_hi,
                                        double stored, DivCheck* ck) {
    const double r = s.e12t * s_lo + s.e12t * s_hi;
    if constexpr (PLAIN) return udiv<DM, QD_E12U>(s, 0.5 * r, s.e12u, ck);
    const double v = cls == NX_HI ? s_hi : cls == NX_LO ? s_lo : udiv<DM, QD_E12U>(s, 0.5 * r, s.e12u, ck);
    return cls == NX_SKIP ? stored : v;
}
template <int DM, bool PLAIN = false>
__device__ __forceinline__ double rc_sv(const Scal& s, unsigned cls, double s_lo, double s_hi,
                                        double stored, DivCheck* ck) {
    const double r = s.e12t * s_lo + s.e12t * s_hi;
    if constexpr (PLAIN) return udiv<DM, QD_E12V>(s, 0.5 * r, s.e12v, ck);
    const double v = cls == NX_HI ? s_hi : cls == NX_LO ? s_lo : udiv<DM, QD_E12V>(s, 0.5 * r, s.e12v, ck);
    return cls == NX_SKIP ? stored : v;
}

// A row's sshn, un, vn; on tiles that are not plain also its class word and
// the stored sshn_u / sshn_v (kept where next leaves a point alone).  Loaded
// with the row, these wait with it: a value fetched when first needed would
// be the youngest load in flight, and vector loads complete in issue order,
// so waiting for it would wait for the rows prefetched ahead too.
template <bool PLAIN = false>
__device__ __forceinline__ void load_col3(Col5F& r, const NState& n, const uint16_t* fl,
                                          unsigned boff) {
    r.un = ldb(n.un, boff);
    r.vn = ldb(n.vn, boff);
    r.ss = ldb(n.sshn, boff);
    if constexpr (PLAIN) {
        r.f = RC_PLAIN;
    } else {
        r.f = ldb(fl, boff >> 2);
        r.su = ldb(n.sshn_u, boff);
        r.sv = ldb(n.sshn_v, boff);
    }
}

struct RcCtx : ShCtx {
    bool ucol, vcol;  // this column holds next-set U / V points (0..jpi / 1..jpi)
    // next's class of the U / V point of this lane in row r (NX_SKIP outside
    // the rows next sets)
    __device__ unsigned ucls_at(unsigned f, int r) const {
        return (ucol && r >= 1 && r <= L.jpj) ? sucls(f) : NX_SKIP;
    }
    __device__ unsigned vcls_at(unsigned f, int r) const {
        return (vcol && r >= 0 && r <= L.jpj) ? svcls(f) : NX_SKIP;
    }
};

struct RcCarry {
    NorthTerms south;         // the row below's north terms
    double un_w, su_w;        // un and sshn_u one column west, this row
    double ss_e;              // sshn one column east, this row
    double ct, cu, cv;        // per-lane checksum sums
};

// What row j computes: the row's north terms, C's sshn_u and B's sshn_v
// (the now-state's interpolations first needed here), the west values
// carried to the next row, and the new sshn, u and v.
struct RcOut {
    NorthTerms nt;
    double csu, bsv, cun_w, csu_w, css_e;
    double s_c, u, v;
};

// Row j's arithmetic, no stores: A, B, C hold rows j-1, j, j+1 (C has just
// arrived).  With DM 2 the divisions' guard is gathered in *ck.
// With PLAIN every point the wave touches has the class word RC_PLAIN (the
// tile table says so), and the class tests drop out.
template <int DM, bool PLAIN = false>
__device__ __forceinline__ RcOut rc_compute(const RcCtx& x, const RcCarry& k, const Col5F& A,
                                            const Col5F& B0, const Col5F& C0, int j, DivCheck* ck) {
    const Scal& s = x.s;
    const double h = 0.5;
    const unsigned f = B0.f;
    RcOut o;
    Col5F B = B0, C = C0;
    o.css_e = lane_next(C.ss);  // (the next row's east sshn too)
    C.su = o.csu = rc_su<DM, PLAIN>(s, x.ucls_at(C.f, j + 1), C.ss, o.css_e, C0.su, ck);
    B.sv = o.bsv = rc_sv<DM, PLAIN>(s, x.vcls_at(f, j), B.ss, C.ss, B0.sv, ck);
    // neighbours by lane shifts (every lane active here)
    const double bun_w = k.un_w, bsu_w = k.su_w, bss_e = k.ss_e;
    const double bvn_w = lane_prev(B.vn), bsv_w = lane_prev(B.sv);
    const double bvn_e = lane_next(B.vn), bsv_e = lane_next(B.sv);
    o.cun_w = lane_prev(C.un);
    o.csu_w = lane_prev(C.su);
    const double cun_w = o.cun_w, csu_w = o.csu_w;
    o.nt = north_terms1<DM>(s, B, bvn_e, bsv_e, C, ck);
    const NorthTerms& nt = o.nt;
    const NorthTerms& st = k.south;

    // continuity (original/nemolite2d.f90:399-403)
    double s_c = 0.0;
    if (PLAIN || (j >= 1 && x.tcol)) {
        if (!PLAIN && (f & T_CLAMP)) {
            s_c = x.tide;
        } else {
            const double r1 = (B.su + s.hu) * B.un;
            const double r2 = (bsu_w + s.hu) * bun_w;
            s_c = B.ss + udiv<DM, QD_E12T>(s, (r2 - r1 + st.r3 - nt.r3) * x.p.rdt, s.e12t, ck);
        }
    }
    o.s_c = s_c;

    // west-face terms of this lane's points; the east ones come from lane+1
    const double un_c = B.un;
    const double u_w = h * (un_c + bun_w) * s.e2t;
    const double depw_u = s.ht + B.ss;
    const double sw_u = sgn_half(u_w);
    const double fw = ((h - sw_u) * un_c + (h + sw_u) * bun_w) * u_w * depw_u;
    const double dudx_w = udiv<DM, QD_E1T>(s, un_c - bun_w, s.e1t, ck) * (s.ht + B.ss);
    const double vn_c = B.vn;
    const double u_wc = h * (bun_w + cun_w);
    const double u_wv = h * u_wc * (s.e2u + s.e2u);
    const double depw_v = h * (s.hu + bsu_w + s.hu + csu_w);
    const double sw_v = sgn_half(u_wv);
    const double gw = ((h - sw_v) * vn_c + (h + sw_v) * bvn_w) * u_wv * depw_v;
    // dvdx_w's quotient; lane+1's is dvdx_e's
    const double qdvdx_w = udiv<DM, QD_E1V2>(s, vn_c - bvn_w, s.e1v + s.e1v, ck);
    const double fe = lane_next(fw);
    const double dudx_e = lane_next(dudx_w);
    const double ge = lane_next(gw);
    const double u_ec = lane_next(u_wc);
    const double qdvdx_e = lane_next(qdvdx_w);

    double u = 0.0, v = 0.0;
    if (x.own) {
        const int i = x.i;
        const unsigned cu = ucls(f), cv = vcls(f);
        if (PLAIN || (j >= 1 && cu == CLS_COMP)) {
            const bool ws = !PLAIN && (f & U_WALL_S) != 0, wn = !PLAIN && (f & U_WALL_N) != 0;
            double ts = st.kn, tn = nt.kn;
            if (ws) ts = ((h - sgn_half(st.vnu)) * un_c) * st.vnu * st.depnu;
            if (wn) tn = ((h + sgn_half(nt.vnu)) * un_c) * nt.vnu * nt.depnu;
            const double adv = fw - fe + ts - tn;
            double dudy_s = 0.0, dudy_n = 0.0;
            if (!ws) dudy_s = st.qdudy * (s.hu + B.su + s.hu + A.su);
            if (!wn) dudy_n = nt.qdudy * (s.hu + B.su + s.hu + C.su);
            double vis = (dudx_e - dudx_w) * s.e2u + (dudy_n - dudy_s) * s.e1u * h;
            vis = x.p.visc * vis;
            const double hsu = s.hu + B.su;
            const double cor = h * (s.fcu * (st.vnc + nt.vnc)) * s.e12u * hsu;
            const double hpg = -GRAV * hsu * s.e2u * (bss_e - B.ss);
            const double q = udiv<DM, QD_E12U>(s, x.p.rdt * (adv + vis + cor + hpg), s.e12u, ck);
            u = udiv<DM, QD_FRIC>(s, udiv<DM, QD_HU>(s, un_c * hsu + q, s.hu + 0.0, ck), x.p.fric, ck);
        }
        if (PLAIN || (i >= 1 && cv == CLS_COMP)) {
            const bool ww = !PLAIN && (f & V_WALL_W) != 0, we = !PLAIN && (f & V_WALL_E) != 0;
            double tw = gw, te = ge;
            if (ww) tw = ((h - sw_v) * vn_c) * u_wv * depw_v;
            if (we) {
                const double u_e = h * u_ec * (s.e2u + s.e2u);
                const double depe = h * (s.hu + B.su + s.hu + C.su);
                te = ((h + sgn_half(u_e)) * vn_c) * u_e * depe;
            }
            const double adv = tw - te + st.hn - nt.hn;
            double dvdx_w = 0.0, dvdx_e = 0.0;
            if (!ww) dvdx_w = qdvdx_w * (s.hv + B.sv + s.hv + bsv_w);
            if (!we) dvdx_e = qdvdx_e * (s.hv + B.sv + s.hv + bsv_e);
            double vis = (nt.dvdyn - st.dvdyn) * s.e1v + (dvdx_e - dvdx_w) * s.e2v * h;
            vis = x.p.visc * vis;
            const double hsv = s.hv + B.sv;
            const double cor = -h * (s.fcv * (u_ec + u_wc)) * s.e12v * hsv;
            const double hpg = -GRAV * hsv * s.e1v * (C.ss - B.ss);
            const double q = udiv<DM, QD_E12V>(s, x.p.rdt * (adv + vis + cor + hpg), s.e12v, ck);
            v = udiv<DM, QD_FRIC>(s, udiv<DM, QD_HV>(s, vn_c * hsv + q, s.hv + 0.0, ck), x.p.fric, ck);
        }
    }
    o.u = u;
    o.v = v;
    return o;
}

// Row j's stores and carries.
template <bool PLAIN = false>
__device__ __forceinline__ void rc_commit(const RcCtx& x, RcCarry& k, const Col5F& A, Col5F& B,
                                          Col5F& C, const RcOut& o, int j) {
    const long P = x.L.pitch;
    const unsigned f = B.f;
    const unsigned o8 = x.rowb(j) + x.lb8;
    if (PLAIN && x.own) {
        stb(x.a.sshn, o8, o.s_c);
        k.ct += fabs(o.s_c);
        stb(x.a.un, o8, o.u);
        k.cu += fabs(o.u);
        stb(x.a.vn, o8, o.v);
        k.cv += fabs(o.v);
    } else if (!PLAIN && x.own) {
        const int i = x.i;
        const unsigned cu = ucls(f), cv = vcls(f);
        if (j >= 1) {
            if (i >= 1) {
                stb(x.a.sshn, o8, o.s_c);
                k.ct += fabs(o.s_c);
            }
            if (cu != CLS_FLATHER) {
                stb(x.a.un, o8, o.u);
                if (i >= 1) k.cu += fabs(o.u);
            }
        }
        if (i >= 1) {
            if (cv != CLS_FLATHER) {
                stb(x.a.vn, o8, o.v);
                if (j >= 1) k.cv += fabs(o.v);
            }
            // bc_flather_v (original/nemolite2d.f90:676-690) at a row-0 point
            // whose inner point is this one; row 0 is outside the checksums
            if (j == 1 && vcls(A.f) == CLS_FLATHER && (A.f & V_FL_NORTH)) {
                const unsigned o08 = o8 - (unsigned)(8 * P);
                stb(x.a.vn, o08, o.v + ldb(x.flv, o08) * (A.sv - o.bsv));
            }
        }
    }
    k.south = o.nt;
    k.un_w = o.cun_w;
    k.su_w = o.csu_w;
    k.ss_e = o.css_e;
    C.su = o.csu;
    B.sv = o.bsv;
}

// One row j: A, B, C hold rows j-1, j, j+1; D receives row j+3 (row j+2
// is in flight).  With DM 2 or 3 the row's divisions go by reciprocal with
// one guard for the whole row; if it fails on any lane, the wave does the
// row again with IEEE divisions.
template <int DM, bool PLAIN = false>
__device__ __forceinline__ void rc_row(const RcCtx& x, RcCarry& k, const Col5F& A, Col5F& B,
                                       Col5F& C, Col5F& D, int j) {
    {
        const int jn = j + 3 <= x.L.jpj + 2 ? j + 3 : x.L.jpj + 2;
        load_col3<PLAIN>(D, x.n, x.fl, x.rowb(jn) + x.lb8);
    }
    if constexpr (DM == 2 || DM == 3) {
        DivCheck ck;
        div_check_reset(ck);
        RcOut o = rc_compute<DM, PLAIN>(x, k, A, B, C, j, &ck);
        if (__builtin_expect(__builtin_amdgcn_ballot_w64(!div_check_ok(ck)) != 0, 0))
            o = rc_compute<0, PLAIN>(x, k, A, B, C, j, nullptr);
        rc_commit<PLAIN>(x, k, A, B, C, o, j);
    } else {
        const RcOut o = rc_compute<DM, PLAIN>(x, k, A, B, C, j, nullptr);
        rc_commit<PLAIN>(x, k, A, B, C, o, j);
    }
}

// One segment rows jb..je of a wave's strip, loads two rows ahead.
template <int DM, bool PLAIN>
__device__ __forceinline__ void rc_march(const RcCtx& x, RcCarry& k, int jb, int je) {
    const Layout& L = x.L;
    const NState& n = x.n;
    const uint16_t* fl = x.fl;
    const Scal& sc_ = x.s;
    Col5F R0, R1, R2, R3, R4;
    load_col3<PLAIN>(R0, n, fl, x.rowb(jb - 1) + x.lb8);
    load_col3<PLAIN>(R1, n, fl, x.rowb(jb) + x.lb8);
    load_col3<PLAIN>(R2, n, fl, x.rowb(jb + 1) + x.lb8);
    load_col3<PLAIN>(R3, n, fl, x.rowb(jb + 2 <= L.jpj + 2 ? jb + 2 : L.jpj + 2) + x.lb8);
    // the segment's first interpolations and south terms, IEEE divisions
    R0.su = rc_su<0, PLAIN>(sc_, x.ucls_at(R0.f, jb - 1), R0.ss, lane_next(R0.ss), R0.su, nullptr);
    k.ss_e = lane_next(R1.ss);
    R1.su = rc_su<0, PLAIN>(sc_, x.ucls_at(R1.f, jb), R1.ss, k.ss_e, R1.su, nullptr);
    R0.sv = rc_sv<0, PLAIN>(sc_, x.vcls_at(R0.f, jb - 1), R0.ss, R1.ss, R0.sv, nullptr);
    k.south = north_terms1<0>(sc_, R0, lane_next(R0.vn), lane_next(R0.sv), R1);
    k.un_w = lane_prev(R1.un);
    k.su_w = lane_prev(R1.su);
    // slot roles rotate, nothing is copied: rows j-1, j, j+1 held, the next
    // two rows in flight or loading
    for (int j = jb; j <= je; j += 5) {
        rc_row<DM, PLAIN>(x, k, R0, R1, R2, R4, j);
        if (j + 1 > je) break;
        rc_row<DM, PLAIN>(x, k, R1, R2, R3, R0, j + 1);
        if (j + 2 > je) break;
        rc_row<DM, PLAIN>(x, k, R2, R3, R4, R1, j + 2);
        if (j + 3 > je) break;
        rc_row<DM, PLAIN>(x, k, R3, R4, R0, R2, j + 3);
        if (j + 4 > je) break;
        rc_row<DM, PLAIN>(x, k, R4, R0, R1, R3, j + 4);
    }
}

}  // namespace nemo2d
