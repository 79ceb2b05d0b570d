// nemo2d: MI355X (gfx950) NEMOLite2D FP64 time-stepper behind the C-ABI of
// include/nemo2d.h.
//
// Two device paths share the point functions of nemo2d_core.h:
//   * stage kernels, one per call of the serial model's step
//     (original/nemolite2d.f90:375-769), each over its Fortran index range,
//     writing the after-arrays ssha/ua/va like the serial model;
//   * a fused step: one HBM pass that reads the now-state once and writes
//     the new now-state (ping-pong buffers), with the boundary stages folded
//     in as per-point classes and the per-step checksums reduced on the fly,
//     followed by a one-block kernel that finishes the checksum reduction.
// The O(N^2) mask scans of the serial model's boundary code become per-point
// flag bits and per-line lists built once on the host.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <thread>
#include <type_traits>
#include <algorithm>
#include <vector>

#include "nemo2d.h"
#include "nemo2d_core.h"
#include "nemo2d_internal.h"
#include "nemo2d_step.h"
#include "nemo2d_march.h"
#include "nemo2d_rc.h"

using namespace nemo2d;

namespace {

thread_local std::string g_err = "no error";

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(NEMO2D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCHCHK(what)                                                                        \
    do {                                                                                       \
        hipError_t e_ = hipGetLastError();                                                     \
        if (e_ != hipSuccess)                                                                  \
            return fail(NEMO2D_ERR_HIP, std::string("launch ") + what + ": " + hipGetErrorString(e_)); \
    } while (0)

enum { F_SSHN = 0, F_SSHN_U, F_SSHN_V, F_UN, F_VN, NF };
enum {
    S_E1T = 0, S_E2T, S_E1U, S_E2U, S_E1V, S_E2V, S_E12T, S_E12U, S_E12V,
    S_HT, S_HU, S_HV, S_FCU, S_FCV, NS
};


}  // namespace

struct nemo2d_model {
    int dev = 0;
    hipStream_t stream = nullptr;
    Layout L{};
    Params prm{};
    int nit000 = 1, nitend = 1, irecord = 1;
    double dx = 0, dy = 0, dep_const = 0;

    double* S[2][NF] = {};  // now-state ping-pong
    int cur = 0;
    double *ssha = nullptr, *ua = nullptr, *va = nullptr;  // stage path after-arrays
    bool a_current = false;  // ua/va hold the last step's after-values

    uint16_t* flags = nullptr;
    double *flu = nullptr, *flv = nullptr;  // Flather sqrt(g/h) at Flather points
    double* st[NS] = {};
    bool uniform = false;
    bool uni_metrics = false;  // metrics and Coriolis uniform (depths may vary)
    bool dep_derived = false;  // interior hu, hv are the means of adjacent ht
    Scal scal{};
    bool fused_ok = false;

    // Flather point lists in the serial model's loop order:
    //   u: rows jj, entries sorted by ji;  v: columns ji, entries sorted by jj.
    int* flu_ptr = nullptr;   // jpj+1 row pointers (rows 1..jpj)
    long* flu_pts = nullptr;
    int* flv_ptr = nullptr;   // jpi+1 column pointers (columns 1..jpi)
    long* flv_pts = nullptr;
    int nflu = 0, nflv = 0;
    int nflv0 = 0;            // v points of row 0 whose inner point is north

    // checksum reduction
    double* partials = nullptr;  // 6 doubles per block
    int npart_max = 0;
    // sshn_u/sshn_v of state set k not written by the last step (they are
    // the interpolations of its sshn; uv_sync fills them in), and whether the
    // current state's sshn_u/sshn_v are known to be those interpolations
    bool uv_stale[2] = {false, false};
    bool uv_derived = false;
    // the stage calls of a step recorded so far (in order) for one fused
    // pass at nemo2d_next, and bc_ssh's time
    int lz_n = 0;
    double lz_rtime = 0.0;
    bool lz_on = true;
    // per (wave strip, 12-row segment) of k_step_march_rc's tiling: 1 where
    // every point the wave touches has the plain interior class word
    uint8_t* rc_plain = nullptr;
    uint8_t* rc_plain_short = nullptr;  // the same for RC_NJ_SHORT-row segments
    double* cs_now = nullptr;    // 3 doubles: last on-demand checksums
    double* chk = nullptr;       // one word: nemo2d_set_state's interpolation check
    double* cs_log = nullptr;    // 3 doubles per logged step
    // Deferred checksum finish: the fused step leaves each logged step's
    // block partials in its own slot of cs_part; reading the log reduces
    // all pending slots in one launch.
    double* cs_part = nullptr;   // cs_cap slots of cs_slot doubles (or null)
    long cs_slot = 0;
    std::vector<int> cs_pend;    // partials waiting in each slot (0: none)
    int* d_pend = nullptr;       // device copy of (slot, nparts) pairs
    bool pend_any = false;
    double* part_dst = nullptr;  // where the step kernel writes its partials
    int cs_cap = 0;
    int cs_next = 0;             // next free slot for stage-path logging
    int cs_n = 0;                // steps logged (max slot + 1)
    double* h_pinned = nullptr;  // 3 doubles of pinned host memory
    // one padded FP64 array's bytes of device memory through which host
    // arrays travel: one contiguous copy across the bus and a strided one
    // on the device (a strided copy from or to pageable host memory goes a
    // row at a time: 0.34 s for the 14 statics of a 2048^2 grid)
    void* stage = nullptr;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // order the closure's hand-overs between this stream and the default
    // stream (buffers free, features or targets out, increments in,
    // increments read) without a host wait
    hipEvent_t ev_free = nullptr, ev_feat = nullptr, ev_incr = nullptr, ev_read = nullptr, ev_tgt = nullptr;
    bool grid_set = false, state_set = false;

    // single-precision form: float state (sshn in double when compensated),
    // float depths, scalar float metrics
    bool mixed = false, compensated = true;
    int variant = 0;              // fused-step kernel variant (0 = default)
    int last_var = 0;             // the variant the last fused step ran (0: none yet)
    double tide_amp = 0.2;                      // original/nemolite2d.f90:629
    double tide_omega = tide_omega_of(12.42);   // :630
    void* X[2][NF] = {};          // float (or double sshn) now-state ping-pong
    float *hf_t = nullptr, *hf_u = nullptr, *hf_v = nullptr;
    double* diag = nullptr;       // 2 doubles: volume, energy

    // asynchronous snapshots (go2d files): pinned copies of sshn, un, vn
    // (padded layout) filled on copy_stream, written by snap_thread
    std::vector<double> out_xt, out_yt, out_ht;  // T layout, from set_output_grid
    double* h_snap[3] = {};
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_state = nullptr, ev_copy = nullptr;
    int snap_src = -1;            // now-state buffer being copied (-1: none)
    std::thread snap_thread;
    int snap_rc = NEMO2D_OK;
    std::string snap_err;

    // device timers of the serial model's regions (nemo2d_stage_timers)
    bool timing = false;
    std::vector<hipEvent_t> tev;  // event pool; tev_used in flight
    int tev_used = 0;
    struct TCall {
        int region, b, e;
    };
    std::vector<TCall> tcalls;
    int topen[5] = {-1, -1, -1, -1, -1};
    long tcount[5] = {};
    double tsum[5] = {}, tsq[5] = {};  // ms, ms^2
};

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
namespace {

// ---- fused step, one thread per cell (kernel wrapper) ---------------------------
template <class S, bool FLIN>
__global__ void __launch_bounds__(TBX * TBY)
k_step_fused(S st, NState n, AOut a, const uint16_t* __restrict__ fl,
             const double* __restrict__ flu, const double* __restrict__ flv, Layout L,
             Params p, double tide, double* __restrict__ partials) {
    DD ct = {0.0, 0.0}, cu = {0.0, 0.0}, cv = {0.0, 0.0};
    cell_step<S, FLIN>(st, n, a, fl, flu, flv, L, p, tide, blockIdx.x * TBX + threadIdx.x,
                       blockIdx.y * TBY + threadIdx.y, ct, cu, cv);
    block_reduce3(ct, cu, cv, partials + 6 * (long)(blockIdx.x + gridDim.x * blockIdx.y));
}

// single-precision form, per-point depths
template <bool COMP>
__global__ void __launch_bounds__(TBX * TBY)
k_step_mixed(ScalF s, MixedState<COMP> n, MixedOut<COMP> a, const float* __restrict__ ht,
             const float* __restrict__ hu, const float* __restrict__ hv,
             const uint16_t* __restrict__ fl, const double* __restrict__ flu,
             const double* __restrict__ flv, Layout L, ParamsT<float> p, double tide,
             double* __restrict__ partials) {
    DD ct = {0.0, 0.0}, cu = {0.0, 0.0}, cv = {0.0, 0.0};
    cell_step_mixed<COMP, false>(s, n, a, ht, hu, hv, fl, flu, flv, L, p, tide,
                                 blockIdx.x * TBX + threadIdx.x, blockIdx.y * TBY + threadIdx.y,
                                 ct, cu, cv);
    block_reduce3(ct, cu, cv, partials + 6 * (long)(blockIdx.x + gridDim.x * blockIdx.y));
}

// ---- fused step, column-marching form (uniform statics) ----------------------
// Each wave owns MOUT = 63 consecutive columns and marches up a segment of
// NJ rows, keeping a 3x3 register window of the now-state per column; each
// new row is one coalesced row of loads.  The new ssh of row j is computed
// once: the column to the east comes from the next lane (lane 63 is the
// east halo of the wave and only supplies that value), the row to the north
// from the next iteration, so next's interpolations need no recomputation.
constexpr int MW = 4;     // waves per block (side by side in i)
constexpr int MOUT = 63;  // output columns per wave

struct RegWin {
    using real = double;
    Scal s;
    double un_[3][3], vn_[3][3], sshn_[3][3], su_[3][3], sv_[3][3];  // [DJ+1][DI+1]
#define NEMO2D_REGF(F, A) \
    template <int DI, int DJ> __device__ double F() const { return A[DJ + 1][DI + 1]; }
    NEMO2D_REGF(un, un_) NEMO2D_REGF(vn, vn_) NEMO2D_REGF(sshn, sshn_) NEMO2D_REGF(sshn_u, su_)
    NEMO2D_REGF(sshn_v, sv_)
#undef NEMO2D_REGF
#define NEMO2D_REGS(F) \
    template <int DI, int DJ> __device__ double F() const { return s.F; }
    NEMO2D_REGS(e1t) NEMO2D_REGS(e2t) NEMO2D_REGS(e1u) NEMO2D_REGS(e2u) NEMO2D_REGS(e1v)
    NEMO2D_REGS(e2v) NEMO2D_REGS(e12t) NEMO2D_REGS(e12u) NEMO2D_REGS(e12v) NEMO2D_REGS(ht)
    NEMO2D_REGS(hu) NEMO2D_REGS(hv) NEMO2D_REGS(fcu) NEMO2D_REGS(fcv)
#undef NEMO2D_REGS
    template <int TAG> __device__ double qd(double x, double d) const { return x / d; }

    // prefetched next row (software pipelining: loaded one iteration ahead)
    double nun_[3], nvn_[3], nsshn_[3], nsu_[3], nsv_[3];
    __device__ __forceinline__ void shift_in_next() {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            un_[0][c] = un_[1][c]; un_[1][c] = un_[2][c]; un_[2][c] = nun_[c];
            vn_[0][c] = vn_[1][c]; vn_[1][c] = vn_[2][c]; vn_[2][c] = nvn_[c];
            sshn_[0][c] = sshn_[1][c]; sshn_[1][c] = sshn_[2][c]; sshn_[2][c] = nsshn_[c];
            su_[0][c] = su_[1][c]; su_[1][c] = su_[2][c]; su_[2][c] = nsu_[c];
            sv_[0][c] = sv_[1][c]; sv_[1][c] = sv_[2][c]; sv_[2][c] = nsv_[c];
        }
    }
    __device__ __forceinline__ void load_next(const NState& n, long rowoff) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            nun_[c] = n.un[rowoff + c];
            nvn_[c] = n.vn[rowoff + c];
            nsshn_[c] = n.sshn[rowoff + c];
            nsu_[c] = n.sshn_u[rowoff + c];
            nsv_[c] = n.sshn_v[rowoff + c];
        }
    }
    // load window row r from storage row j, columns i-1..i+1 (rowoff = offset
    // of (i-1, j))
    __device__ __forceinline__ void load(int r, const NState& n, long rowoff) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            un_[r][c] = n.un[rowoff + c];
            vn_[r][c] = n.vn[rowoff + c];
            sshn_[r][c] = n.sshn[rowoff + c];
            su_[r][c] = n.sshn_u[rowoff + c];
            sv_[r][c] = n.sshn_v[rowoff + c];
        }
    }
};

// The five-array marching step for grids whose arrays exceed the shared-term
// kernels' 32-bit byte offsets: 16-row segments, the next row's loads issued
// one iteration ahead, IEEE divisions, Flather points left to the epilogue.
constexpr int M16 = 16;
__global__ void __launch_bounds__(64 * MW)
k_step_march(Scal sc_, NState n, AOut a, const uint16_t* __restrict__ fl, Layout L, Params p,
             double tide, double* __restrict__ partials, int nbx, int ntiles) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const long P = L.pitch;
    // per-lane sums of at most NJ terms in plain double (relative error
    // below NJ*2^-53), merged across lanes and blocks in double-double
    double ct_ = 0.0, cu_ = 0.0, cv_ = 0.0;
    if (t < ntiles) {
        const int bx = t % nbx, by = t / nbx;
        const int i = (bx * MW + wave) * MOUT + lane;
        const int il = i < L.jpi + 1 ? i : L.jpi + 1;  // clamped column for loads
        const bool own = lane < MOUT && i <= L.jpi;
        const bool tcol = i >= 1 && i <= L.jpi;        // column holds T points
        const int jb = by * M16;
        const int je = jb + M16 - 1 < L.jpj ? jb + M16 - 1 : L.jpj;
        StUniform st{sc_};
        RegWin w;
        w.s = sc_;
        w.load(1, n, L.at(il - 1, jb - 1));
        w.load(2, n, L.at(il - 1, jb));
        w.load_next(n, L.at(il - 1, jb + 1));
        unsigned f_nxt = fl[L.at(il, jb)];
        double s_prev = 0.0;   // new ssh at (i, j-1)
        double sv_prev = 0.0;  // old sshn_v at (i, j-1)
        unsigned f_prev = 0;
        for (int j = jb; j <= je; ++j) {
            const long o = L.at(il, j);
            // rows j-1..j+1 from the window and the prefetched row; the loads
            // of row j+2 (and the flags of row j+1) have a whole iteration to
            // arrive
            w.shift_in_next();
            const unsigned f = f_nxt;
            const int jn = j + 2 <= L.jpj + 2 ? j + 2 : L.jpj + 2;
            w.load_next(n, L.at(il - 1, jn));
            f_nxt = fl[L.at(il, j + 1 <= L.jpj + 1 ? j + 1 : L.jpj + 1)];
            double s_c = 0.0;
            if (j >= 1 && tcol) s_c = (f & T_CLAMP) ? tide : continuity_w(w, p.rdt);
            const double s_e = lane_next(s_c);
            if (own) {
                if (j >= 1) {
                    if (i >= 1) {
                        a.sshn[o] = s_c;
                        ct_ += fabs(s_c);
                    }
                    const unsigned c = ucls(f);
                    if (c != CLS_FLATHER) {  // Flather points: the epilogue
                        const double u = c == CLS_COMP ? momentum_u_w(w, f, p) : 0.0;
                        a.un[o] = u;
                        if (i >= 1) cu_ += fabs(u);
                    }
                    a.sshn_u[o] = next_su(st, sucls(f), o, s_c, s_e, w.su_[1][1]);
                }
                if (i >= 1) {
                    const unsigned c = vcls(f);
                    if (c != CLS_FLATHER) {  // Flather points: the epilogue
                        const double v = c == CLS_COMP ? momentum_v_w(w, f, p) : 0.0;
                        a.vn[o] = v;
                        if (j >= 1) cv_ += fabs(v);
                    }
                    if (j > jb) a.sshn_v[o - P] = next_sv(st, svcls(f_prev), o - P, P, s_prev, s_c, sv_prev);
                }
            }
            s_prev = s_c;
            sv_prev = w.sv_[1][1];
            f_prev = f;
        }
        // the last row's sshn_v needs the new ssh one row up
        if (own && i >= 1) {
            const long o = L.at(il, je);
            const unsigned c = svcls(f_prev);
            double s_n = 0.0;
            if (je + 1 <= L.jpj && (c == NX_AVG || c == NX_HI)) {
                const long on = o + P;
                s_n = (fl[on] & T_CLAMP) ? tide : continuity_at(st, n, on, P, p.rdt);
            }
            a.sshn_v[o] = next_sv(st, c, o, P, s_prev, s_n, sv_prev);
        }
    }
    block_reduce3(DD{ct_, 0.0}, DD{cu_, 0.0}, DD{cv_, 0.0}, partials + 6 * (long)blockIdx.x);
}

// sum|sshn|, sum|un|, sum|vn| over (1:jpi, 1:jpj) of a state of element
// types TS (ssh) and TV (the rest).
template <class TS, class TV>
__global__ void __launch_bounds__(TBX * TBY)
k_checksum_partials_t(const TS* __restrict__ sshn, const TV* __restrict__ un,
                      const TV* __restrict__ vn, Layout L, double* __restrict__ partials) {
    const int i = blockIdx.x * TBX + threadIdx.x;
    const int j = blockIdx.y * TBY + threadIdx.y;
    DD ct = {0.0, 0.0}, cu = {0.0, 0.0}, cv = {0.0, 0.0};
    if (in_rect(i, j, 1, L.jpi, 1, L.jpj)) {
        const long o = L.at(i, j);
        dd_add(ct, fabs((double)sshn[o]));
        dd_add(cu, fabs((double)un[o]));
        dd_add(cv, fabs((double)vn[o]));
    }
    block_reduce3(ct, cu, cv, partials + 6 * (long)(blockIdx.x + gridDim.x * blockIdx.y));
}

// Volume and energy diagnostics in FP64 over the T points:
//   volume = sum e12t*ssh
//   energy = sum e12t*(0.5*(ht+ssh)*(ubar^2+vbar^2) + 0.5*g*ssh^2)
// with ubar, vbar the averages of the two faces of the cell.  The third
// accumulator carries sum e12t (the wet area).
template <class TS, class TV, class TH>
__global__ void __launch_bounds__(TBX * TBY)
k_diag_partials(const TS* __restrict__ sshn, const TV* __restrict__ un, const TV* __restrict__ vn,
                const TH* __restrict__ ht, double ht0, const double* __restrict__ e12t, double e12t0,
                Layout L, double* __restrict__ partials) {
    const int i = blockIdx.x * TBX + threadIdx.x;
    const int j = blockIdx.y * TBY + threadIdx.y;
    DD a0 = {0.0, 0.0}, a1 = {0.0, 0.0}, a2 = {0.0, 0.0};
    if (in_rect(i, j, 1, L.jpi, 1, L.jpj)) {
        const long o = L.at(i, j);
        const double ar = e12t ? e12t[o] : e12t0;
        const double h = ht ? (double)ht[o] : ht0;
        const double z = (double)sshn[o];
        const double ub = 0.5 * ((double)un[o - 1] + (double)un[o]);
        const double vb = 0.5 * ((double)vn[o - L.pitch] + (double)vn[o]);
        dd_add(a0, ar * z);
        dd_add(a1, ar * (0.5 * (h + z) * (ub * ub + vb * vb) + 0.5 * GRAV * z * z));
        dd_add(a2, ar);
    }
    block_reduce3(a0, a1, a2, partials + 6 * (long)(blockIdx.x + gridDim.x * blockIdx.y));
}

// ---- step epilogue -------------------------------------------------------------
// One block.  Phase 1: the Flather points of the fused step (bc_flather_u/v,
// original/nemolite2d.f90:669-695) from the new after-velocities of their
// inner neighbours and the old sshn_u/sshn_v, one thread per line in loop
// order; their |values| inside (1:jpi,1:jpj) go to partial slot nparts.
// Phase 2 (when out is given): the fixed-order reduction of the nparts+1
// partials into the step's checksums.
__global__ void __launch_bounds__(RED_THREADS)
k_epilogue(const int* __restrict__ uptr, const long* __restrict__ upts, int nul,
           const int* __restrict__ vptr, const long* __restrict__ vpts, int nvl,
           const uint16_t* __restrict__ fl, const double* __restrict__ flu,
           const double* __restrict__ flv, const double* __restrict__ sshn_u,
           const double* __restrict__ sshn_v, double* __restrict__ un, double* __restrict__ vn,
           Layout L, double* __restrict__ partials, int nparts, double* __restrict__ out,
           int skip_v0) {
    const long P = L.pitch;
    DD cu = {0.0, 0.0}, cv = {0.0, 0.0};
    for (int line = threadIdx.x; line < nul; line += RED_THREADS)
        for (int k = uptr[line]; k < uptr[line + 1]; ++k) {
            const long o = upts[k];
            const long oi = (fl[o] & U_FL_EAST) ? o + 1 : o - 1;
            const double u = un[oi] + flu[o] * (sshn_u[o] - sshn_u[oi]);
            un[o] = u;
            if ((o - L.base) - (long)(line + 1) * P >= 1) dd_add(cu, fabs(u));  // i >= 1
        }
    for (int line = threadIdx.x; line < nvl; line += RED_THREADS)
        for (int k = vptr[line]; k < vptr[line + 1]; ++k) {
            const long o = vpts[k];
            // row-0 points with the inner point north: done by the step kernel
            if (skip_v0 && (fl[o] & V_FL_NORTH) && (o - L.base) / P == 0) continue;
            const long oi = (fl[o] & V_FL_NORTH) ? o + P : o - P;
            const double v = vn[oi] + flv[o] * (sshn_v[o] - sshn_v[oi]);
            vn[o] = v;
            if ((o - L.base) / P >= 1) dd_add(cv, fabs(v));  // j >= 1
        }
    block_reduce3(DD{0.0, 0.0}, cu, cv, partials + 6 * (long)nparts);
    __syncthreads();
    if (out) finish_partials(partials, nparts + 1, out);
}

// ---- fused step, column-marching with shared face terms ----------------------------
// Same work split and results as k_step_march (uniform metrics and depths),
// with fewer operations per point:
//  * terms on a face shared by two points of one column are computed once,
//    at the lower point as its north term, and carried up a row as the
//    upper point's south term (continuity's v flux, momentum_u's south
//    advective flux, v average and du/dy quotient, momentum_v's south
//    advective flux and dv/dy);
//  * terms on a face shared by two points of one row are computed once, by
//    the lane of the east point as its west term, and handed to the west
//    lane by a lane shuffle (momentum_u's east advective flux and du/dx,
//    momentum_v's east advective flux, u average and dv/dx quotient);
//  * the register window is a ring of four rows (three in use, one being
//    loaded) whose roles rotate with the row, so no row is ever copied.
// Every shared term is the same expression tree as the reference's own (an
// IEEE sum a+b equals b+a), so the results are bit-identical.  At a wall the
// shared advective term differs from the walled point's, which recomputes
// its own.
struct Row5 {
    double un[3], vn[3], ss[3], su[3], sv[3];  // columns i-1, i, i+1
};
// row at byte offset boff of column i-1
__device__ __forceinline__ void load_row5(Row5& r, const NState& n, unsigned boff) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        r.un[c] = *(const double*)((const char*)n.un + boff + 8 * c);
        r.vn[c] = *(const double*)((const char*)n.vn + boff + 8 * c);
        r.ss[c] = *(const double*)((const char*)n.sshn + boff + 8 * c);
        r.su[c] = *(const double*)((const char*)n.sshn_u + boff + 8 * c);
        r.sv[c] = *(const double*)((const char*)n.sshn_v + boff + 8 * c);
    }
}

__device__ __forceinline__ Col5 col5(const Row5& r, int c) {
    return Col5{r.un[c], r.vn[c], r.ss[c], r.su[c], r.sv[c]};
}
__device__ __forceinline__ NorthTerms north_terms(const Scal& s, const Row5& c, const Row5& nr) {
    return north_terms1(s, col5(c, 1), c.vn[2], c.sv[2], col5(nr, 1));
}

struct ShCarry {
    NorthTerms south;          // the row below's north terms
    double s_prev, sv_prev;    // new ssh and old sshn_v at (i, j-1)
    unsigned f_prev, f_nxt;
    double ct, cu, cv;         // per-lane checksum sums
};


// One row j: A, B, C hold rows j-1, j, j+1; C is loaded here.
__device__ __forceinline__ void sh_row(const ShCtx& x, ShCarry& k, const Row5& A, const Row5& B,
                                       Row5& C, int j) {
    const Layout& L = x.L;
    const Scal& s = x.s;
    const long P = L.pitch;
    const double h = 0.5;
    const long o = L.at(x.il, j);
    const unsigned f = k.f_nxt;
    load_row5(C, x.n, x.rowb(j + 1) + x.lb8);
    {
        const int jf = j + 1 <= L.jpj + 1 ? j + 1 : L.jpj + 1;
        k.f_nxt = ldb(x.fl, (x.rowb(jf) + x.lb8 + 8) >> 2);
    }
    const unsigned o8 = x.rowb(j) + x.lb8 + 8;  // byte offset of (i, j)
    const NorthTerms nt = north_terms(s, B, C);
    const NorthTerms& st = k.south;

    // continuity (original/nemolite2d.f90:399-403)
    double s_c = 0.0;
    if (j >= 1 && x.tcol) {
        if (f & T_CLAMP) {
            s_c = x.tide;
        } else {
            const double r1 = (B.su[1] + s.hu) * B.un[1];
            const double r2 = (B.su[0] + s.hu) * B.un[0];
            s_c = B.ss[1] + (r2 - r1 + st.r3 - nt.r3) * x.p.rdt / s.e12t;
        }
    }
    const double s_e = lane_next(s_c);

    // west-face terms of this lane's points; the east ones come from lane+1
    const double un_c = B.un[1];
    const double u_w = h * (un_c + B.un[0]) * s.e2t;
    const double depw_u = s.ht + B.ss[1];
    const double sw_u = sgn_half(u_w);
    const double fw = ((h - sw_u) * un_c + (h + sw_u) * B.un[0]) * u_w * depw_u;
    const double dudx_w = (un_c - B.un[0]) / s.e1t * (s.ht + B.ss[1]);
    const double vn_c = B.vn[1];
    const double u_wc = h * (B.un[0] + C.un[0]);
    const double u_wv = h * u_wc * (s.e2u + s.e2u);
    const double depw_v = h * (s.hu + B.su[0] + s.hu + C.su[0]);
    const double sw_v = sgn_half(u_wv);
    const double gw = ((h - sw_v) * vn_c + (h + sw_v) * B.vn[0]) * u_wv * depw_v;
    // dvdx_w's quotient; lane+1's is dvdx_e's
    const double qdvdx_w = (vn_c - B.vn[0]) / (s.e1v + s.e1v);
    const double fe = lane_next(fw);
    const double dudx_e = lane_next(dudx_w);
    const double ge = lane_next(gw);
    const double u_ec = lane_next(u_wc);
    const double qdvdx_e = lane_next(qdvdx_w);

    const double su_old = B.su[1], sv_old = B.sv[1];
    if (x.own) {
        const int i = x.i;
        double u = 0.0, v = 0.0;
        const unsigned cu = ucls(f), cv = vcls(f);
        if (j >= 1 && cu == CLS_COMP) {
            const bool ws = (f & U_WALL_S) != 0, wn = (f & U_WALL_N) != 0;
            double ts = st.kn, tn = nt.kn;
            if (ws) ts = ((h - sgn_half(st.vnu)) * un_c) * st.vnu * st.depnu;
            if (wn) tn = ((h + sgn_half(nt.vnu)) * un_c) * nt.vnu * nt.depnu;
            const double adv = fw - fe + ts - tn;
            double dudy_s = 0.0, dudy_n = 0.0;
            if (!ws) dudy_s = st.qdudy * (s.hu + B.su[1] + s.hu + A.su[1]);
            if (!wn) dudy_n = nt.qdudy * (s.hu + B.su[1] + s.hu + C.su[1]);
            double vis = (dudx_e - dudx_w) * s.e2u + (dudy_n - dudy_s) * s.e1u * h;
            vis = x.p.visc * vis;
            const double hsu = s.hu + B.su[1];
            const double cor = h * (s.fcu * (st.vnc + nt.vnc)) * s.e12u * hsu;
            const double hpg = -GRAV * hsu * s.e2u * (B.ss[2] - B.ss[1]);
            u = (un_c * hsu + x.p.rdt * (adv + vis + cor + hpg) / s.e12u) / (s.hu + 0.0) / x.p.fric;
        }
        if (i >= 1 && cv == CLS_COMP) {
            const bool ww = (f & V_WALL_W) != 0, we = (f & V_WALL_E) != 0;
            double tw = gw, te = ge;
            if (ww) tw = ((h - sw_v) * vn_c) * u_wv * depw_v;
            if (we) {
                const double u_e = h * u_ec * (s.e2u + s.e2u);
                const double depe = h * (s.hu + B.su[1] + s.hu + C.su[1]);
                te = ((h + sgn_half(u_e)) * vn_c) * u_e * depe;
            }
            const double adv = tw - te + st.hn - nt.hn;
            double dvdx_w = 0.0, dvdx_e = 0.0;
            if (!ww) dvdx_w = qdvdx_w * (s.hv + B.sv[1] + s.hv + B.sv[0]);
            if (!we) dvdx_e = qdvdx_e * (s.hv + B.sv[1] + s.hv + B.sv[2]);
            double vis = (nt.dvdyn - st.dvdyn) * s.e1v + (dvdx_e - dvdx_w) * s.e2v * h;
            vis = x.p.visc * vis;
            const double hsv = s.hv + B.sv[1];
            const double cor = -h * (s.fcv * (u_ec + u_wc)) * s.e12v * hsv;
            const double hpg = -GRAV * hsv * s.e1v * (C.ss[1] - B.ss[1]);
            v = (vn_c * hsv + x.p.rdt * (adv + vis + cor + hpg) / s.e12v) / (s.hv + 0.0) / x.p.fric;
        }
        const StUniform stu{s};
        if (j >= 1) {
            if (i >= 1) {
                stb(x.a.sshn, o8, s_c);
                k.ct += fabs(s_c);
            }
            if (cu != CLS_FLATHER) {
                stb(x.a.un, o8, u);
                if (i >= 1) k.cu += fabs(u);
            }
            stb(x.a.sshn_u, o8, next_su(stu, sucls(f), o, s_c, s_e, su_old));
        }
        if (i >= 1) {
            if (cv != CLS_FLATHER) {
                stb(x.a.vn, o8, v);
                if (j >= 1) k.cv += fabs(v);
            }
            // bc_flather_v (original/nemolite2d.f90:676-690) at a row-0 point
            // whose inner point is this one; row 0 is outside the checksums
            if (j == 1 && vcls(k.f_prev) == CLS_FLATHER && (k.f_prev & V_FL_NORTH)) {
                const unsigned o08 = o8 - (unsigned)(8 * P);
                stb(x.a.vn, o08, v + ldb(x.flv, o08) * (A.sv[1] - B.sv[1]));
            }
            if (k.f_prev != 0xffffffffu)
                stb(x.a.sshn_v, o8 - (unsigned)(8 * P),
                    next_sv(stu, svcls(k.f_prev), o - P, P, k.s_prev, s_c, k.sv_prev));
        }
    }
    k.south = nt;
    k.s_prev = s_c;
    k.sv_prev = sv_old;
    k.f_prev = f;
}

template <int NJ>
__global__ void __launch_bounds__(64 * MW)
k_step_march_sh(Scal sc_, NState n, AOut a, const uint16_t* __restrict__ fl,
                const double* __restrict__ flv, Layout L, Params p, double tide,
                double* __restrict__ partials, int nbx, int ntiles) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    ShCarry k;
    k.ct = k.cu = k.cv = 0.0;
    if (t < ntiles) {
        const int bx = t % nbx, by = t / nbx;
        ShCtx x;
        x.s = sc_;
        x.p = p;
        x.n = n;
        x.a = a;
        x.fl = fl;
        x.flv = flv;
        x.L = L;
        x.tide = tide;
        x.i = (bx * MW + wave) * MOUT + lane;
        x.il = x.i < L.jpi + 1 ? x.i : L.jpi + 1;  // clamped column for loads
        x.lb8 = (unsigned)(8 * (x.il - 1));
        x.own = lane < MOUT && x.i <= L.jpi;
        x.tcol = x.i >= 1 && x.i <= L.jpi;
        const int jb = by * NJ;
        const int je = jb + NJ - 1 < L.jpj ? jb + NJ - 1 : L.jpj;
        Row5 R0, R1, R2;
        load_row5(R0, n, x.rowb(jb - 1) + x.lb8);
        load_row5(R1, n, x.rowb(jb) + x.lb8);
        k.f_nxt = fl[L.at(x.il, jb)];
        k.f_prev = 0xffffffffu;  // no row below inside this segment
        k.s_prev = 0.0;
        k.sv_prev = 0.0;
        k.south = north_terms(sc_, R0, R1);
        for (int j = jb; j <= je; j += 3) {
            sh_row(x, k, R0, R1, R2, j);
            if (j + 1 > je) break;
            sh_row(x, k, R1, R2, R0, j + 1);
            if (j + 2 > je) break;
            sh_row(x, k, R2, R0, R1, j + 2);
        }
        // the last row's sshn_v needs the new ssh one row up
        if (x.own && x.i >= 1) {
            StUniform st{sc_};
            const long P = L.pitch;
            const long o = L.at(x.il, je);
            const unsigned c = svcls(k.f_prev);
            double s_n = 0.0;
            if (je + 1 <= L.jpj && (c == NX_AVG || c == NX_HI)) {
                const long on = o + P;
                s_n = (fl[on] & T_CLAMP) ? tide : continuity_at(st, n, on, P, p.rdt);
            }
            a.sshn_v[o] = next_sv(st, c, o, P, k.s_prev, s_n, k.sv_prev);
        }
    }
    block_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)blockIdx.x);
}

// ---- fused step, own-column marching (DPP neighbours) -----------------------------
// The shared-face-term algorithm above with each lane loading only its own
// column of the five state arrays per row: one 512-byte coalesced load per
// array and wave.  The west and east neighbours come from lanes l-1 and l+1
// by DPP wavefront shifts, so lanes 1..62 own the output columns and lanes 0
// and 63 are the west and east halos.  A row is 10 VGPRs instead of 30, so
// the ring of four rows keeps the loads of row j+2 in flight while row j is
// computed, within the register budget of four waves per SIMD.
constexpr int DOUT = 62;  // output columns per wave

__device__ __forceinline__ void load_col5(Col5F& r, const NState& n, const uint16_t* fl,
                                          unsigned boff) {
    r.un = ldb(n.un, boff);
    r.vn = ldb(n.vn, boff);
    r.ss = ldb(n.sshn, boff);
    r.su = ldb(n.sshn_u, boff);
    r.sv = ldb(n.sshn_v, boff);
    r.f = ldb(fl, boff >> 2);
}

struct DpCarry {
    NorthTerms south;         // the row below's north terms
    double s_prev, sv_prev;   // new ssh and old sshn_v at (i, j-1)
    double un_w, su_w;        // un and sshn_u one column west, this row
    unsigned f_prev;
    double ct, cu, cv;        // per-lane checksum sums
};

// next's interpolation (next_su / next_sv with uniform statics; E12 = QD_E12U
// or QD_E12V)
template <bool FD, int E12>
__device__ __forceinline__ double dp_next(const Scal& s, unsigned cls, double s_lo, double s_hi,
                                          double keep) {
    switch (cls) {
        case NX_AVG: {
            const double r = s.e12t * s_lo + s.e12t * s_hi;
            return udiv<FD, E12>(s, 0.5 * r, E12 == QD_E12U ? s.e12u : s.e12v);
        }
        case NX_HI: return s_hi;
        case NX_LO: return s_lo;
        default: return keep;
    }
}

// One row j: A, B, C hold rows j-1, j, j+1 (own column); D receives row j+2.
template <bool FD, bool BAR>
__device__ __forceinline__ void dp_row(const ShCtx& x, DpCarry& k, const Col5F& A, const Col5F& B,
                                       const Col5F& C, Col5F& D, int j) {
    if constexpr (BAR) __syncthreads();  // the block's waves start the row together
    const Layout& L = x.L;
    const Scal& s = x.s;
    const long P = L.pitch;
    const double h = 0.5;
    const unsigned f = B.f;
    {
        const int jn = j + 2 <= L.jpj + 2 ? j + 2 : L.jpj + 2;
        load_col5(D, x.n, x.fl, x.rowb(jn) + x.lb8);
    }
    const unsigned o8 = x.rowb(j) + x.lb8;  // byte offset of (i, j)
    // neighbours by lane shifts (every lane active here)
    const double bun_w = k.un_w, bsu_w = k.su_w;
    const double bvn_w = lane_prev(B.vn), bsv_w = lane_prev(B.sv);
    const double bvn_e = lane_next(B.vn), bsv_e = lane_next(B.sv), bss_e = lane_next(B.ss);
    const double cun_w = lane_prev(C.un), csu_w = lane_prev(C.su);
    const NorthTerms nt = north_terms1<FD>(s, B, bvn_e, bsv_e, C);
    const NorthTerms& st = k.south;

    // continuity (original/nemolite2d.f90:399-403)
    double s_c = 0.0;
    if (j >= 1 && x.tcol) {
        if (f & T_CLAMP) {
            s_c = x.tide;
        } else {
            const double r1 = (B.su + s.hu) * B.un;
            const double r2 = (bsu_w + s.hu) * bun_w;
            s_c = B.ss + udiv<FD, QD_E12T>(s, (r2 - r1 + st.r3 - nt.r3) * x.p.rdt, s.e12t);
        }
    }
    const double s_e = lane_next(s_c);

    // west-face terms of this lane's points; the east ones come from lane+1
    const double un_c = B.un;
    const double u_w = h * (un_c + bun_w) * s.e2t;
    const double depw_u = s.ht + B.ss;
    const double sw_u = sgn_half(u_w);
    const double fw = ((h - sw_u) * un_c + (h + sw_u) * bun_w) * u_w * depw_u;
    const double dudx_w = udiv<FD, QD_E1T>(s, un_c - bun_w, s.e1t) * (s.ht + B.ss);
    const double vn_c = B.vn;
    const double u_wc = h * (bun_w + cun_w);
    const double u_wv = h * u_wc * (s.e2u + s.e2u);
    const double depw_v = h * (s.hu + bsu_w + s.hu + csu_w);
    const double sw_v = sgn_half(u_wv);
    const double gw = ((h - sw_v) * vn_c + (h + sw_v) * bvn_w) * u_wv * depw_v;
    // dvdx_w's quotient; lane+1's is dvdx_e's
    const double qdvdx_w = udiv<FD, QD_E1V2>(s, vn_c - bvn_w, s.e1v + s.e1v);
    const double fe = lane_next(fw);
    const double dudx_e = lane_next(dudx_w);
    const double ge = lane_next(gw);
    const double u_ec = lane_next(u_wc);
    const double qdvdx_e = lane_next(qdvdx_w);

    const double su_old = B.su, sv_old = B.sv;
    if (x.own) {
        const int i = x.i;
        double u = 0.0, v = 0.0;
        const unsigned cu = ucls(f), cv = vcls(f);
        if (j >= 1 && cu == CLS_COMP) {
            const bool ws = (f & U_WALL_S) != 0, wn = (f & U_WALL_N) != 0;
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
            const double q = udiv<FD, QD_E12U>(s, x.p.rdt * (adv + vis + cor + hpg), s.e12u);
            u = udiv<FD, QD_FRIC>(s, udiv<FD, QD_HU>(s, un_c * hsu + q, s.hu + 0.0), x.p.fric);
        }
        if (i >= 1 && cv == CLS_COMP) {
            const bool ww = (f & V_WALL_W) != 0, we = (f & V_WALL_E) != 0;
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
            const double q = udiv<FD, QD_E12V>(s, x.p.rdt * (adv + vis + cor + hpg), s.e12v);
            v = udiv<FD, QD_FRIC>(s, udiv<FD, QD_HV>(s, vn_c * hsv + q, s.hv + 0.0), x.p.fric);
        }
        if (j >= 1) {
            if (i >= 1) {
                stb(x.a.sshn, o8, s_c);
                k.ct += fabs(s_c);
            }
            if (cu != CLS_FLATHER) {
                stb(x.a.un, o8, u);
                if (i >= 1) k.cu += fabs(u);
            }
            stb(x.a.sshn_u, o8, dp_next<FD, QD_E12U>(s, sucls(f), s_c, s_e, su_old));
        }
        if (i >= 1) {
            if (cv != CLS_FLATHER) {
                stb(x.a.vn, o8, v);
                if (j >= 1) k.cv += fabs(v);
            }
            // bc_flather_v (original/nemolite2d.f90:676-690) at a row-0 point
            // whose inner point is this one; row 0 is outside the checksums
            if (j == 1 && vcls(k.f_prev) == CLS_FLATHER && (k.f_prev & V_FL_NORTH)) {
                const unsigned o08 = o8 - (unsigned)(8 * P);
                stb(x.a.vn, o08, v + ldb(x.flv, o08) * (A.sv - B.sv));
            }
            if (k.f_prev != 0xffffffffu)
                stb(x.a.sshn_v, o8 - (unsigned)(8 * P),
                    dp_next<FD, QD_E12V>(s, svcls(k.f_prev), k.s_prev, s_c, k.sv_prev));
        }
    }
    k.south = nt;
    k.s_prev = s_c;
    k.sv_prev = sv_old;
    k.f_prev = f;
    k.un_w = cun_w;
    k.su_w = csu_w;
}

template <int NJ, bool FD, bool BAR>
__global__ void __launch_bounds__(64 * MW)
k_step_march_dp(Scal sc_, NState n, AOut a, const uint16_t* __restrict__ fl,
                const double* __restrict__ flv, Layout L, Params p, double tide,
                double* __restrict__ partials, int nbx, int ntiles) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    DpCarry k;
    k.ct = k.cu = k.cv = 0.0;
    if (t < ntiles) {
        const int bx = t % nbx, by = t / nbx;
        ShCtx x;
        x.s = sc_;
        x.p = p;
        x.n = n;
        x.a = a;
        x.fl = fl;
        x.flv = flv;
        x.L = L;
        x.tide = tide;
        // lane 0 is column i0 - 1 (>= -1, inside the left padding)
        x.i = (bx * MW + wave) * DOUT + lane - 1;
        x.il = x.i < L.jpi + 1 ? x.i : L.jpi + 1;  // clamped column for loads
        x.lb8 = (unsigned)(8 * x.il);              // wraps for -1; sums stay in range
        x.own = lane >= 1 && lane <= DOUT && x.i <= L.jpi;
        x.tcol = x.i >= 1 && x.i <= L.jpi;
        const int jb = by * NJ;
        const int je = jb + NJ - 1 < L.jpj ? jb + NJ - 1 : L.jpj;
        Col5F R0, R1, R2, R3;
        load_col5(R0, n, fl, x.rowb(jb - 1) + x.lb8);
        load_col5(R1, n, fl, x.rowb(jb) + x.lb8);
        load_col5(R2, n, fl, x.rowb(jb + 1) + x.lb8);
        k.f_prev = 0xffffffffu;  // no row below inside this segment
        k.s_prev = 0.0;
        k.sv_prev = 0.0;
        k.south = north_terms1<FD>(sc_, R0, lane_next(R0.vn), lane_next(R0.sv), R1);
        k.un_w = lane_prev(R1.un);
        k.su_w = lane_prev(R1.su);
        for (int j = jb; j <= je; j += 4) {
            dp_row<FD, BAR>(x, k, R0, R1, R2, R3, j);
            if (j + 1 > je) break;
            dp_row<FD, BAR>(x, k, R1, R2, R3, R0, j + 1);
            if (j + 2 > je) break;
            dp_row<FD, BAR>(x, k, R2, R3, R0, R1, j + 2);
            if (j + 3 > je) break;
            dp_row<FD, BAR>(x, k, R3, R0, R1, R2, j + 3);
        }
        // the last row's sshn_v needs the new ssh one row up
        if (x.own && x.i >= 1) {
            StUniform st{sc_};
            const long P = L.pitch;
            const long o = L.at(x.il, je);
            const unsigned c = svcls(k.f_prev);
            double s_n = 0.0;
            if (je + 1 <= L.jpj && (c == NX_AVG || c == NX_HI)) {
                const long on = o + P;
                s_n = (fl[on] & T_CLAMP) ? tide : continuity_at(st, n, on, P, p.rdt);
            }
            a.sshn_v[o] = next_sv(st, c, o, P, k.s_prev, s_n, k.sv_prev);
        }
    }
    block_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)blockIdx.x);
}

// A wave whose tile the table marks plain (every point it touches, halos
// included, has the class word RC_PLAIN) runs the class-free code.
template <int DM>
__global__ void __launch_bounds__(64 * MW)
k_step_march_rc(Scal sc_, NState n, AOut a, const uint16_t* __restrict__ fl,
                const double* __restrict__ flv, Layout L, Params p, double tide,
                double* __restrict__ partials, int nbx, int ntiles,
                const uint8_t* __restrict__ plain, int nj) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    RcCarry k;
    k.ct = k.cu = k.cv = 0.0;
    if (t < ntiles) {
        const int bx = t % nbx, by = t / nbx;
        RcCtx x;
        x.s = sc_;
        x.p = p;
        x.n = n;
        x.a = a;
        x.fl = fl;
        x.flv = flv;
        x.L = L;
        x.tide = tide;
        // lane 0 is column i0 - 1 (>= -1, inside the left padding)
        x.i = (bx * MW + wave) * ROUT + lane - 1;
        x.il = x.i < L.jpi + 1 ? x.i : L.jpi + 1;  // clamped column for loads
        x.lb8 = (unsigned)(8 * x.il);              // wraps for -1; sums stay in range
        x.own = lane >= 1 && lane <= ROUT && x.i <= L.jpi;
        x.tcol = x.i >= 1 && x.i <= L.jpi;
        x.ucol = x.i >= 0 && x.i <= L.jpi;
        x.vcol = x.i >= 1 && x.i <= L.jpi;
        const int jb = by * nj;  // (segments of nj rows: the plain table's)
        const int je = jb + nj - 1 < L.jpj ? jb + nj - 1 : L.jpj;
        const int nseg = ntiles / nbx;
        // (tiles that are not plain divide by IEEE division: their general
        // code with the guard's bookkeeping would not fit three waves per SIMD)
        if (plain[(long)(bx * MW + wave) * nseg + by])
            rc_march<DM, true>(x, k, jb, je);
        else
            rc_march<0, false>(x, k, jb, je);
    }
    block_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)blockIdx.x);
}

// ---- streaming reference ---------------------------------------------------------
// Reads five arrays and writes five, elementwise: the fused step's algorithmic
// traffic without its stencil, for the practical bandwidth ceiling.
struct Stream5 {
    const double* in[5];
    double* out[5];
};
__global__ void __launch_bounds__(256) k_stream5(Stream5 s, long n) {
    const long stride = (long)gridDim.x * 256;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
#pragma unroll
        for (int k = 0; k < 5; ++k) s.out[k][i] = s.in[k][i] * 1.0000001;
    }
}

// ---- checksum ----------------------------------------------------------------
// sum|sshn|, sum|un|, sum|vn| over (1:jpi, 1:jpj) of the given state.
__global__ void __launch_bounds__(TBX * TBY)
k_checksum_partials(const double* __restrict__ sshn, const double* __restrict__ un,
                    const double* __restrict__ vn, Layout L, double* __restrict__ partials) {
    const int i = blockIdx.x * TBX + threadIdx.x;
    const int j = blockIdx.y * TBY + threadIdx.y;
    DD ct = {0.0, 0.0}, cu = {0.0, 0.0}, cv = {0.0, 0.0};
    if (in_rect(i, j, 1, L.jpi, 1, L.jpj)) {
        const long o = L.at(i, j);
        dd_add(ct, fabs(sshn[o]));
        dd_add(cu, fabs(un[o]));
        dd_add(cv, fabs(vn[o]));
    }
    block_reduce3(ct, cu, cv, partials + 6 * (long)(blockIdx.x + gridDim.x * blockIdx.y));
}

// Finish: reduce nparts block partials in a fixed order into out[0..2].
__global__ void __launch_bounds__(RED_THREADS)
k_checksum_finish(const double* __restrict__ partials, int nparts, double* __restrict__ out) {
    finish_partials(partials, nparts, out);
}

// Pending logged steps: block b finishes slot pend[2b] from its pend[2b+1]
// partials.
__global__ void __launch_bounds__(RED_THREADS)
k_checksum_finish_batch(const double* __restrict__ part, long slot_doubles,
                        const int* __restrict__ pend, double* __restrict__ log) {
    const int slot = pend[2 * blockIdx.x], np = pend[2 * blockIdx.x + 1];
    finish_partials(part + slot * slot_doubles, np, log + 3 * (long)slot);
}

// ---- stage kernels (one thread per point of the stage's Fortran range) ------
template <class S>
__global__ void k_continuity(S st, NState n, double* __restrict__ ssha, Layout L, double rdt) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (!in_rect(i, j, 1, L.jpi, 1, L.jpj)) return;
    const long o = L.at(i, j);
    ssha[o] = continuity_at(st, n, o, L.pitch, rdt);
}

template <class S>
__global__ void k_momentum_u(S st, NState n, const uint16_t* __restrict__ fl,
                             double* __restrict__ ua, Layout L, Params p) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (!in_rect(i, j, 1, L.jpi - 1, 1, L.jpj)) return;
    const long o = L.at(i, j);
    const unsigned f = fl[o];
    if (ucls(f) == CLS_COMP) ua[o] = momentum_u_at(st, n, o, L.pitch, f, p);
}

template <class S>
__global__ void k_momentum_v(S st, NState n, const uint16_t* __restrict__ fl,
                             double* __restrict__ va, Layout L, Params p) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (!in_rect(i, j, 1, L.jpi, 1, L.jpj - 1)) return;
    const long o = L.at(i, j);
    const unsigned f = fl[o];
    if (vcls(f) == CLS_COMP) va[o] = momentum_v_at(st, n, o, L.pitch, f, p);
}

__global__ void k_bc_ssh(const uint16_t* __restrict__ fl, double* __restrict__ ssha, Layout L,
                         double tide) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (!in_rect(i, j, 1, L.jpi, 1, L.jpj)) return;
    const long o = L.at(i, j);
    if (fl[o] & T_CLAMP) ssha[o] = tide;
}

__global__ void k_bc_solid_u(const uint16_t* __restrict__ fl, double* __restrict__ ua, Layout L) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (!in_rect(i, j, 0, L.jpi, 1, L.jpj)) return;
    const long o = L.at(i, j);
    if (ucls(fl[o]) == CLS_SOLID) ua[o] = 0.0;
}

__global__ void k_bc_solid_v(const uint16_t* __restrict__ fl, double* __restrict__ va, Layout L) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (!in_rect(i, j, 1, L.jpi, 0, L.jpj)) return;
    const long o = L.at(i, j);
    if (vcls(fl[o]) == CLS_SOLID) va[o] = 0.0;
}

// Flather: one thread per line walks that line's points in the serial
// model's loop order, so chained points see exactly the values it would.
__global__ void k_bc_flather_u(const int* __restrict__ ptr, const long* __restrict__ pts,
                               int nlines, const uint16_t* __restrict__ fl,
                               const double* __restrict__ flu, const double* __restrict__ sshn_u,
                               double* __restrict__ ua) {
    const int line = blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nlines) return;
    for (int k = ptr[line]; k < ptr[line + 1]; ++k) {
        const long o = pts[k];
        const long oi = (fl[o] & U_FL_EAST) ? o + 1 : o - 1;
        ua[o] = ua[oi] + flu[o] * (sshn_u[o] - sshn_u[oi]);
    }
}

__global__ void k_bc_flather_v(const int* __restrict__ ptr, const long* __restrict__ pts,
                               int nlines, const uint16_t* __restrict__ fl,
                               const double* __restrict__ flv, const double* __restrict__ sshn_v,
                               double* __restrict__ va, long P) {
    const int line = blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= nlines) return;
    for (int k = ptr[line]; k < ptr[line + 1]; ++k) {
        const long o = pts[k];
        const long oi = (fl[o] & V_FL_NORTH) ? o + P : o - P;
        va[o] = va[oi] + flv[o] * (sshn_v[o] - sshn_v[oi]);
    }
}

// next, original/nemolite2d.f90:704-769: un=ua, vn=va, sshn=ssha, then the
// interpolations from the new ssh (read from ssha, which equals it).  With
// CS the block also sums |sshn|, |un|, |vn| of the new state over
// (1:jpi, 1:jpj) into its checksum partial.
template <class S, bool CS>
__global__ void __launch_bounds__(TBX * TBY)
k_next(S st, const uint16_t* __restrict__ fl, const double* __restrict__ ssha,
       const double* __restrict__ ua, const double* __restrict__ va, double* __restrict__ sshn,
       double* __restrict__ sshn_u, double* __restrict__ sshn_v, double* __restrict__ un,
       double* __restrict__ vn, Layout L, double* __restrict__ partials) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    DD ct = {0.0, 0.0}, cu = {0.0, 0.0}, cv = {0.0, 0.0};
    if (i <= L.jpi && j <= L.jpj) {
        const long o = L.at(i, j), P = L.pitch;
        const unsigned f = fl[o];
        const double s = ssha[o];
        if (i >= 1 && j >= 1) {
            sshn[o] = s;
            if (CS) dd_add(ct, fabs(s));
        }
        if (j >= 1) {
            const double u = ua[o];
            un[o] = u;
            if (CS && i >= 1) dd_add(cu, fabs(u));
            const unsigned c = sucls(f);
            if (c != NX_SKIP) sshn_u[o] = next_su(st, c, o, s, ssha[o + 1], 0.0);
        }
        if (i >= 1) {
            const double v = va[o];
            vn[o] = v;
            if (CS && j >= 1) dd_add(cv, fabs(v));
            const unsigned c = svcls(f);
            if (c != NX_SKIP) sshn_v[o] = next_sv(st, c, o, P, s, ssha[o + P], 0.0);
        }
    }
    if (CS) block_reduce3(ct, cu, cv, partials + 6 * (long)(blockIdx.x + gridDim.x * blockIdx.y));
}

// ---- learned closure on the device state ------------------------------------------
// Pixel p = (j-1)*jpi + (i-1) stands for T cell (i, j) with its east face U(i, j)
// and north face V(i, j).
struct Norm4 {
    double m[4], s[4];
};
struct Scale3 {
    double s[3];
};

template <class S>
__global__ void __launch_bounds__(TBX * TBY)
k_closure_features(S st, const double* __restrict__ sshn, const double* __restrict__ un,
                   const double* __restrict__ vn, Layout L, Norm4 nm, uint16_t* __restrict__ x) {
    const int i = blockIdx.x * TBX + threadIdx.x + 1, j = blockIdx.y * TBY + threadIdx.y + 1;
    if (i > L.jpi || j > L.jpj) return;
    const long o = L.at(i, j);
    const long p = (long)(j - 1) * L.jpi + (i - 1);
    const double v[4] = {sshn[o], un[o], vn[o], st.ht(o)};
    uint16_t* xp = x + 8 * p;
#pragma unroll
    for (int c = 0; c < 4; ++c) xp[c] = bf16_rne((float)((v[c] - nm.m[c]) / nm.s[c]));
#pragma unroll
    for (int c = 4; c < 8; ++c) xp[c] = 0;
}

// Increments: sshn at every T cell, un on the interior east faces
// (i < jpi), vn on the interior north faces (j < jpj).
__global__ void __launch_bounds__(TBX * TBY)
k_closure_add(const float* __restrict__ z, int zs, Scale3 sc, Layout L, double* __restrict__ sshn,
              double* __restrict__ un, double* __restrict__ vn) {
    const int i = blockIdx.x * TBX + threadIdx.x + 1, j = blockIdx.y * TBY + threadIdx.y + 1;
    if (i > L.jpi || j > L.jpj) return;
    const long o = L.at(i, j);
    const float* zp = z + (long)zs * ((long)(j - 1) * L.jpi + (i - 1));
    sshn[o] = sshn[o] + (double)zp[0] * sc.s[0];
    if (i < L.jpi) un[o] = un[o] + (double)zp[1] * sc.s[1];
    if (j < L.jpj) vn[o] = vn[o] + (double)zp[2] * sc.s[2];
}

// Restriction of a fine state onto the 2x coarser grid (coarse cell (I, J)
// covers fine cells 2I-1..2I x 2J-1..2J): ssh the mean of the 4 fine cells,
// u on a coarse U face the depth-weighted mean of the 2 fine U faces on it,
// v likewise.  One thread per coarse point (I, J), I in 0..jpi, J in 0..jpj.
template <class H>
__global__ void __launch_bounds__(TBX * TBY)
k_restrict(const double* __restrict__ fs, const double* __restrict__ fu, const double* __restrict__ fv,
           H hu, H hv, Layout F, Layout C, double* __restrict__ cs, double* __restrict__ cu,
           double* __restrict__ cv) {
    const int I = blockIdx.x * TBX + threadIdx.x, J = blockIdx.y * TBY + threadIdx.y;
    if (I > C.jpi || J > C.jpj) return;
    const long o = C.at(I, J);
    if (I >= 1 && J >= 1) {
        const long a = F.at(2 * I - 1, 2 * J - 1), b = a + 1, c = a + F.pitch, d = c + 1;
        cs[o] = 0.25 * (((fs[a] + fs[b]) + fs[c]) + fs[d]);
    }
    if (J >= 1) {  // U face (I, J): fine faces (2I, 2J-1) and (2I, 2J)
        const long a = F.at(2 * I, 2 * J - 1), b = a + F.pitch;
        const double h0 = hu(a), h1 = hu(b);
        cu[o] = (h0 * fu[a] + h1 * fu[b]) / (h0 + h1);
    }
    if (I >= 1) {  // V face (I, J): fine faces (2I-1, 2J) and (2I, 2J)
        const long a = F.at(2 * I - 1, 2 * J), b = a + 1;
        const double h0 = hv(a), h1 = hv(b);
        cv[o] = (h0 * fv[a] + h1 * fv[b]) / (h0 + h1);
    }
}

struct DepthArr {
    const double* p;
    __device__ double operator()(long o) const { return p[o]; }
};
struct DepthConst {
    double v;
    __device__ double operator()(long) const { return v; }
};

// Closure training target: y[p][c] = FP32((truth_c - pred_c) / scale[c]) for
// the cell's ssh (c = 0), east-face u (1) and north-face v (2); 0 for c >= 3.
__global__ void __launch_bounds__(TBX * TBY)
k_closure_target(const double* __restrict__ ts, const double* __restrict__ tu, const double* __restrict__ tv,
                 const double* __restrict__ ps, const double* __restrict__ pu, const double* __restrict__ pv,
                 Layout L, Scale3 sc, float* __restrict__ y) {
    const int i = blockIdx.x * TBX + threadIdx.x + 1, j = blockIdx.y * TBY + threadIdx.y + 1;
    if (i > L.jpi || j > L.jpj) return;
    const long o = L.at(i, j);
    float* yp = y + 8 * ((long)(j - 1) * L.jpi + (i - 1));
    yp[0] = (float)((ts[o] - ps[o]) / sc.s[0]);
    yp[1] = (float)((tu[o] - pu[o]) / sc.s[1]);
    yp[2] = (float)((tv[o] - pv[o]) / sc.s[2]);
#pragma unroll
    for (int c = 3; c < 8; ++c) yp[c] = 0.f;
}

// sshn_u, sshn_v from sshn by the next-stage rules (original/nemolite2d.f90:
// 735-762); the points next leaves alone keep their values.
template <class S>
__global__ void __launch_bounds__(TBX * TBY)
k_interp(S st, const uint16_t* __restrict__ fl, const double* __restrict__ sshn, Layout L,
                 double* __restrict__ sshn_u, double* __restrict__ sshn_v) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (i > L.jpi || j > L.jpj) return;
    const long o = L.at(i, j), P = L.pitch;
    const unsigned f = fl[o];
    if (j >= 1) {
        const unsigned c = sucls(f);
        if (c != NX_SKIP) sshn_u[o] = next_su(st, c, o, sshn[o], sshn[o + 1], 0.0);
    }
    if (i >= 1) {
        const unsigned c = svcls(f);
        if (c != NX_SKIP) sshn_v[o] = next_sv(st, c, o, P, sshn[o], sshn[o + P], 0.0);
    }
}

// Whether given sshn_u/sshn_v are next's interpolations of sshn, bit for bit,
// at every point next sets: any point that differs sets *differs.
template <class S>
__global__ void __launch_bounds__(TBX * TBY)
k_interp_check(S st, const uint16_t* __restrict__ fl, const double* __restrict__ sshn, Layout L,
               const double* __restrict__ sshn_u, const double* __restrict__ sshn_v,
               unsigned* __restrict__ differs) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (i > L.jpi || j > L.jpj) return;
    const long o = L.at(i, j), P = L.pitch;
    const unsigned f = fl[o];
    bool d = false;
    if (j >= 1) {
        const unsigned c = sucls(f);
        if (c != NX_SKIP)
            d = __double_as_longlong(sshn_u[o]) !=
                __double_as_longlong(next_su(st, c, o, sshn[o], sshn[o + 1], 0.0));
    }
    if (i >= 1) {
        const unsigned c = svcls(f);
        if (c != NX_SKIP)
            d = d || __double_as_longlong(sshn_v[o]) !=
                         __double_as_longlong(next_sv(st, c, o, P, sshn[o], sshn[o + P], 0.0));
    }
    if (d) atomicOr(differs, 1u);
}

// The same for the single-precision form's state (FP32 sshn_u/sshn_v from an
// FP64 or FP32 sshn, FP32 arithmetic and areas, as its step forms them).
template <class SS>
__global__ void __launch_bounds__(TBX * TBY)
k_interp_mixed(const uint16_t* __restrict__ fl, const SS* __restrict__ sshn, Layout L, float e12t,
               float e12u, float e12v, float* __restrict__ sshn_u, float* __restrict__ sshn_v) {
    const int i = blockIdx.x * TBX + threadIdx.x, j = blockIdx.y * TBY + threadIdx.y;
    if (i > L.jpi || j > L.jpj) return;
    const long o = L.at(i, j), P = L.pitch;
    const unsigned f = fl[o];
    if (j >= 1) {
        const unsigned c = sucls(f);
        if (c != NX_SKIP) sshn_u[o] = next_g<float>(c, (float)sshn[o], (float)sshn[o + 1], 0.f, e12t, e12u);
    }
    if (i >= 1) {
        const unsigned c = svcls(f);
        if (c != NX_SKIP) sshn_v[o] = next_g<float>(c, (float)sshn[o], (float)sshn[o + P], 0.f, e12t, e12v);
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
dim3 grid2d(int ni, int nj) { return dim3((ni + TBX - 1) / TBX, (nj + TBY - 1) / TBY); }



NState nstate(const nemo2d_model* m, int k) {
    NState n;
    n.sshn = m->S[k][F_SSHN];
    n.sshn_u = m->S[k][F_SSHN_U];
    n.sshn_v = m->S[k][F_SSHN_V];
    n.un = m->S[k][F_UN];
    n.vn = m->S[k][F_VN];
    return n;
}

StArrays starrays(const nemo2d_model* m) {
    StArrays s;
    s.a_e1t = m->st[S_E1T]; s.a_e2t = m->st[S_E2T];
    s.a_e1u = m->st[S_E1U]; s.a_e2u = m->st[S_E2U];
    s.a_e1v = m->st[S_E1V]; s.a_e2v = m->st[S_E2V];
    s.a_e12t = m->st[S_E12T]; s.a_e12u = m->st[S_E12U]; s.a_e12v = m->st[S_E12V];
    s.a_ht = m->st[S_HT]; s.a_hu = m->st[S_HU]; s.a_hv = m->st[S_HV];
    s.a_fcu = m->st[S_FCU]; s.a_fcv = m->st[S_FCV];
    return s;
}

// Dispatch a templated launch on the statics policy.
#define WITH_STATICS(m, STV, BODY)          \
    do {                                    \
        if ((m)->uniform) {                 \
            StUniform STV{(m)->scal};       \
            BODY;                           \
        } else {                            \
            StArrays STV = starrays(m);     \
            BODY;                           \
        }                                   \
    } while (0)

// Fill in sshn_u/sshn_v of the current state set when the last fused step
// left them unwritten (k_step_march_rc, k_step_march_gr): next's
// interpolation of its sshn, in the state's own arithmetic.
int uv_sync(nemo2d_model* m) {
    if (!m->uv_stale[m->cur]) return NEMO2D_OK;
    const Layout& L = m->L;
    const dim3 g = grid2d(L.jpi + 1, L.jpj + 1);
    if (m->mixed) {
        void** x = m->X[m->cur];
        const Scal& d = m->scal;
        if (m->compensated)
            k_interp_mixed<double><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                m->flags, (const double*)x[F_SSHN], L, (float)d.e12t, (float)d.e12u, (float)d.e12v,
                (float*)x[F_SSHN_U], (float*)x[F_SSHN_V]);
        else
            k_interp_mixed<float><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                m->flags, (const float*)x[F_SSHN], L, (float)d.e12t, (float)d.e12u, (float)d.e12v,
                (float*)x[F_SSHN_U], (float*)x[F_SSHN_V]);
    } else {
        double** s = m->S[m->cur];
        WITH_STATICS(m, st, (k_interp<<<g, dim3(TBX, TBY), 0, m->stream>>>(
                                st, m->flags, s[F_SSHN], L, s[F_SSHN_U], s[F_SSHN_V])));
    }
    LAUNCHCHK("interp");
    m->uv_stale[m->cur] = false;
    return NEMO2D_OK;
}


void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// Before work on m->stream writes now-state buffer `target` (0 or 1; -1:
// both): if a snapshot is still copying that buffer, wait for the copy.
int snap_guard(nemo2d_model* m, int target) {
    if (m->snap_src >= 0 && (target < 0 || target == m->snap_src)) {
        HIPCHK(hipStreamWaitEvent(m->stream, m->ev_copy, 0));
        m->snap_src = -1;
    }
    return NEMO2D_OK;
}

// Finish the snapshot thread (if any); returns its status.
int snap_join(nemo2d_model* m) {
    if (m->snap_thread.joinable()) m->snap_thread.join();
    const int rc = m->snap_rc;
    m->snap_rc = NEMO2D_OK;
    if (rc) return fail(rc, m->snap_err);
    return NEMO2D_OK;
}

// ---- region timers: Continuity, Momentum, BCs, Next (the serial model's
// dl_timer regions, original/nemolite2d.f90:391-769) and the fused step
enum { TR_CONT = 0, TR_MOM, TR_BCS, TR_NEXT, TR_FUSED, TR_N };
// a region opened by a stage call that was recorded for the fused pass
constexpr int TEV_DEFER = -2;

// fold the finished calls into the running sums and recycle their events
void timers_fold(nemo2d_model* m) {
    if (m->tcalls.empty()) return;
    (void)hipStreamSynchronize(m->stream);
    for (const auto& c : m->tcalls) {
        float ms = 0.f;  // a recorded (deferred) call did no GPU work of its own
        if (c.b != TEV_DEFER && hipEventElapsedTime(&ms, m->tev[c.b], m->tev[c.e]) != hipSuccess) continue;
        m->tcount[c.region] += 1;
        m->tsum[c.region] += ms;
        m->tsq[c.region] += (double)ms * ms;
    }
    m->tcalls.clear();
    m->tev_used = 0;
}

int timer_event(nemo2d_model* m) {
    if (m->tev_used == (int)m->tev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        m->tev.push_back(e);
    }
    const int k = m->tev_used++;
    (void)hipEventRecord(m->tev[k], m->stream);
    return k;
}

// deferred: the call is only recorded (no event; the region counts the
// call with no time unless a call that does run work ends it)
void region_begin(nemo2d_model* m, int r, bool deferred) {
    if (!m->timing) return;
    bool open = false;
    for (int k = 0; k < TR_N; ++k) open = open || m->topen[k] >= 0;
    if (!open && m->tev_used > 8192) timers_fold(m);
    m->topen[r] = deferred ? TEV_DEFER : timer_event(m);
}

void region_end(nemo2d_model* m, int r, bool deferred) {
    if (!m->timing || m->topen[r] == -1) return;
    if (deferred && m->topen[r] == TEV_DEFER) {
        m->tcalls.push_back({r, TEV_DEFER, TEV_DEFER});
    } else if (m->topen[r] >= 0) {
        const int e = timer_event(m);
        if (e >= 0) m->tcalls.push_back({r, m->topen[r], e});
    }
    m->topen[r] = -1;
}

// begin and/or end region r around the enclosing call
struct RegionScope {
    nemo2d_model* m;
    int r;
    bool end, deferred;
    RegionScope(nemo2d_model* m_, int r_, bool begin, bool end_, bool deferred_ = false)
        : m(m_), r(r_), end(end_), deferred(deferred_) {
        if (begin) {
            region_begin(m, r, deferred);
        } else if (!deferred && m->timing && m->topen[r] == TEV_DEFER) {
            m->topen[r] = timer_event(m);  // opened by a recorded call: time from here
        }
    }
    ~RegionScope() {
        if (end) region_end(m, r, deferred);
    }
};

// Forget logged steps whose checksums were not finished yet.
void drop_pending(nemo2d_model* m) {
    std::fill(m->cs_pend.begin(), m->cs_pend.end(), 0);
    m->pend_any = false;
}

// Allocate n doubles and zero them on stream s.  The zeroing is queued on
// the stream that later fills or reads the buffer: a fill on the null stream
// is not ordered against a stream created hipStreamNonBlocking, so a late
// null-stream zeroing could land after an upload made on the model's stream.
int dalloc(double** p, size_t n, hipStream_t s) {
    HIPCHK(hipMalloc((void**)p, n * sizeof(double)));
    HIPCHK(hipMemsetAsync(*p, 0, n * sizeof(double), s));
    return NEMO2D_OK;
}

// Host Fortran array (lb1:ub1, lb2:ub2) -> device layout.
// Host array (lb1:ub1, lb2:ub2) of E, contiguous, <-> device padded array
// of E, through the model's staging buffer (in stream order, so one buffer
// serves consecutive copies).
template <class E>
int to_device(const nemo2d_model* m, E* dst, const E* src, int lb1, int ub1, int lb2, int ub2) {
    const size_t w = (size_t)(ub1 - lb1 + 1) * sizeof(E), rows = (size_t)(ub2 - lb2 + 1);
    HIPCHK(hipMemcpyAsync(m->stage, src, w * rows, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpy2DAsync(dst + m->L.at(lb1, lb2), m->L.pitch * sizeof(E), m->stage, w, w, rows,
                            hipMemcpyDeviceToDevice, m->stream));
    return NEMO2D_OK;
}
template <class E>
int to_host(const nemo2d_model* m, E* dst, const E* src, int lb1, int ub1, int lb2, int ub2) {
    const size_t w = (size_t)(ub1 - lb1 + 1) * sizeof(E), rows = (size_t)(ub2 - lb2 + 1);
    HIPCHK(hipMemcpy2DAsync(m->stage, w, src + m->L.at(lb1, lb2), m->L.pitch * sizeof(E), w, rows,
                            hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(dst, m->stage, w * rows, hipMemcpyDeviceToHost, m->stream));
    return NEMO2D_OK;
}

int upload(const nemo2d_model* m, double* dst, const double* src, int lb1, int ub1, int lb2,
           int ub2) {
    return to_device<double>(m, dst, src, lb1, ub1, lb2, ub2);
}

int download(const nemo2d_model* m, double* dst, const double* src, int lb1, int ub1, int lb2,
             int ub2) {
    return to_host<double>(m, dst, src, lb1, ub1, lb2, ub2);
}

bool all_equal(const double* a, size_t n) {
    for (size_t k = 1; k < n; ++k)
        if (memcmp(&a[k], &a[0], sizeof(double)) != 0) return false;
    return true;
}

// Host Fortran array (lb1:ub1, lb2:ub2) of doubles -> device padded array
// of element type E (converted on the host).
template <class E>
int upload_as(const nemo2d_model* m, E* dst, const double* src, int lb1, int ub1, int lb2, int ub2) {
    const long n1 = ub1 - lb1 + 1;
    std::vector<E> rows((size_t)n1 * (ub2 - lb2 + 1));
    for (size_t k = 0; k < rows.size(); ++k) rows[k] = (E)src[k];
    const int rc = to_device<E>(m, dst, rows.data(), lb1, ub1, lb2, ub2);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));  // rows is a local buffer
    return NEMO2D_OK;
}

template <class E>
int download_as(const nemo2d_model* m, double* dst, const E* src, int lb1, int ub1, int lb2, int ub2) {
    const long n1 = ub1 - lb1 + 1;
    std::vector<E> rows((size_t)n1 * (ub2 - lb2 + 1));
    const int rc = to_host<E>(m, rows.data(), src, lb1, ub1, lb2, ub2);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    for (size_t k = 0; k < rows.size(); ++k) dst[k] = (double)rows[k];
    return NEMO2D_OK;
}

int checksum_into(nemo2d_model* m, int k, double* out) {
    const Layout& L = m->L;
    dim3 g = grid2d(L.jpi + 1, L.jpj + 1);
    if (m->mixed) {
        const float* un = (const float*)m->X[k][F_UN];
        const float* vn = (const float*)m->X[k][F_VN];
        if (m->compensated)
            k_checksum_partials_t<double, float><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                (const double*)m->X[k][F_SSHN], un, vn, L, m->partials);
        else
            k_checksum_partials_t<float, float><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                (const float*)m->X[k][F_SSHN], un, vn, L, m->partials);
    } else {
        k_checksum_partials<<<g, dim3(TBX, TBY), 0, m->stream>>>(
            m->S[k][F_SSHN], m->S[k][F_UN], m->S[k][F_VN], L, m->partials);
    }
    LAUNCHCHK("checksum_partials");
    k_checksum_finish<<<1, RED_THREADS, 0, m->stream>>>(m->partials, (int)(g.x * g.y), out);
    LAUNCHCHK("checksum_finish");
    return NEMO2D_OK;
}

}  // namespace

namespace nemo2d {

int set_error(int code, const std::string& msg) { return fail(code, msg); }

// Reciprocals of the uniform divisors (formed as the point functions form
// them) for the division-by-reciprocal path; off unless every divisor is in
// the range where that path is exact.
void set_reciprocals(Scal& s, double fric) {
    const double d[QD_N] = {s.e12t, s.e1t, s.e2t, s.e2u + s.e2u, s.e1v + s.e1v, s.e12u, s.e12v,
                            s.hu + 0.0, s.hv + 0.0, fric};
    bool ok = true, one = true, one_met = true;
    for (int k = 0; k < QD_N; ++k) {
        ok = ok && d[k] >= 0x1p-60 && d[k] <= 0x1p60;
        s.rcp[k] = 1.0 / d[k];
        // |d*RN(1/d) - 1| <= 2^-54 (the FMA evaluates it to within a relative
        // 2^-53; the margin covers that): RN(x*y) is then a faithful x/d
        const bool q = fabs(fma(d[k], s.rcp[k], -1.0)) <= 0x1p-54 * (1.0 - 0x1p-40);
        one = one && q;
        if (k != QD_HU && k != QD_HV) one_met = one_met && q;
    }
    s.fast_div = ok ? 1 : 0;
    s.one_corr = ok && one ? 1 : 0;
    s.one_corr_met = ok && one_met ? 1 : 0;
}


// Per-point classes of a mask (see nemo2d_core.h), the Flather coefficients
// sqrt(g/h) and the Flather point lists in the serial model's loop order.
// Rejects masks the serial model would index outside its arrays with.
int classify(const Layout& L, const int* pt, const double* hu, const double* hv, Classes& c,
             std::string& err) {
    const int jpi = L.jpi, jpj = L.jpj;
    const int pw = jpi + 2;
    auto PT = [&](int i, int j) { return pt[i + (long)j * pw]; };
    for (long k = 0; k < (long)pw * (jpj + 2); ++k)
        if (pt[k] < -1 || pt[k] > 1) {
            err = "mask values must be -1, 0 or 1";
            return NEMO2D_ERR_MASK;
        }
    auto U = [&](const double* a, int i, int j) { return a[i + (long)(j - 1) * (jpi + 1)]; };
    auto V = [&](const double* a, int i, int j) { return a[(i - 1) + (long)j * jpi]; };
    // ---- per-point classes --------------------------------------------------
    std::vector<uint16_t>& fl = c.fl;
    std::vector<double>& hflu = c.flu;
    std::vector<double>& hflv = c.flv;
    fl.assign((size_t)L.size(), 0);
    hflu.assign((size_t)L.size(), 0.0);
    hflv.assign((size_t)L.size(), 0.0);
    auto err_oob = [&](const char* what, int i, int j) {
        err = std::string("mask makes the serial model read outside its arrays (") + what +
              " at i=" + std::to_string(i) + ", j=" + std::to_string(j) + ")";
        return NEMO2D_ERR_MASK;
    };
    for (int j = 1; j <= jpj; ++j)
        for (int i = 0; i <= jpi; ++i) {  // U points
            unsigned f = 0;
            const int a = PT(i, j), b = PT(i + 1, j);
            unsigned cls = CLS_NONE;
            if (i >= 1 && i <= jpi - 1 && !(a + b <= 0) && !(a <= 0 || b <= 0)) {
                cls = CLS_COMP;
                if (PT(i, j - 1) <= 0 || PT(i + 1, j - 1) <= 0) f |= U_WALL_S;
                else if (j == 1) return err_oob("momentum_u south", i, j);
                if (PT(i, j + 1) <= 0 || PT(i + 1, j + 1) <= 0) f |= U_WALL_N;
                else if (j == jpj) return err_oob("momentum_u north", i, j);
            }
            if (a * b == 0) cls = CLS_SOLID;
            if (!(a + b <= -1) && (a < 0 || b < 0)) {
                cls = CLS_FLATHER;
                const bool east = a < 0;
                if (east) f |= U_FL_EAST;
                const int ii = east ? i + 1 : i - 1;
                if (ii < 0 || ii > jpi) return err_oob("bc_flather_u", i, j);
                hflu[L.at(i, j)] = sqrt(GRAV / U(hu, i, j));
            }
            f |= cls << U_CLS_SHIFT;
            unsigned nx = NX_SKIP;
            if (!(a + b <= 0)) {
                if (a * b > 0) nx = NX_AVG;
                else if (a <= 0) nx = NX_HI;
                else if (b <= 0) nx = NX_LO;
            }
            if ((nx == NX_AVG && (i == 0 || i == jpi)) || (nx == NX_HI && i == jpi) ||
                (nx == NX_LO && i == 0))
                return err_oob("next sshn_u", i, j);
            f |= nx << SU_SHIFT;
            fl[L.at(i, j)] |= (uint16_t)f;
        }
    for (int j = 0; j <= jpj; ++j)
        for (int i = 1; i <= jpi; ++i) {  // V points
            unsigned f = 0;
            const int a = PT(i, j), b = PT(i, j + 1);
            unsigned cls = CLS_NONE;
            // the first skip test looks east (original/nemolite2d.f90:523)
            if (j >= 1 && j <= jpj - 1 && !(a + PT(i + 1, j) <= 0) && !(a <= 0 || b <= 0)) {
                cls = CLS_COMP;
                if (PT(i - 1, j) <= 0 || PT(i - 1, j + 1) <= 0) f |= V_WALL_W;
                else if (i == 1) return err_oob("momentum_v west", i, j);
                if (PT(i + 1, j) <= 0 || PT(i + 1, j + 1) <= 0) f |= V_WALL_E;
                else if (i == jpi) return err_oob("momentum_v east", i, j);
            }
            if (a * b == 0) cls = CLS_SOLID;
            if (!(a + b <= -1) && (a < 0 || b < 0)) {
                cls = CLS_FLATHER;
                const bool north = a < 0;
                if (north) f |= V_FL_NORTH;
                const int jj = north ? j + 1 : j - 1;
                if (jj < 0 || jj > jpj) return err_oob("bc_flather_v", i, j);
                hflv[L.at(i, j)] = sqrt(GRAV / V(hv, i, j));
            }
            f |= cls << V_CLS_SHIFT;
            unsigned nx = NX_SKIP;
            if (!(a + b <= 0)) {
                if (a * b > 0) nx = NX_AVG;
                else if (a <= 0) nx = NX_HI;
                else if (b <= 0) nx = NX_LO;
            }
            if ((nx == NX_AVG && (j == 0 || j == jpj)) || (nx == NX_HI && j == jpj) ||
                (nx == NX_LO && j == 0))
                return err_oob("next sshn_v", i, j);
            f |= nx << SV_SHIFT;
            fl[L.at(i, j)] |= (uint16_t)f;
        }
    for (int j = 1; j <= jpj; ++j)
        for (int i = 1; i <= jpi; ++i) {  // T points: bc_ssh
            if (PT(i, j) <= 0) continue;
            if (PT(i, j - 1) < 0 || PT(i, j + 1) < 0 || PT(i + 1, j) < 0 || PT(i - 1, j) < 0)
                fl[L.at(i, j)] |= T_CLAMP;
        }

    // Flather lines, in loop order; the fused path needs every Flather
    // point's inner neighbour to be a non-Flather point.
    std::vector<int>& uptr = c.uptr;
    std::vector<int>& vptr = c.vptr;
    std::vector<long>& upts = c.upts;
    std::vector<long>& vpts = c.vpts;
    uptr.assign(jpj + 1, 0);
    vptr.assign(jpi + 1, 0);
    upts.clear();
    vpts.clear();
    bool& chained = c.chained;
    chained = false;
    for (int j = 1; j <= jpj; ++j) {
        uptr[j - 1] = (int)upts.size();
        for (int i = 0; i <= jpi; ++i) {
            const long o = L.at(i, j);
            if (ucls(fl[o]) == CLS_FLATHER) {
                upts.push_back(o);
                const long oi = (fl[o] & U_FL_EAST) ? o + 1 : o - 1;
                if (ucls(fl[oi]) == CLS_FLATHER) chained = true;
            }
        }
    }
    uptr[jpj] = (int)upts.size();
    // V lines are columns (i outer, j inner in the serial loop); the class
    // words are scanned row by row (a column walk strides a whole row per
    // point), counting per column, then filled in at each column's offset
    std::vector<int> cnt(jpi + 1, 0);
    for (int j = 0; j <= jpj; ++j)
        for (int i = 1; i <= jpi; ++i)
            if (vcls(fl[L.at(i, j)]) == CLS_FLATHER) ++cnt[i];
    for (int i = 1; i <= jpi; ++i) vptr[i] = vptr[i - 1] + cnt[i];
    vpts.resize(vptr[jpi]);
    std::vector<int> nxt(vptr.begin(), vptr.end() - 1);  // next slot of column i at nxt[i - 1]
    for (int j = 0; j <= jpj; ++j)
        for (int i = 1; i <= jpi; ++i) {
            const long o = L.at(i, j);
            if (vcls(fl[o]) == CLS_FLATHER) {
                vpts[nxt[i - 1]++] = o;
                const long oi = (fl[o] & V_FL_NORTH) ? o + L.pitch : o - L.pitch;
                if (vcls(fl[oi]) == CLS_FLATHER) chained = true;
            }
        }
    return NEMO2D_OK;
}

}  // namespace nemo2d

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* nemo2d_last_error(void) { return g_err.c_str(); }

const char* nemo2d_version(void) { return "nemo2d 0.1 gfx950 fp64"; }

int nemo2d_create(nemo2d_model** out, int jpi, int jpj, double dx, double dy, double dep_const,
                  int nit000, int nitend, int irecord, double rdt, double cbfr, double visc) {
    if (!out) return fail(NEMO2D_ERR_ARG, "nemo2d_create: null handle pointer");
    *out = nullptr;
    if (jpi < 1 || jpj < 1 || jpi > (1 << 20) || jpj > (1 << 20))
        return fail(NEMO2D_ERR_ARG, "nemo2d_create: grid size must be >= 1 (jpi=" +
                                        std::to_string(jpi) + ", jpj=" + std::to_string(jpj) + ")");
    if (!(rdt > 0.0) || !isfinite(cbfr) || !isfinite(visc))
        return fail(NEMO2D_ERR_ARG, "nemo2d_create: rdt must be > 0 and cbfr, visc finite");
    nemo2d_model* m = new nemo2d_model();
    m->prm.rdt = rdt;
    m->prm.cbfr = cbfr;
    m->prm.visc = visc;
    m->prm.fric = 1.0 + cbfr * rdt;
    m->nit000 = nit000;
    m->nitend = nitend;
    m->irecord = irecord;
    m->dx = dx;
    m->dy = dy;
    m->dep_const = dep_const;
    m->L = Layout::of(jpi, jpj);
    const Layout& L = m->L;
    int rc = NEMO2D_OK;
    auto bail = [&](int code) {
        nemo2d_destroy(m);
        return code;
    };
    if (hipGetDevice(&m->dev) != hipSuccess) return bail(fail(NEMO2D_ERR_HIP, "no HIP device"));
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(NEMO2D_ERR_HIP, "hipStreamCreate failed"));
    const size_t n = (size_t)L.size();
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < NF; ++f)
            if ((rc = dalloc(&m->S[k][f], n, m->stream))) return bail(rc);
    if ((rc = dalloc(&m->ssha, n, m->stream)) || (rc = dalloc(&m->ua, n, m->stream)) ||
        (rc = dalloc(&m->va, n, m->stream)) || (rc = dalloc(&m->flu, n, m->stream)) ||
        (rc = dalloc(&m->flv, n, m->stream)))
        return bail(rc);
    if (hipMalloc((void**)&m->flags, n * sizeof(uint16_t)) != hipSuccess ||
        hipMemsetAsync(m->flags, 0, n * sizeof(uint16_t), m->stream) != hipSuccess)
        return bail(fail(NEMO2D_ERR_HIP, "flags allocation failed"));
    dim3 g = grid2d(jpi + 1, jpj + 1);
    // partial slots: the most blocks any step kernel launches (the
    // thread-per-cell grid, or a marching grid of segments of at least 8 rows
    // padded to 8 XCDs)
    {
        // marching grids: segments of at least 6 rows
        const long mb = (long)((jpi + 1 + MW * ROUT - 1) / (MW * ROUT)) * ((jpj + 1 + 7) / 8);
        static_assert(ROUT <= DOUT && DOUT <= MOUT && ROUT <= GOUT, "the narrowest strips");
        const long mbp = (mb + 7) / 8 * 8;
        m->npart_max = (int)((long)(g.x * g.y) > mbp ? (long)(g.x * g.y) : mbp);
    }
    if ((rc = dalloc(&m->partials, 6 * ((size_t)m->npart_max + 1), m->stream)) ||
        (rc = dalloc(&m->cs_now, 3, m->stream)) ||
        (rc = dalloc(&m->chk, 1, m->stream)))
        return bail(rc);
    if (hipStreamSynchronize(m->stream) != hipSuccess)
        return bail(fail(NEMO2D_ERR_HIP, "nemo2d_create: zeroing the device arrays failed"));
    if (hipHostMalloc((void**)&m->h_pinned, 3 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(fail(NEMO2D_ERR_HIP, "pinned allocation failed"));
    if (hipMalloc(&m->stage, (size_t)m->L.size() * sizeof(double)) != hipSuccess)
        return bail(fail(NEMO2D_ERR_HIP, "staging allocation failed"));
    if (hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_free, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_tgt, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_feat, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_incr, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_read, hipEventDisableTiming) != hipSuccess)
        return bail(fail(NEMO2D_ERR_HIP, "event creation failed"));
    *out = m;
    return NEMO2D_OK;
}

// Runs the stage calls recorded for the fused pass (defined below).
static int lz_flush(nemo2d_model* m);

int nemo2d_set_grid(nemo2d_model* m, const int* pt, const double* e1t, const double* e2t,
                    const double* e1u, const double* e2u, const double* e1v, const double* e2v,
                    const double* e12t, const double* e12u, const double* e12v,
                    const double* gphiu, const double* gphiv, const double* ht, const double* hu,
                    const double* hv) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_set_grid: null model");
    const double* arrs[] = {e1t, e2t, e1u, e2u, e1v, e2v, e12t, e12u, e12v, gphiu, gphiv, ht, hu, hv};
    for (const double* a : arrs)
        if (!a) return fail(NEMO2D_ERR_ARG, "nemo2d_set_grid: null array");
    if (!pt) return fail(NEMO2D_ERR_ARG, "nemo2d_set_grid: null mask");
    if (m->lz_n > 0) {  // recorded stages act on the grid this replaces
        const int frc = lz_flush(m);
        if (frc) return frc;
    }
    const Layout& L = m->L;
    const int jpi = L.jpi, jpj = L.jpj;
    const long nu = (long)(jpi + 1) * jpj, nv = (long)jpi * (jpj + 1), nt = (long)jpi * jpj;
    Classes cls;
    {
        std::string err;
        const int rc = classify(L, pt, hu, hv, cls, err);
        if (rc) return fail(rc, "nemo2d_set_grid: " + err);
    }
    std::vector<uint16_t>& fl = cls.fl;
    std::vector<double>& hflu = cls.flu;
    std::vector<double>& hflv = cls.flv;
    std::vector<int>& uptr = cls.uptr;
    std::vector<int>& vptr = cls.vptr;
    std::vector<long>& upts = cls.upts;
    std::vector<long>& vpts = cls.vpts;
    m->fused_ok = !cls.chained;

    // ---- statics ------------------------------------------------------------
    std::vector<double> fcu((size_t)nu), fcv((size_t)nv);
    const double d2r = PI / 180.0;
    // 2 Omega sin(latitude) per point; a run of bitwise equal latitudes
    // (every point of a constant-latitude grid) reuses the value: sin is
    // the costliest part of set-up on large grids
    const auto coriolis = [d2r](const double* g, long n, double* f) {
        for (long k = 0; k < n; ++k)
            f[k] = (k > 0 && memcmp(&g[k], &g[k - 1], sizeof(double)) == 0) ? f[k - 1]
                                                                            : 2.0 * OMEGA * sin(g[k] * d2r);
    };
    coriolis(gphiu, nu, fcu.data());
    coriolis(gphiv, nv, fcv.data());
    struct {
        const double* a;
        int lb1, ub1, lb2, ub2;
        long n;
    } sd[NS] = {
        {e1t, 1, jpi, 1, jpj, nt},   {e2t, 1, jpi, 1, jpj, nt},   {e1u, 0, jpi, 1, jpj, nu},
        {e2u, 0, jpi, 1, jpj, nu},   {e1v, 1, jpi, 0, jpj, nv},   {e2v, 1, jpi, 0, jpj, nv},
        {e12t, 1, jpi, 1, jpj, nt},  {e12u, 0, jpi, 1, jpj, nu},  {e12v, 1, jpi, 0, jpj, nv},
        {ht, 1, jpi, 1, jpj, nt},    {hu, 0, jpi, 1, jpj, nu},    {hv, 1, jpi, 0, jpj, nv},
        {fcu.data(), 0, jpi, 1, jpj, nu}, {fcv.data(), 1, jpi, 0, jpj, nv}};
    bool uni = true, unim = true;
    for (int s = 0; s < NS; ++s) {
        const bool e = all_equal(sd[s].a, (size_t)sd[s].n);
        uni = uni && e;
        if (s != S_HT && s != S_HU && s != S_HV) unim = unim && e;
    }
    m->uniform = uni;
    m->uni_metrics = unim;
    {
        // hu(i,j) = 0.5*(ht(i,j)+ht(i+1,j)) for i = 1..jpi-1 and hv(i,j) =
        // 0.5*(ht(i,j)+ht(i,j+1)) for j = 1..jpj-1, bit for bit (arrays as
        // the caller passes them: ht(1:jpi,1:jpj), hu(0:jpi,1:jpj),
        // hv(1:jpi,0:jpj), column-major)
        bool d = true;
        for (int j = 1; j <= jpj && d; ++j)
            for (int i = 1; i <= jpi - 1 && d; ++i) {
                const double w = 0.5 * (ht[(i - 1) + (long)(j - 1) * jpi] + ht[i + (long)(j - 1) * jpi]);
                d = memcmp(&w, &hu[i + (long)(j - 1) * (jpi + 1)], sizeof(double)) == 0;
            }
        for (int j = 1; j <= jpj - 1 && d; ++j)
            for (int i = 1; i <= jpi && d; ++i) {
                const double w = 0.5 * (ht[(i - 1) + (long)(j - 1) * jpi] + ht[(i - 1) + (long)j * jpi]);
                d = memcmp(&w, &hv[(i - 1) + (long)j * jpi], sizeof(double)) == 0;
            }
        m->dep_derived = d;
    }
    m->scal = Scal{e1t[0], e2t[0], e1u[0], e2u[0], e1v[0], e2v[0], e12t[0], e12u[0], e12v[0],
                   ht[0], hu[0], hv[0], fcu[0], fcv[0], {}, 0};
    set_reciprocals(m->scal, m->prm.fric);
    const size_t n = (size_t)L.size();
    for (int s = 0; s < NS; ++s) {
        if (!m->st[s]) {
            int rc = dalloc(&m->st[s], n, m->stream);  // zeroed before the upload below
            if (rc) return rc;
        }
        int rc = upload(m, m->st[s], sd[s].a, sd[s].lb1, sd[s].ub1, sd[s].lb2, sd[s].ub2);
        if (rc) return rc;
    }
    HIPCHK(hipMemcpyAsync(m->flags, fl.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice, m->stream));
    // plain tiles of k_step_march_rc and k_step_march_gr (strips of ROUT
    // columns with lanes at c0-1 .. c0+62, segments of nj rows with rows
    // jb-1 .. je+1+RC_LA loaded): every class word there equal to RC_PLAIN
    // (the host tables live until the stream is synchronised below)
    std::vector<uint8_t> tabs[2];
    {
        const int nbx = (jpi + 1 + MW * ROUT - 1) / (MW * ROUT);
        const int nstrip = nbx * MW;
        for (const int nj : {RC_NJ, RC_NJ_SHORT}) {
            const int nseg = (jpj + 1 + nj - 1) / nj;
            std::vector<uint8_t>& tab = tabs[nj == RC_NJ ? 0 : 1];
            tab.assign((size_t)nstrip * nseg, 0);
            for (int w = 0; w < nstrip; ++w) {
                const int i0 = w * ROUT - 1, i1 = w * ROUT + 62;
                if (i0 < 0 || i1 > jpi + 1) continue;
                for (int q = 0; q < nseg; ++q) {
                    const int jb = q * nj, je = jb + nj - 1 < jpj ? jb + nj - 1 : jpj;
                    const int j0 = jb - 1, j1 = je + 1 + RC_LA;
                    if (j0 < 0 || j1 > jpj + 1) continue;
                    bool ok = true;
                    for (int j = j0; j <= j1 && ok; ++j)
                        for (int i = i0; i <= i1 && ok; ++i) ok = fl[L.at(i, j)] == RC_PLAIN;
                    tab[(size_t)w * nseg + q] = ok ? 1 : 0;
                }
            }
            uint8_t*& dst = nj == RC_NJ ? m->rc_plain : m->rc_plain_short;
            dfree(dst);
            dst = nullptr;
            HIPCHK(hipMalloc((void**)&dst, tab.size()));
            HIPCHK(hipMemcpyAsync(dst, tab.data(), tab.size(), hipMemcpyHostToDevice, m->stream));
        }
    }
    HIPCHK(hipMemcpyAsync(m->flu, hflu.data(), n * sizeof(double), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->flv, hflv.data(), n * sizeof(double), hipMemcpyHostToDevice, m->stream));
    dfree(m->flu_ptr); dfree(m->flu_pts); dfree(m->flv_ptr); dfree(m->flv_pts);
    m->flu_ptr = nullptr; m->flu_pts = nullptr; m->flv_ptr = nullptr; m->flv_pts = nullptr;
    m->nflu = (int)upts.size();
    m->nflv = (int)vpts.size();
    m->nflv0 = 0;
    for (long o : vpts)
        if ((o - L.base) / L.pitch == 0 && (fl[o] & V_FL_NORTH)) ++m->nflv0;
    HIPCHK(hipMalloc((void**)&m->flu_ptr, uptr.size() * sizeof(int)));
    HIPCHK(hipMalloc((void**)&m->flv_ptr, vptr.size() * sizeof(int)));
    HIPCHK(hipMalloc((void**)&m->flu_pts, (upts.size() + 1) * sizeof(long)));
    HIPCHK(hipMalloc((void**)&m->flv_pts, (vpts.size() + 1) * sizeof(long)));
    HIPCHK(hipMemcpyAsync(m->flu_ptr, uptr.data(), uptr.size() * sizeof(int), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->flv_ptr, vptr.data(), vptr.size() * sizeof(int), hipMemcpyHostToDevice, m->stream));
    if (!upts.empty())
        HIPCHK(hipMemcpyAsync(m->flu_pts, upts.data(), upts.size() * sizeof(long), hipMemcpyHostToDevice, m->stream));
    if (!vpts.empty())
        HIPCHK(hipMemcpyAsync(m->flv_pts, vpts.data(), vpts.size() * sizeof(long), hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (m->mixed) {
        // the single-precision form takes per-point depths and uniform
        // metrics and Coriolis parameter
        for (int s2 = 0; s2 < NS; ++s2)
            if (s2 != S_HT && s2 != S_HU && s2 != S_HV && !all_equal(sd[s2].a, (size_t)sd[s2].n))
                return fail(NEMO2D_ERR_ARG, "nemo2d_set_grid: the single-precision form needs "
                                            "uniform metrics and latitude (depth may vary)");
        float** hf[3] = {&m->hf_t, &m->hf_u, &m->hf_v};
        const int sidx[3] = {S_HT, S_HU, S_HV};
        // when hu, hv are the means of adjacent T depths, the FP32 ones are
        // the FP32 means of the FP32 T depths (so the marching kernel can
        // form them from ht on plain tiles); edge values as given
        std::vector<double> dhu, dhv;
        if (m->dep_derived) {
            dhu.assign(hu, hu + nu);
            dhv.assign(hv, hv + nv);
            for (int j = 1; j <= jpj; ++j)
                for (int i = 1; i <= jpi - 1; ++i)
                    dhu[i + (long)(j - 1) * (jpi + 1)] = (double)(0.5f * ((float)ht[(i - 1) + (long)(j - 1) * jpi] +
                                                                          (float)ht[i + (long)(j - 1) * jpi]));
            for (int j = 1; j <= jpj - 1; ++j)
                for (int i = 1; i <= jpi; ++i)
                    dhv[(i - 1) + (long)j * jpi] = (double)(0.5f * ((float)ht[(i - 1) + (long)(j - 1) * jpi] +
                                                                    (float)ht[(i - 1) + (long)j * jpi]));
        }
        for (int q = 0; q < 3; ++q) {
            if (!*hf[q]) {
                HIPCHK(hipMalloc((void**)hf[q], n * sizeof(float)));
                HIPCHK(hipMemsetAsync(*hf[q], 0, n * sizeof(float), m->stream));
            }
            const auto& d = sd[sidx[q]];
            const double* src = d.a;
            if (m->dep_derived && q == 1) src = dhu.data();
            if (m->dep_derived && q == 2) src = dhv.data();
            int rc = upload_as<float>(m, *hf[q], src, d.lb1, d.ub1, d.lb2, d.ub2);
            if (rc) return rc;
        }
    }
    m->grid_set = true;
    m->uv_derived = false;  // the state's sshn_u/sshn_v were not formed on this grid
    return NEMO2D_OK;
}

int nemo2d_set_state(nemo2d_model* m, const double* sshn, const double* sshn_u, const double* sshn_v,
                     const double* un, const double* vn) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_set_state: null model");
    if (!sshn || !sshn_u || !sshn_v || !un || !vn)
        return fail(NEMO2D_ERR_ARG, "nemo2d_set_state: null array");
    const int jpi = m->L.jpi, jpj = m->L.jpj;
    const size_t bytes = (size_t)m->L.size() * sizeof(double);
    int grc = snap_guard(m, -1);
    if (grc) return grc;
    m->cur = 0;
    if (m->mixed) {
        const size_t nb[NF] = {m->compensated ? bytes : bytes / 2, bytes / 2, bytes / 2, bytes / 2,
                               bytes / 2};
        for (int k = 0; k < 2; ++k)
            for (int f = 0; f < NF; ++f) HIPCHK(hipMemsetAsync(m->X[k][f], 0, nb[f], m->stream));
        int rc;
        if (m->compensated)
            rc = upload_as<double>(m, (double*)m->X[0][F_SSHN], sshn, 1, jpi, 1, jpj);
        else
            rc = upload_as<float>(m, (float*)m->X[0][F_SSHN], sshn, 1, jpi, 1, jpj);
        if (rc || (rc = upload_as<float>(m, (float*)m->X[0][F_SSHN_U], sshn_u, 0, jpi, 1, jpj)) ||
            (rc = upload_as<float>(m, (float*)m->X[0][F_SSHN_V], sshn_v, 1, jpi, 0, jpj)) ||
            (rc = upload_as<float>(m, (float*)m->X[0][F_UN], un, 0, jpi, 1, jpj)) ||
            (rc = upload_as<float>(m, (float*)m->X[0][F_VN], vn, 1, jpi, 0, jpj)))
            return rc;
        for (int f = 0; f < NF; ++f)
            HIPCHK(hipMemcpyAsync(m->X[1][f], m->X[0][f], nb[f], hipMemcpyDeviceToDevice, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        m->uv_stale[0] = m->uv_stale[1] = false;
        m->uv_derived = false;
        m->lz_n = 0;
        m->a_current = false;
        m->state_set = true;
        m->cs_next = 0;
        m->cs_n = 0;
        drop_pending(m);
        return NEMO2D_OK;
    }
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < NF; ++f) HIPCHK(hipMemsetAsync(m->S[k][f], 0, bytes, m->stream));
    HIPCHK(hipMemsetAsync(m->ssha, 0, bytes, m->stream));
    HIPCHK(hipMemsetAsync(m->ua, 0, bytes, m->stream));
    HIPCHK(hipMemsetAsync(m->va, 0, bytes, m->stream));
    int rc;
    if ((rc = upload(m, m->S[0][F_SSHN], sshn, 1, jpi, 1, jpj)) ||
        (rc = upload(m, m->S[0][F_SSHN_U], sshn_u, 0, jpi, 1, jpj)) ||
        (rc = upload(m, m->S[0][F_SSHN_V], sshn_v, 1, jpi, 0, jpj)) ||
        (rc = upload(m, m->S[0][F_UN], un, 0, jpi, 1, jpj)) ||
        (rc = upload(m, m->S[0][F_VN], vn, 1, jpi, 0, jpj)))
        return rc;
    // the skip-class points of sshn_u/sshn_v are carried, not recomputed:
    // give the second buffer the same values
    for (int f = 0; f < NF; ++f)
        HIPCHK(hipMemcpyAsync(m->S[1][f], m->S[0][f], bytes, hipMemcpyDeviceToDevice, m->stream));
    // sshn_u/sshn_v as given: when they are sshn's interpolations bit for bit
    // (a state from rest, or one this library wrote), the first step may
    // already take a kernel that forms them from sshn
    HIPCHK(hipMemsetAsync(m->chk, 0, sizeof(double), m->stream));
    {
        const Layout& L = m->L;
        double** s0 = m->S[0];
        WITH_STATICS(m, st, (k_interp_check<<<grid2d(L.jpi + 1, L.jpj + 1), dim3(TBX, TBY), 0, m->stream>>>(
                                st, m->flags, s0[F_SSHN], L, s0[F_SSHN_U], s0[F_SSHN_V], (unsigned*)m->chk)));
        LAUNCHCHK("interp_check");
    }
    HIPCHK(hipMemcpyAsync(m->h_pinned, m->chk, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->uv_stale[0] = m->uv_stale[1] = false;
    m->uv_derived = m->grid_set && *(const unsigned*)m->h_pinned == 0u;
    m->lz_n = 0;  // recorded stage calls act on the state this replaces
    m->a_current = true;
    m->state_set = true;
    m->cs_next = 0;
    m->cs_n = 0;
    drop_pending(m);
    return NEMO2D_OK;
}


#define READY_NF(m, who)                                                                   \
    do {                                                                                   \
        if (!(m)) return fail(NEMO2D_ERR_ARG, std::string(who) + ": null model");          \
        if (!(m)->grid_set || !(m)->state_set)                                             \
            return fail(NEMO2D_ERR_STATE, std::string(who) + ": grid and state must be set first"); \
    } while (0)

// every entry point but the stage calls first runs the stage calls that
// were recorded for a fused pass
#define READY(m, who)                                                                      \
    do {                                                                                   \
        READY_NF(m, who);                                                                  \
        const int frc_ = lz_flush(m);                                                      \
        if (frc_) return frc_;                                                             \
    } while (0)

#define READY_STAGE(m, who)                                                                \
    do {                                                                                   \
        READY_NF(m, who);                                                                  \
        if ((m)->mixed)                                                                    \
            return fail(NEMO2D_ERR_STATE, std::string(who) +                               \
                                              ": the single-precision form has the fused step only"); \
    } while (0)

// ---- the nine stage calls -------------------------------------------------------
// Each stage's kernel (the eager path).  original/nemolite2d.f90:375-386
// calls them in this order every step; made in that order on a grid the
// fused step handles, they are recorded instead (stage_call, stage_defers)
// and nemo2d_next runs the fused step, which gives the nine stages' and
// next's results bit for bit.
// A call out of order, or any other entry point, first runs the recorded
// stages one by one (lz_flush).
enum { LZ_CONT = 0, LZ_MOMU, LZ_MOMV, LZ_SSH, LZ_SOLU, LZ_SOLV, LZ_FLU, LZ_FLV, LZ_N };

static int stage_run(nemo2d_model* m, int stage) {
    int rc = uv_sync(m);  // the stages read sshn_u/sshn_v
    if (rc) return rc;
    const Layout& L = m->L;
    NState n = nstate(m, m->cur);
    switch (stage) {
        case LZ_CONT:
            WITH_STATICS(m, st, (k_continuity<<<grid2d(L.jpi + 1, L.jpj + 1), dim3(TBX, TBY), 0, m->stream>>>(
                                    st, n, m->ssha, L, m->prm.rdt)));
            LAUNCHCHK("continuity");
            break;
        case LZ_MOMU:
            WITH_STATICS(m, st, (k_momentum_u<<<grid2d(L.jpi + 1, L.jpj + 1), dim3(TBX, TBY), 0, m->stream>>>(
                                    st, n, m->flags, m->ua, L, m->prm)));
            LAUNCHCHK("momentum_u");
            m->a_current = true;
            break;
        case LZ_MOMV:
            WITH_STATICS(m, st, (k_momentum_v<<<grid2d(L.jpi + 1, L.jpj + 1), dim3(TBX, TBY), 0, m->stream>>>(
                                    st, n, m->flags, m->va, L, m->prm)));
            LAUNCHCHK("momentum_v");
            m->a_current = true;
            break;
        case LZ_SSH:
            k_bc_ssh<<<grid2d(L.jpi + 1, L.jpj + 1), dim3(TBX, TBY), 0, m->stream>>>(
                m->flags, m->ssha, L, tide_eval(m->tide_amp, m->tide_omega, m->lz_rtime));
            LAUNCHCHK("bc_ssh");
            break;
        case LZ_SOLU:
            k_bc_solid_u<<<grid2d(L.jpi + 1, L.jpj + 1), dim3(TBX, TBY), 0, m->stream>>>(m->flags, m->ua, L);
            LAUNCHCHK("bc_solid_u");
            break;
        case LZ_SOLV:
            k_bc_solid_v<<<grid2d(L.jpi + 1, L.jpj + 1), dim3(TBX, TBY), 0, m->stream>>>(m->flags, m->va, L);
            LAUNCHCHK("bc_solid_v");
            break;
        case LZ_FLU:
            if (m->nflu > 0) {
                const int lines = L.jpj;
                k_bc_flather_u<<<(lines + 255) / 256, 256, 0, m->stream>>>(
                    m->flu_ptr, m->flu_pts, lines, m->flags, m->flu, m->S[m->cur][F_SSHN_U], m->ua);
                LAUNCHCHK("bc_flather_u");
            }
            break;
        case LZ_FLV:
            if (m->nflv > 0) {
                const int lines = L.jpi;
                k_bc_flather_v<<<(lines + 255) / 256, 256, 0, m->stream>>>(
                    m->flv_ptr, m->flv_pts, lines, m->flags, m->flv, m->S[m->cur][F_SSHN_V], m->va, L.pitch);
                LAUNCHCHK("bc_flather_v");
            }
            break;
        default: break;
    }
    return NEMO2D_OK;
}

static int lz_flush(nemo2d_model* m) {
    const int n = m->lz_n;
    m->lz_n = 0;
    for (int st = 0; st < n; ++st) {
        const int rc = stage_run(m, st);
        if (rc) return rc;
    }
    return NEMO2D_OK;
}

// True when a call of this stage would be recorded for the fused pass.
static bool stage_defers(const nemo2d_model* m, int stage) {
    return m->lz_on && m->fused_ok && !m->mixed && m->lz_n == stage;
}

// Records the call for the fused pass when it comes in order; else runs the
// recorded stages and then this one.
static int stage_call(nemo2d_model* m, int stage) {
    if (stage_defers(m, stage)) {
        m->lz_n = stage + 1;
        return NEMO2D_OK;
    }
    int rc = lz_flush(m);
    if (rc) return rc;
    return stage_run(m, stage);
}

int nemo2d_continuity(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_continuity");
    RegionScope rs_(m, TR_CONT, true, true, stage_defers(m, LZ_CONT));
    return stage_call(m, LZ_CONT);
}

int nemo2d_momentum_u(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_momentum_u");
    RegionScope rs_(m, TR_MOM, true, false, stage_defers(m, LZ_MOMU));
    return stage_call(m, LZ_MOMU);
}

int nemo2d_momentum_v(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_momentum_v");
    RegionScope rs_(m, TR_MOM, false, true, stage_defers(m, LZ_MOMV));
    return stage_call(m, LZ_MOMV);
}

int nemo2d_bc_ssh(nemo2d_model* m, double rtime) {
    READY_STAGE(m, "nemo2d_bc_ssh");
    RegionScope rs_(m, TR_BCS, true, false, stage_defers(m, LZ_SSH));
    // a recorded bc_ssh keeps its time; an earlier recorded one has been run
    // (in order) before this call replaces it
    if (m->lz_n != LZ_SSH) {
        const int rc = lz_flush(m);
        if (rc) return rc;
    }
    m->lz_rtime = rtime;
    return stage_call(m, LZ_SSH);
}

int nemo2d_bc_solid_u(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_bc_solid_u");
    return stage_call(m, LZ_SOLU);
}

int nemo2d_bc_solid_v(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_bc_solid_v");
    return stage_call(m, LZ_SOLV);
}

int nemo2d_bc_flather_u(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_bc_flather_u");
    return stage_call(m, LZ_FLU);
}

int nemo2d_bc_flather_v(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_bc_flather_v");
    RegionScope rs_(m, TR_BCS, false, true, stage_defers(m, LZ_FLV));
    return stage_call(m, LZ_FLV);
}

static int fused_step(nemo2d_model* m, double tide, int slot);

int nemo2d_next(nemo2d_model* m) {
    READY_STAGE(m, "nemo2d_next");
    RegionScope rs_(m, TR_NEXT, true, true);
    if (m->lz_n == LZ_N) {
        // the step's nine stage calls were recorded in order: one fused pass,
        // logging into the stage path's next slot
        m->lz_n = 0;
        const int slot = (m->cs_cap > 0 && m->cs_next < m->cs_cap) ? m->cs_next : -1;
        return fused_step(m, tide_eval(m->tide_amp, m->tide_omega, m->lz_rtime), slot);
    }
    {
        const int rc = lz_flush(m);
        if (rc) return rc;
    }
    {
        const int rc = uv_sync(m);
        if (rc) return rc;
    }
    int grc = snap_guard(m, m->cur);
    if (grc) return grc;
    const Layout& L = m->L;
    double** s = m->S[m->cur];
    const dim3 g = grid2d(L.jpi + 1, L.jpj + 1);
    // with recording on, the checksums come from the same pass: block
    // partials in the step's slot (reduced when the log is read) or in the
    // scratch partials (reduced at once)
    const bool rec = m->cs_cap > 0 && m->cs_next < m->cs_cap;
    const int slot = m->cs_next;
    double* part = rec && m->cs_part ? m->cs_part + slot * m->cs_slot : m->partials;
    if (rec) {
        WITH_STATICS(m, st, (k_next<decltype(st), true><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                                st, m->flags, m->ssha, m->ua, m->va, s[F_SSHN], s[F_SSHN_U],
                                s[F_SSHN_V], s[F_UN], s[F_VN], L, part)));
    } else {
        WITH_STATICS(m, st, (k_next<decltype(st), false><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                                st, m->flags, m->ssha, m->ua, m->va, s[F_SSHN], s[F_SSHN_U],
                                s[F_SSHN_V], s[F_UN], s[F_VN], L, nullptr)));
    }
    LAUNCHCHK("next");
    m->a_current = true;
    m->uv_derived = true;
    if (rec) {
        const int nparts = (int)(g.x * g.y);
        if (m->cs_part) {
            m->cs_pend[slot] = nparts;
            m->pend_any = true;
        } else {
            k_checksum_finish<<<1, RED_THREADS, 0, m->stream>>>(part, nparts, m->cs_log + 3 * (long)slot);
            LAUNCHCHK("checksum_finish");
        }
        ++m->cs_next;
        if (m->cs_next > m->cs_n) m->cs_n = m->cs_next;
    }
    return NEMO2D_OK;
}


static int run_stages(nemo2d_model* m, int istp) {
    int rc;
    if ((rc = nemo2d_continuity(m)) || (rc = nemo2d_momentum_u(m)) || (rc = nemo2d_momentum_v(m)) ||
        (rc = nemo2d_bc_ssh(m, (double)istp * m->prm.rdt)) || (rc = nemo2d_bc_solid_u(m)) ||
        (rc = nemo2d_bc_solid_v(m)) || (rc = nemo2d_bc_flather_u(m)) || (rc = nemo2d_bc_flather_v(m)))
        return rc;
    // next logs into the next free slot; keep fused and stage slots aligned
    const int slot = istp - m->nit000;
    if (m->cs_cap > 0 && slot >= 0 && slot < m->cs_cap) m->cs_next = slot;
    return nemo2d_next(m);
}

}  // extern "C"

// Fused-step kernel variants (nemo2d_set_variant); 0 picks the default.
enum {
    VAR_AUTO = 0,
    VAR_CELL = 1,           // one thread per cell, Flather in the epilogue
    VAR_CELL_FLIN = 2,      // one thread per cell, Flather inline
    VAR_MARCH16 = 3,        // five-array column-marching, 16-row segments (the fallback
                            // where the shared-term kernels' 32-bit offsets do not fit)
    VAR_MARCH12_SHN = 4,    // shared face terms, 12-row segments
    VAR_MARCH12_SHD = 5,    // shared face terms, 12 rows, per-point depths
    VAR_MARCH12_DP = 6,     // shared face terms, own-column rows, DPP neighbours
    VAR_MARCH12_DPF = 7,    //   the same dividing by precomputed reciprocals
    VAR_MARCH12_DPB = 8,    // 6 and 7 with the block's waves in step per row
    VAR_MARCH12_DPFB = 9,
    VAR_MARCH12_RC = 10,    // reads sshn, un, vn only (sshn_u/v recomputed), loads two rows
                            // ahead, reciprocal divisions under one guard per row, class-free
                            // code on plain interior tiles
    VAR_MARCH12_GRD = 11,   // 5 reading sshn, un, vn (and depths) only, plain tiles
    VAR_MARCH8_GRD = 12,    // 11 with 8-row segments (more waves on mid-size grids)
    VAR_MARCH8_RC = 13,     // 10 with 8-row segments
    VAR_LAST = 13
};

static int launch_march16(nemo2d_model* m, const NState& n, const AOut& a, double tide, int& nparts) {
    const Layout& L = m->L;
    const int nbx = (L.jpi + 1 + MW * MOUT - 1) / (MW * MOUT);
    const int nby = (L.jpj + 1 + M16 - 1) / M16;
    const int ntiles = nbx * nby;
    const int nblocks = (ntiles + 7) / 8 * 8;
    if (nblocks > m->npart_max) return fail(NEMO2D_ERR_ARG, "partials buffer too small");
    k_step_march<<<nblocks, 64 * MW, 0, m->stream>>>(m->scal, n, a, m->flags, L, m->prm, tide,
                                                      m->part_dst, nbx, ntiles);
    LAUNCHCHK("step_march");
    nparts = nblocks;
    return NEMO2D_OK;
}

// the shared-term kernels address the arrays by 32-bit byte offsets
static bool sh_fits(const Layout& L) { return (unsigned long)L.size() * sizeof(double) < (1ul << 32); }

template <int NJ>
static int launch_march_sh(nemo2d_model* m, const NState& n, const AOut& a, double tide, int& nparts) {
    const Layout& L = m->L;
    if (!sh_fits(L)) return launch_march16(m, n, a, tide, nparts);
    const int nbx = (L.jpi + 1 + MW * MOUT - 1) / (MW * MOUT);
    const int nby = (L.jpj + 1 + NJ - 1) / NJ;
    const int ntiles = nbx * nby;
    const int nblocks = (ntiles + 7) / 8 * 8;
    if (nblocks > m->npart_max) return fail(NEMO2D_ERR_ARG, "partials buffer too small");
    k_step_march_sh<NJ><<<nblocks, 64 * MW, 0, m->stream>>>(
        m->scal, n, a, m->flags, m->flv, L, m->prm, tide, m->part_dst, nbx, ntiles);
    LAUNCHCHK("step_march_sh");
    nparts = nblocks;
    return NEMO2D_OK;
}

template <int NJ, bool FD, bool BAR>
static int launch_march_dp(nemo2d_model* m, const NState& n, const AOut& a, double tide, int& nparts) {
    const Layout& L = m->L;
    if (!sh_fits(L)) return launch_march16(m, n, a, tide, nparts);
    const int nbx = (L.jpi + 1 + MW * DOUT - 1) / (MW * DOUT);
    const int nby = (L.jpj + 1 + NJ - 1) / NJ;
    const int ntiles = nbx * nby;
    const int nblocks = (ntiles + 7) / 8 * 8;
    if (nblocks > m->npart_max) return fail(NEMO2D_ERR_ARG, "partials buffer too small");
    if (FD && !m->scal.fast_div)
        return launch_march_dp<NJ, false, BAR>(m, n, a, tide, nparts);
    k_step_march_dp<NJ, FD, BAR><<<nblocks, 64 * MW, 0, m->stream>>>(
        m->scal, n, a, m->flags, m->flv, L, m->prm, tide, m->part_dst, nbx, ntiles);
    LAUNCHCHK("step_march_dp");
    nparts = nblocks;
    return NEMO2D_OK;
}

template <int DM, int NJ = RC_NJ>
static int launch_march_rc(nemo2d_model* m, const NState& n, const AOut& a, double tide, int& nparts) {
    const Layout& L = m->L;
    if (DM != 0 && !m->scal.fast_div) return launch_march_rc<0, NJ>(m, n, a, tide, nparts);
    if (DM == 2 && m->scal.one_corr) return launch_march_rc<3, NJ>(m, n, a, tide, nparts);
    static_assert(RC_LA >= 2, "the plain-tile table covers the rows loaded ahead");
    static_assert(NJ == RC_NJ || NJ == RC_NJ_SHORT, "the plain-tile tables");
    const int nbx = (L.jpi + 1 + MW * ROUT - 1) / (MW * ROUT);
    const int nby = (L.jpj + 1 + NJ - 1) / NJ;
    const int ntiles = nbx * nby;
    const int nblocks = (ntiles + 7) / 8 * 8;
    if (nblocks > m->npart_max) return fail(NEMO2D_ERR_ARG, "partials buffer too small");
    k_step_march_rc<DM><<<nblocks, 64 * MW, 0, m->stream>>>(
        m->scal, n, a, m->flags, m->flv, L, m->prm, tide, m->part_dst, nbx, ntiles,
        NJ == RC_NJ ? m->rc_plain : m->rc_plain_short, NJ);
    LAUNCHCHK("step_march_rc");
    nparts = nblocks;
    return NEMO2D_OK;
}

// Flather points only on row 0 with the inner point north (the reference
// basin's open boundary): the marching kernels set those themselves.
static bool flather_row0_only(const nemo2d_model* m) { return m->nflu == 0 && m->nflv == m->nflv0; }

template <class R, class SS, bool DARR, int NJ, int WPE = 0>
static int launch_march_g(nemo2d_model* m, const MetG<R>& mt, const ParamsT<R>& p,
                          const StateG<R, SS, DARR>& n, const OutG<R, SS>& a, double tide,
                          int& nparts) {
    const Layout& L = m->L;
    const int nbx = (L.jpi + 1 + GW * GOUT - 1) / (GW * GOUT);
    const int nby = (L.jpj + 1 + NJ - 1) / NJ;
    const int ntiles = nbx * nby;
    const int nblocks = (ntiles + 7) / 8 * 8;
    if (nblocks > m->npart_max) return fail(NEMO2D_ERR_ARG, "partials buffer too small");
    k_step_march_g<R, SS, DARR, NJ, WPE><<<nblocks, 64 * GW, 0, m->stream>>>(
        mt, p, n, a, m->flags, m->flv, L, tide, m->part_dst, nbx, ntiles);
    LAUNCHCHK("step_march_g");
    nparts = nblocks;
    return NEMO2D_OK;
}

static MetG<double> metg(const Scal& d) {
    return MetG<double>{d.e1t, d.e2t, d.e1u, d.e2u, d.e1v, d.e2v, d.e12t, d.e12u, d.e12v,
                        d.fcu, d.fcv, d.ht, d.hu, d.hv};
}

// FP64 fused step by the generic marching kernel with per-point depths
// (within 168 VGPRs: 3 waves per SIMD; measured 3% faster than 2 waves at
// 170 VGPRs despite a 12-byte spill)
// The generic step that reads sshn, un, vn (and depths) only; the
// plain-tile table of k_step_march_rc serves it (same strips, segments and
// rows ahead).
template <class R, class SS, bool DARR, int NJ, int LA, bool DV = false>
static int launch_march_gr(nemo2d_model* m, const MetG<R>& mt, const ParamsT<R>& p,
                           const StateG<R, SS, DARR>& n, const OutG<R, SS>& a, double tide,
                           int& nparts) {
    static_assert(GROUT == ROUT && GW == MW && (NJ == RC_NJ || NJ == RC_NJ_SHORT) && LA <= RC_LA,
                  "the plain-tile tables");
    const uint8_t* plain = NJ == RC_NJ ? m->rc_plain : m->rc_plain_short;
    const Layout& L = m->L;
    const int nbx = (L.jpi + 1 + GW * GROUT - 1) / (GW * GROUT);
    const int nby = (L.jpj + 1 + NJ - 1) / NJ;
    const int ntiles = nbx * nby;
    const int nblocks = (ntiles + 7) / 8 * 8;
    if (nblocks > m->npart_max) return fail(NEMO2D_ERR_ARG, "partials buffer too small");
    RcpG rc;
    for (int q = 0; q < QD_N; ++q) rc.r[q] = m->scal.rcp[q];
    // FP64 with every metric divisor passing the one-correction test: plain
    // tiles divide by the metrics' reciprocals (the per-point depths by IEEE)
    if (sizeof(R) == sizeof(double) && m->scal.one_corr_met)
        k_step_march_gr<R, SS, DARR, NJ, LA, DV, sizeof(R) == sizeof(double) ? 3 : 0>
            <<<nblocks, 64 * GW, 0, m->stream>>>(mt, p, n, a, m->flags, m->flv, L, tide, m->part_dst, nbx,
                                                 ntiles, plain, rc);
    else
        k_step_march_gr<R, SS, DARR, NJ, LA, DV><<<nblocks, 64 * GW, 0, m->stream>>>(
            mt, p, n, a, m->flags, m->flv, L, tide, m->part_dst, nbx, ntiles, plain, rc);
    LAUNCHCHK("step_march_gr");
    nparts = nblocks;
    return NEMO2D_OK;
}

template <int NJ>
static int launch_march_depths_rc(nemo2d_model* m, const NState& n, const AOut& a, double tide,
                                  int& nparts) {
    StateG<double, double, true> ng{n.sshn, n.sshn_u, n.sshn_v, n.un, n.vn, m->st[S_HT], m->st[S_HU],
                                    m->st[S_HV]};
    OutG<double, double> ag{a.sshn, a.sshn_u, a.sshn_v, a.un, a.vn};
    if (m->dep_derived)
        return launch_march_gr<double, double, true, NJ, 2, true>(m, metg(m->scal), m->prm, ng, ag, tide,
                                                                  nparts);
    return launch_march_gr<double, double, true, NJ, 2>(m, metg(m->scal), m->prm, ng, ag, tide, nparts);
}

static int launch_march_depths(nemo2d_model* m, const NState& n, const AOut& a, double tide, int& nparts) {
    StateG<double, double, true> ng{n.sshn, n.sshn_u, n.sshn_v, n.un, n.vn, m->st[S_HT], m->st[S_HU],
                                    m->st[S_HV]};
    OutG<double, double> ag{a.sshn, a.sshn_u, a.sshn_v, a.un, a.vn};
    return launch_march_g<double, double, true, 12, 3>(m, metg(m->scal), m->prm, ng, ag, tide, nparts);
}

template <bool FLIN>
static int launch_cell(nemo2d_model* m, const NState& n, const AOut& a, double tide, int& nparts) {
    const Layout& L = m->L;
    dim3 g = grid2d(L.jpi + 1, L.jpj + 1);
    WITH_STATICS(m, st, (k_step_fused<decltype(st), FLIN><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                            st, n, a, m->flags, m->flu, m->flv, L, m->prm, tide, m->part_dst)));
    LAUNCHCHK("step_fused");
    nparts = (int)(g.x * g.y);
    return NEMO2D_OK;
}

// Grid sizes (T cells) up to which the default kernel is one thread per cell
// (FP64 with uniform statics: Flather points inline where they are all on
// row 0, else by the epilogue / with per-point statics: Flather points by the
// epilogue / the single-precision form), and up to which
// a uniform FP64 grid takes the five-array marching kernel.  On smaller grids
// a marching step lasts as long as one wave's march of its segment, whatever
// the grid size.
// (device us per step, one thread per cell / 8-row marching: uniform 448^2
// 14.0 / 14.2, 512^2 16.2 / 14.5; seamount FP64 384^2 21.2 / 17.3; seamount
// FP32 448^2 11.9 / 13.4, 512^2 13.9 / 13.4)
constexpr long AUTO_CELL_MAX_UNIFORM = 230000;
constexpr long AUTO_CELL_MAX_DEPTHS = 100000;
constexpr long AUTO_CELL_MAX_MIXED = 230000;
// Uniform statics: the sshn-only kernel with 8-row segments up to this many
// cells, 12-row ones beyond (device us per step, reference basin, 12 / 8
// rows / the five-array shared-term kernel: 768^2 26.1 / 20.4 / 28.3,
// 1024^2 27.1 / 25.4 / 29.2, 1536^2 42.0 / 38.7 / 49.3, 2048^2 55.8 / 55.5,
// 3072^2 118.5 / 121.2, 4096^2 191.7 / 197.2).
constexpr long AUTO_SHORT_SEG_MAX_UNIFORM = 3000000;
// Per-point depths: 8-row segments (more waves in flight) up to this many
// cells, 12-row ones (less halo work) beyond.  Measured on MI355X (device us
// per step, seamount basins, 12 / 8 rows): 512^2 23.1 / 17.5, 1024^2 33.5 /
// 29.3, 2048^2 71.8 / 68.9, 3072^2 150.7 / 148.8, 4096^2 238.3 / 244.3.
constexpr long AUTO_SHORT_SEG_MAX_DEPTHS = 10000000;

extern "C" {

static int step_mixed(nemo2d_model* m, int istp) {
    if (!m->fused_ok)
        return fail(NEMO2D_ERR_MASK, "nemo2d_step: chained Flather points are not supported by "
                                     "the single-precision form");
    const Layout& L = m->L;
    const int k = m->cur;
    const Scal& d = m->scal;
    ScalF sf{(float)d.e1t, (float)d.e2t, (float)d.e1u, (float)d.e2u, (float)d.e1v, (float)d.e2v,
             (float)d.e12t, (float)d.e12u, (float)d.e12v, (float)d.fcu, (float)d.fcv, (float)d.ht, (float)d.hu, (float)d.hv};
    ParamsT<float> pf{(float)m->prm.rdt, (float)m->prm.cbfr, (float)m->prm.visc, (float)m->prm.fric};
    const double tide = tide_eval(m->tide_amp, m->tide_omega, (double)istp * m->prm.rdt);
    dim3 g = grid2d(L.jpi + 1, L.jpj + 1);
    void** x = m->X[k];
    void** y = m->X[1 - k];
    // one thread per cell when chosen, or by default on small grids (FP32
    // seamount basins, device us per step, marching / per cell: 512^2 20.3 /
    // 14.3, 1024^2 29.3 / 36.7)
    const bool small = m->variant == VAR_AUTO && (long)L.jpi * L.jpj <= AUTO_CELL_MAX_MIXED;
    if (m->variant != VAR_CELL && !small && flather_row0_only(m) && sh_fits(L)) {
        // the kernel that reads sshn, un, vn only once sshn_u/sshn_v are sshn's
        // interpolations (after any step)
        const bool gr = m->uv_derived && m->variant != VAR_MARCH12_SHD;
        if (!gr) {
            const int urc = uv_sync(m);
            if (urc) return urc;
        }
        const MetG<float> mt{sf.e1t, sf.e2t, sf.e1u, sf.e2u, sf.e1v, sf.e2v, sf.e12t, sf.e12u, sf.e12v,
                             sf.fcu, sf.fcv, sf.ht, sf.hu, sf.hv};
        // 8-row segments when chosen, or by default up to the FP64 form's size
        const bool short_seg = m->variant == VAR_MARCH8_GRD ||
                               (m->variant == VAR_AUTO && (long)L.jpi * L.jpj <= AUTO_SHORT_SEG_MAX_DEPTHS);
        int nparts = 0, rc;
        m->part_dst = m->partials;
        if (m->compensated) {
            StateG<float, double, true> n{(const double*)x[F_SSHN], (const float*)x[F_SSHN_U],
                                          (const float*)x[F_SSHN_V], (const float*)x[F_UN],
                                          (const float*)x[F_VN], m->hf_t, m->hf_u, m->hf_v};
            OutG<float, double> a{(double*)y[F_SSHN], (float*)y[F_SSHN_U], (float*)y[F_SSHN_V],
                                  (float*)y[F_UN], (float*)y[F_VN]};
            rc = !gr ? launch_march_g<float, double, true, 12>(m, mt, pf, n, a, tide, nparts)
                 : short_seg
                     ? (m->dep_derived
                            ? launch_march_gr<float, double, true, RC_NJ_SHORT, 2, true>(m, mt, pf, n, a, tide, nparts)
                            : launch_march_gr<float, double, true, RC_NJ_SHORT, 2>(m, mt, pf, n, a, tide, nparts))
                 : m->dep_derived ? launch_march_gr<float, double, true, 12, 2, true>(m, mt, pf, n, a, tide, nparts)
                                  : launch_march_gr<float, double, true, 12, 2>(m, mt, pf, n, a, tide, nparts);
        } else {
            StateG<float, float, true> n{(const float*)x[F_SSHN], (const float*)x[F_SSHN_U],
                                         (const float*)x[F_SSHN_V], (const float*)x[F_UN],
                                         (const float*)x[F_VN], m->hf_t, m->hf_u, m->hf_v};
            OutG<float, float> a{(float*)y[F_SSHN], (float*)y[F_SSHN_U], (float*)y[F_SSHN_V],
                                 (float*)y[F_UN], (float*)y[F_VN]};
            rc = !gr ? launch_march_g<float, float, true, 12>(m, mt, pf, n, a, tide, nparts)
                 : short_seg
                     ? (m->dep_derived
                            ? launch_march_gr<float, float, true, RC_NJ_SHORT, 2, true>(m, mt, pf, n, a, tide, nparts)
                            : launch_march_gr<float, float, true, RC_NJ_SHORT, 2>(m, mt, pf, n, a, tide, nparts))
                 : m->dep_derived ? launch_march_gr<float, float, true, 12, 2, true>(m, mt, pf, n, a, tide, nparts)
                                  : launch_march_gr<float, float, true, 12, 2>(m, mt, pf, n, a, tide, nparts);
        }
        if (rc) return rc;
        m->uv_stale[1 - k] = gr;
        m->uv_derived = true;
        m->cur = 1 - k;
        const int slot = istp - m->nit000;
        if (m->cs_cap > 0 && slot >= 0 && slot < m->cs_cap) {
            k_checksum_finish<<<1, RED_THREADS, 0, m->stream>>>(m->partials, nparts, m->cs_log + 3 * (long)slot);
            LAUNCHCHK("checksum_finish");
            if (slot + 1 > m->cs_n) m->cs_n = slot + 1;
            m->cs_next = slot + 1;
        }
        return NEMO2D_OK;
    }
    {
        const int urc = uv_sync(m);
        if (urc) return urc;
    }
    if (m->compensated) {
        MixedState<true> n{(const double*)x[F_SSHN], (const float*)x[F_SSHN_U], (const float*)x[F_SSHN_V],
                           (const float*)x[F_UN], (const float*)x[F_VN]};
        MixedOut<true> a{(double*)y[F_SSHN], (float*)y[F_SSHN_U], (float*)y[F_SSHN_V], (float*)y[F_UN],
                         (float*)y[F_VN]};
        k_step_mixed<true><<<g, dim3(TBX, TBY), 0, m->stream>>>(sf, n, a, m->hf_t, m->hf_u, m->hf_v,
                                                                m->flags, m->flu, m->flv, L, pf, tide,
                                                                m->partials);
    } else {
        MixedState<false> n{(const float*)x[F_SSHN], (const float*)x[F_SSHN_U], (const float*)x[F_SSHN_V],
                            (const float*)x[F_UN], (const float*)x[F_VN]};
        MixedOut<false> a{(float*)y[F_SSHN], (float*)y[F_SSHN_U], (float*)y[F_SSHN_V], (float*)y[F_UN],
                          (float*)y[F_VN]};
        k_step_mixed<false><<<g, dim3(TBX, TBY), 0, m->stream>>>(sf, n, a, m->hf_t, m->hf_u, m->hf_v,
                                                                 m->flags, m->flu, m->flv, L, pf, tide,
                                                                 m->partials);
    }
    LAUNCHCHK("step_mixed");
    m->uv_stale[1 - k] = false;
    m->uv_derived = true;
    m->cur = 1 - k;
    const int slot = istp - m->nit000;
    if (m->cs_cap > 0 && slot >= 0 && slot < m->cs_cap) {
        k_checksum_finish<<<1, RED_THREADS, 0, m->stream>>>(m->partials, (int)(g.x * g.y),
                                                            m->cs_log + 3 * (long)slot);
        LAUNCHCHK("checksum_finish");
        if (slot + 1 > m->cs_n) m->cs_n = slot + 1;
        m->cs_next = slot + 1;
    }
    return NEMO2D_OK;
}

int nemo2d_step(nemo2d_model* m, int istp) {
    READY(m, "nemo2d_step");
    RegionScope rs_(m, TR_FUSED, m->fused_ok || m->mixed, m->fused_ok || m->mixed);
    if (m->mixed) return step_mixed(m, istp);
    if (!m->fused_ok) return run_stages(m, istp);
    return fused_step(m, tide_eval(m->tide_amp, m->tide_omega, (double)istp * m->prm.rdt),
                      istp - m->nit000);
}

// One fused step of the FP64 model with the given tide; logs the checksums
// into slot `slot` when recording covers it.
static int fused_step(nemo2d_model* m, double tide, int slot) {
    const Layout& L = m->L;
    const int k = m->cur;
    NState n = nstate(m, k);
    AOut a{m->S[1 - k][F_SSHN], m->S[1 - k][F_SSHN_U], m->S[1 - k][F_SSHN_V], m->S[1 - k][F_UN],
           m->S[1 - k][F_VN]};
    int var = m->variant;
    const bool depth_march_ok = m->uni_metrics && flather_row0_only(m) && sh_fits(L);
    if (var == VAR_AUTO) {
        // A marching wave walks its segment's rows one after another, so on a
        // small grid (few waves) the step is the latency of one wave's march;
        // there one thread per cell is faster, and on mid-size grids shorter
        // segments (more waves).  Crossovers measured on MI355X (device us
        // per step, marching / one thread per cell): uniform 512^2 20.4 /
        // 16.0, 1024^2 25.4 / 42.5; per-point depths 256^2 22.3 / 12.5,
        // 512^2 17.5 / 26.2.
        const long cells = (long)L.jpi * L.jpj;
        if (cells <= (m->uniform ? AUTO_CELL_MAX_UNIFORM : AUTO_CELL_MAX_DEPTHS))
            var = (m->uniform && flather_row0_only(m)) ? VAR_CELL_FLIN : VAR_CELL;
        else if (m->uniform)
            var = cells <= AUTO_SHORT_SEG_MAX_UNIFORM ? VAR_MARCH8_RC : VAR_MARCH12_RC;
        else
            var = !depth_march_ok ? VAR_CELL
                                  : cells <= AUTO_SHORT_SEG_MAX_DEPTHS ? VAR_MARCH8_GRD : VAR_MARCH12_GRD;
    }
    const bool depth_var = var == VAR_MARCH12_SHD || var == VAR_MARCH12_GRD || var == VAR_MARCH8_GRD;
    if (depth_var && !depth_march_ok) var = VAR_CELL;
    // marching with scalar depths needs uniform statics
    if (!m->uniform && var != VAR_CELL_FLIN && !depth_var) var = VAR_CELL;
    // the kernels that recompute sshn_u/sshn_v need them to be sshn's
    // interpolations (true after any step) and no Flather points off row 0
    if ((var == VAR_MARCH12_RC || var == VAR_MARCH8_RC) && !(m->uv_derived && flather_row0_only(m) && sh_fits(L)))
        var = VAR_MARCH12_DPFB;
    if ((var == VAR_MARCH12_GRD || var == VAR_MARCH8_GRD) && !m->uv_derived) var = VAR_MARCH12_SHD;
    const bool rc_kernel = var == VAR_MARCH12_RC || var == VAR_MARCH8_RC || var == VAR_MARCH12_GRD ||
                           var == VAR_MARCH8_GRD;
    if (!rc_kernel && m->uv_stale[k]) {
        const int urc = uv_sync(m);
        if (urc) return urc;
    }
    int nparts = 0, rc = NEMO2D_OK;
    bool flin = false;
    double* out = (m->cs_cap > 0 && slot >= 0 && slot < m->cs_cap) ? m->cs_log + 3 * (long)slot : nullptr;
    const bool defer = out && m->cs_part;
    m->part_dst = defer ? m->cs_part + slot * m->cs_slot : m->partials;
    if ((rc = snap_guard(m, 1 - k))) return rc;
    switch (var) {
        case VAR_CELL: rc = launch_cell<false>(m, n, a, tide, nparts); break;
        case VAR_CELL_FLIN: rc = launch_cell<true>(m, n, a, tide, nparts); flin = true; break;
        case VAR_MARCH12_SHN: rc = launch_march_sh<12>(m, n, a, tide, nparts); break;
        case VAR_MARCH12_SHD: rc = launch_march_depths(m, n, a, tide, nparts); break;
        case VAR_MARCH12_GRD: rc = launch_march_depths_rc<RC_NJ>(m, n, a, tide, nparts); break;
        case VAR_MARCH8_GRD: rc = launch_march_depths_rc<RC_NJ_SHORT>(m, n, a, tide, nparts); break;
        case VAR_MARCH12_DP: rc = launch_march_dp<12, false, false>(m, n, a, tide, nparts); break;
        case VAR_MARCH12_DPF: rc = launch_march_dp<12, true, false>(m, n, a, tide, nparts); break;
        case VAR_MARCH12_DPB: rc = launch_march_dp<12, false, true>(m, n, a, tide, nparts); break;
        case VAR_MARCH12_DPFB: rc = launch_march_dp<12, true, true>(m, n, a, tide, nparts); break;
        case VAR_MARCH12_RC: rc = launch_march_rc<2>(m, n, a, tide, nparts); break;
        case VAR_MARCH8_RC: rc = launch_march_rc<2, RC_NJ_SHORT>(m, n, a, tide, nparts); break;
        default: rc = launch_march16(m, n, a, tide, nparts); break;
    }
    if (rc) return rc;
    m->last_var = var;
    m->uv_stale[1 - k] = rc_kernel;
    m->uv_derived = true;
    m->cur = 1 - k;
    m->a_current = false;
    // the shared-term kernels set the row-0 Flather v points themselves
    const bool sh = var >= VAR_MARCH12_SHN && var <= VAR_LAST && sh_fits(L) &&
                    ((var != VAR_MARCH12_SHD && var != VAR_MARCH12_GRD && var != VAR_MARCH8_GRD) ||
                     depth_march_ok);
    const bool has_fl = (m->nflu + m->nflv - (sh ? m->nflv0 : 0)) > 0;
    if (!flin && has_fl) {
        // Flather points of the new state, then (if recording) the checksums
        // (partial nparts holds the Flather points' sums)
        k_epilogue<<<1, RED_THREADS, 0, m->stream>>>(
            m->flu_ptr, m->flu_pts, m->nflu ? L.jpj : 0, m->flv_ptr, m->flv_pts, m->nflv ? L.jpi : 0,
            m->flags, m->flu, m->flv, n.sshn_u, n.sshn_v, a.un, a.vn, L, m->part_dst, nparts,
            defer ? nullptr : out, sh ? 1 : 0);
        LAUNCHCHK("epilogue");
        ++nparts;
    } else if (out && !defer) {
        k_checksum_finish<<<1, RED_THREADS, 0, m->stream>>>(m->part_dst, nparts, out);
        LAUNCHCHK("checksum_finish");
    }
    if (defer) {
        m->cs_pend[slot] = nparts;
        m->pend_any = true;
    }
    if (out) {
        if (slot + 1 > m->cs_n) m->cs_n = slot + 1;
        m->cs_next = slot + 1;
    }
    return NEMO2D_OK;
}

int nemo2d_stream_bandwidth(long n, int repeats, double* gbps) {
    if (n < 1 || repeats < 1 || !gbps) return fail(NEMO2D_ERR_ARG, "nemo2d_stream_bandwidth: bad argument");
    double* a[10] = {};
    int rc = NEMO2D_OK;
    for (int k = 0; k < 10 && rc == NEMO2D_OK; ++k) rc = dalloc(&a[k], (size_t)n, nullptr);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == NEMO2D_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(NEMO2D_ERR_HIP, "nemo2d_stream_bandwidth: event");
    if (rc == NEMO2D_OK) {
        Stream5 s;
        for (int k = 0; k < 5; ++k) {
            s.in[k] = a[k];
            s.out[k] = a[5 + k];
        }
        const unsigned grid = 65536;
        for (int it = 0; it < 3; ++it) k_stream5<<<grid, 256>>>(s, n);
        (void)hipEventRecord(e0, nullptr);
        for (int it = 0; it < repeats; ++it) k_stream5<<<grid, 256>>>(s, n);
        (void)hipEventRecord(e1, nullptr);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
            hipGetLastError() != hipSuccess)
            rc = fail(NEMO2D_ERR_HIP, "nemo2d_stream_bandwidth: launch");
        else
            *gbps = 80.0 * (double)n * repeats / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (int k = 0; k < 10; ++k) dfree(a[k]);
    return rc;
}

int nemo2d_set_tide(nemo2d_model* m, double amp, double period_hours) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_set_tide: null model");
    if (!isfinite(amp) || !(period_hours > 0.0))
        return fail(NEMO2D_ERR_ARG, "nemo2d_set_tide: amplitude must be finite, period > 0");
    if (m->lz_n > 0) {  // a recorded bc_ssh keeps the tide it was called with
        const int rc = lz_flush(m);
        if (rc) return rc;
    }
    m->tide_amp = amp;
    m->tide_omega = tide_omega_of(period_hours);
    return NEMO2D_OK;
}

int nemo2d_set_stage_fusion(nemo2d_model* m, int on) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_set_stage_fusion: null model");
    if (m->grid_set && m->state_set) {
        const int rc = lz_flush(m);
        if (rc) return rc;
    }
    m->lz_n = 0;
    m->lz_on = on != 0;
    return NEMO2D_OK;
}

int nemo2d_last_variant(nemo2d_model* m, int* variant) {
    if (!m || !variant) return fail(NEMO2D_ERR_ARG, "nemo2d_last_variant: bad argument");
    *variant = m->last_var;
    return NEMO2D_OK;
}

int nemo2d_set_variant(nemo2d_model* m, int variant) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_set_variant: null model");
    if (variant < 0 || variant > VAR_LAST)
        return fail(NEMO2D_ERR_ARG, "nemo2d_set_variant: unknown variant " + std::to_string(variant));
    m->variant = variant;
    return NEMO2D_OK;
}

int nemo2d_run(nemo2d_model* m, int nit000, int nitend) {
    READY(m, "nemo2d_run");
    for (int istp = nit000; istp <= nitend; ++istp) {
        int rc = nemo2d_step(m, istp);
        if (rc) return rc;
    }
    return NEMO2D_OK;
}

// Reduce the partials of every pending logged step into the log.
static int flush_checksums(nemo2d_model* m) {
    if (!m->pend_any) return NEMO2D_OK;
    std::vector<int> pairs;
    for (int k = 0; k < m->cs_cap; ++k)
        if (m->cs_pend[k] > 0) {
            pairs.push_back(k);
            pairs.push_back(m->cs_pend[k]);
            m->cs_pend[k] = 0;
        }
    m->pend_any = false;
    if (pairs.empty()) return NEMO2D_OK;
    HIPCHK(hipMemcpyAsync(m->d_pend, pairs.data(), pairs.size() * sizeof(int), hipMemcpyHostToDevice,
                          m->stream));
    k_checksum_finish_batch<<<(unsigned)(pairs.size() / 2), RED_THREADS, 0, m->stream>>>(
        m->cs_part, m->cs_slot, m->d_pend, m->cs_log);
    LAUNCHCHK("checksum_finish_batch");
    HIPCHK(hipStreamSynchronize(m->stream));
    return NEMO2D_OK;
}

// Slot buffers for the deferred finish up to this many bytes; beyond it
// each logged step finishes its checksums at once.
constexpr size_t CS_DEFER_MAX_BYTES = size_t(8) << 30;

int nemo2d_record_checksums(nemo2d_model* m, int capacity) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_record_checksums: null model");
    if (capacity < 0) return fail(NEMO2D_ERR_ARG, "nemo2d_record_checksums: negative capacity");
    HIPCHK(hipStreamSynchronize(m->stream));
    dfree(m->cs_log);
    dfree(m->cs_part);
    dfree(m->d_pend);
    m->cs_log = nullptr;
    m->cs_part = nullptr;
    m->d_pend = nullptr;
    m->cs_pend.clear();
    m->pend_any = false;
    m->cs_cap = 0;
    m->cs_n = 0;
    m->cs_next = 0;
    if (capacity > 0) {
        int rc = dalloc(&m->cs_log, 3 * (size_t)capacity, m->stream);
        if (rc) return rc;
        m->cs_cap = capacity;
        m->cs_slot = 6 * ((long)m->npart_max + 1);
        const size_t bytes = (size_t)capacity * (size_t)m->cs_slot * sizeof(double);
        if (bytes <= CS_DEFER_MAX_BYTES) {
            if (hipMalloc((void**)&m->cs_part, bytes) != hipSuccess ||
                hipMalloc((void**)&m->d_pend, 2 * (size_t)capacity * sizeof(int)) != hipSuccess) {
                (void)hipGetLastError();  // no room: finish at once instead
                dfree(m->cs_part);
                dfree(m->d_pend);
                m->cs_part = nullptr;
                m->d_pend = nullptr;
            } else {
                m->cs_pend.assign((size_t)capacity, 0);
            }
        }
    }
    return NEMO2D_OK;
}

int nemo2d_checksum_log(nemo2d_model* m, double* out, int max_steps, int* nsteps) {
    if (!m || (!out && max_steps > 0)) return fail(NEMO2D_ERR_ARG, "nemo2d_checksum_log: bad argument");
    HIPCHK(hipStreamSynchronize(m->stream));
    int rc = flush_checksums(m);
    if (rc) return rc;
    const int n = m->cs_n < max_steps ? m->cs_n : max_steps;
    if (n > 0) HIPCHK(hipMemcpy(out, m->cs_log, 3 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    if (nsteps) *nsteps = m->cs_n;
    return NEMO2D_OK;
}

int nemo2d_checksums(nemo2d_model* m, double out[3]) {
    READY(m, "nemo2d_checksums");
    if (!out) return fail(NEMO2D_ERR_ARG, "nemo2d_checksums: null output");
    int rc = checksum_into(m, m->cur, m->cs_now);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(m->h_pinned, m->cs_now, 3 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    for (int k = 0; k < 3; ++k) out[k] = m->h_pinned[k];
    for (int k = 0; k < 3; ++k)
        if (!isfinite(out[k])) return fail(NEMO2D_ERR_NONFINITE, "nemo2d_checksums: non-finite checksum");
    return NEMO2D_OK;
}

// The network runs on the default stream.  A buffer it reads (features,
// targets) is rewritten on a model's stream only after the default stream's
// work queued so far (which may still read it): the model's stream waits.
static int after_default_stream(nemo2d_model* m) {
    HIPCHK(hipEventRecord(m->ev_free, nullptr));
    HIPCHK(hipStreamWaitEvent(m->stream, m->ev_free, 0));
    return NEMO2D_OK;
}
// ... and the default stream waits for what m's stream wrote.
static int before_default_stream(nemo2d_model* m) {
    HIPCHK(hipEventRecord(m->ev_feat, m->stream));
    HIPCHK(hipStreamWaitEvent(nullptr, m->ev_feat, 0));
    return NEMO2D_OK;
}

int nemo2d_closure_features(nemo2d_model* m, const double* mean, const double* scale, void* x) {
    READY_STAGE(m, "nemo2d_closure_features");
    if (!mean || !scale || !x) return fail(NEMO2D_ERR_ARG, "nemo2d_closure_features: null argument");
    Norm4 nm;
    for (int c = 0; c < 4; ++c) {
        nm.m[c] = mean[c];
        nm.s[c] = scale[c];
        if (!(scale[c] != 0.0)) return fail(NEMO2D_ERR_ARG, "nemo2d_closure_features: zero scale");
    }
    const Layout& L = m->L;
    double** s = m->S[m->cur];
    int rc = after_default_stream(m);
    if (rc) return rc;
    WITH_STATICS(m, st, (k_closure_features<<<grid2d(L.jpi, L.jpj), dim3(TBX, TBY), 0, m->stream>>>(
                            st, s[F_SSHN], s[F_UN], s[F_VN], L, nm, (uint16_t*)x)));
    LAUNCHCHK("closure_features");
    return before_default_stream(m);
}

int nemo2d_closure_apply(nemo2d_model* m, const float* z, int zstride, const double* scale) {
    READY_STAGE(m, "nemo2d_closure_apply");
    if (!z || !scale || zstride < 3) return fail(NEMO2D_ERR_ARG, "nemo2d_closure_apply: bad argument");
    {
        const int frc = lz_flush(m);  // recorded stages act on the state before the increments
        if (frc) return frc;
    }
    // the increments come from work on the default stream: this stream waits for it
    HIPCHK(hipEventRecord(m->ev_incr, nullptr));
    HIPCHK(hipStreamWaitEvent(m->stream, m->ev_incr, 0));
    int grc = snap_guard(m, m->cur);
    if (grc) return grc;
    const Layout& L = m->L;
    double** s = m->S[m->cur];
    const Scale3 sc{{scale[0], scale[1], scale[2]}};
    k_closure_add<<<grid2d(L.jpi, L.jpj), dim3(TBX, TBY), 0, m->stream>>>(z, zstride, sc, L, s[F_SSHN],
                                                                          s[F_UN], s[F_VN]);
    LAUNCHCHK("closure_add");
    // sshn_u/sshn_v are now next's interpolations of the new sshn (where
    // next sets them) and not yet written: the sshn-only step kernels form
    // them from sshn, and uv_sync fills them in before anything reads them
    m->uv_stale[m->cur] = true;
    // later default-stream work (the next network pass rewriting z) waits for the read of z
    HIPCHK(hipEventRecord(m->ev_read, m->stream));
    HIPCHK(hipStreamWaitEvent(nullptr, m->ev_read, 0));
    m->a_current = false;
    m->uv_derived = true;
    return NEMO2D_OK;
}

int nemo2d_closure_restrict(nemo2d_model* f, nemo2d_model* c) {
    READY_STAGE(f, "nemo2d_closure_restrict");
    if (!c || c == f) return fail(NEMO2D_ERR_ARG, "nemo2d_closure_restrict: bad coarse model");
    if (!c->grid_set || c->mixed)
        return fail(NEMO2D_ERR_STATE, "nemo2d_closure_restrict: the coarse model needs a grid (FP64)");
    if (f->L.jpi != 2 * c->L.jpi || f->L.jpj != 2 * c->L.jpj)
        return fail(NEMO2D_ERR_ARG, "nemo2d_closure_restrict: the fine grid must be 2x the coarse grid");
    if (c->lz_n > 0) c->lz_n = 0;  // recorded stage calls act on the state this replaces
    int rc = lz_flush(f);  // recorded stage calls of the fine model run first
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(f->stream));  // the fine state is final
    if ((rc = snap_guard(c, -1))) return rc;
    c->cur = 0;
    const size_t bytes = (size_t)c->L.size() * sizeof(double);
    double** fs = f->S[f->cur];
    double** cs = c->S[0];
    for (int k = 0; k < NF; ++k) HIPCHK(hipMemsetAsync(cs[k], 0, bytes, c->stream));
    HIPCHK(hipMemsetAsync(c->ssha, 0, bytes, c->stream));
    HIPCHK(hipMemsetAsync(c->ua, 0, bytes, c->stream));
    HIPCHK(hipMemsetAsync(c->va, 0, bytes, c->stream));
    const dim3 g = grid2d(c->L.jpi + 1, c->L.jpj + 1);
    if (f->uniform)
        k_restrict<<<g, dim3(TBX, TBY), 0, c->stream>>>(fs[F_SSHN], fs[F_UN], fs[F_VN],
                                                         DepthConst{f->scal.hu}, DepthConst{f->scal.hv},
                                                         f->L, c->L, cs[F_SSHN], cs[F_UN], cs[F_VN]);
    else
        k_restrict<<<g, dim3(TBX, TBY), 0, c->stream>>>(fs[F_SSHN], fs[F_UN], fs[F_VN],
                                                         DepthArr{f->st[S_HU]}, DepthArr{f->st[S_HV]},
                                                         f->L, c->L, cs[F_SSHN], cs[F_UN], cs[F_VN]);
    LAUNCHCHK("restrict");
    // sshn_u/sshn_v by the next-stage rules; the points next leaves alone stay 0
    WITH_STATICS(c, st, (k_interp<<<g, dim3(TBX, TBY), 0, c->stream>>>(st, c->flags, cs[F_SSHN], c->L,
                                                                       cs[F_SSHN_U], cs[F_SSHN_V])));
    LAUNCHCHK("restrict_interp");
    for (int k = 0; k < NF; ++k)
        HIPCHK(hipMemcpyAsync(c->S[1][k], cs[k], bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->uv_stale[0] = c->uv_stale[1] = false;
    c->uv_derived = false;
    c->a_current = true;
    c->state_set = true;
    c->cs_next = 0;
    c->cs_n = 0;
    drop_pending(c);
    return NEMO2D_OK;
}

int nemo2d_closure_target(nemo2d_model* truth, nemo2d_model* pred, const double* scale, float* y) {
    READY_STAGE(truth, "nemo2d_closure_target");
    READY_STAGE(pred, "nemo2d_closure_target");
    if (!scale || !y) return fail(NEMO2D_ERR_ARG, "nemo2d_closure_target: null argument");
    if (truth->L.jpi != pred->L.jpi || truth->L.jpj != pred->L.jpj)
        return fail(NEMO2D_ERR_ARG, "nemo2d_closure_target: models of different sizes");
    for (int k = 0; k < 3; ++k)
        if (!(scale[k] != 0.0)) return fail(NEMO2D_ERR_ARG, "nemo2d_closure_target: zero scale");
    int rc = lz_flush(truth);
    if (rc || (rc = lz_flush(pred))) return rc;
    // pred's stream writes y after truth's state is final and after the
    // default stream's work queued so far (which may still read y)
    HIPCHK(hipEventRecord(truth->ev_feat, truth->stream));
    HIPCHK(hipStreamWaitEvent(pred->stream, truth->ev_feat, 0));
    if ((rc = after_default_stream(pred))) return rc;
    const Layout& L = pred->L;
    double** t = truth->S[truth->cur];
    double** q = pred->S[pred->cur];
    k_closure_target<<<grid2d(L.jpi, L.jpj), dim3(TBX, TBY), 0, pred->stream>>>(
        t[F_SSHN], t[F_UN], t[F_VN], q[F_SSHN], q[F_UN], q[F_VN], L, Scale3{{scale[0], scale[1], scale[2]}},
        y);
    LAUNCHCHK("closure_target");
    // later work on truth's stream (a new state) waits for that read
    HIPCHK(hipEventRecord(pred->ev_tgt, pred->stream));
    HIPCHK(hipStreamWaitEvent(truth->stream, pred->ev_tgt, 0));
    return before_default_stream(pred);
}

int nemo2d_retrieve_results(nemo2d_model* m, double* sshn, double* un, double* vn, double* ua,
                            double* va) {
    READY(m, "nemo2d_retrieve_results");
    const int jpi = m->L.jpi, jpj = m->L.jpj;
    if (m->mixed) {
        void** x = m->X[m->cur];
        int rc = NEMO2D_OK;
        if (sshn) {
            rc = m->compensated ? download_as<double>(m, sshn, (const double*)x[F_SSHN], 1, jpi, 1, jpj)
                                : download_as<float>(m, sshn, (const float*)x[F_SSHN], 1, jpi, 1, jpj);
            if (rc) return rc;
        }
        if (un && (rc = download_as<float>(m, un, (const float*)x[F_UN], 0, jpi, 1, jpj))) return rc;
        if (vn && (rc = download_as<float>(m, vn, (const float*)x[F_VN], 1, jpi, 0, jpj))) return rc;
        if (ua && (rc = download_as<float>(m, ua, (const float*)x[F_UN], 0, jpi, 1, jpj))) return rc;
        if (va && (rc = download_as<float>(m, va, (const float*)x[F_VN], 1, jpi, 0, jpj))) return rc;
        return NEMO2D_OK;
    }
    double** s = m->S[m->cur];
    // after a fused step, ua/va equal un/vn (next copied them)
    const double* dua = m->a_current ? m->ua : s[F_UN];
    const double* dva = m->a_current ? m->va : s[F_VN];
    int rc = NEMO2D_OK;
    if (sshn && (rc = download(m, sshn, s[F_SSHN], 1, jpi, 1, jpj))) return rc;
    if (un && (rc = download(m, un, s[F_UN], 0, jpi, 1, jpj))) return rc;
    if (vn && (rc = download(m, vn, s[F_VN], 1, jpi, 0, jpj))) return rc;
    if (ua && (rc = download(m, ua, dua, 0, jpi, 1, jpj))) return rc;
    if (va && (rc = download(m, va, dva, 1, jpi, 0, jpj))) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return NEMO2D_OK;
}

int nemo2d_retrieve_aux(nemo2d_model* m, double* sshn_u, double* sshn_v, double* ssha) {
    READY(m, "nemo2d_retrieve_aux");
    const int jpi = m->L.jpi, jpj = m->L.jpj;
    if (m->mixed) {
        const int urc = uv_sync(m);
        if (urc) return urc;
        void** x = m->X[m->cur];
        int rc = NEMO2D_OK;
        if (sshn_u && (rc = download_as<float>(m, sshn_u, (const float*)x[F_SSHN_U], 0, jpi, 1, jpj))) return rc;
        if (sshn_v && (rc = download_as<float>(m, sshn_v, (const float*)x[F_SSHN_V], 1, jpi, 0, jpj))) return rc;
        if (ssha) {
            rc = m->compensated ? download_as<double>(m, ssha, (const double*)x[F_SSHN], 1, jpi, 1, jpj)
                                : download_as<float>(m, ssha, (const float*)x[F_SSHN], 1, jpi, 1, jpj);
            if (rc) return rc;
        }
        return NEMO2D_OK;
    }
    int rc = uv_sync(m);
    if (rc) return rc;
    double** s = m->S[m->cur];
    if (sshn_u && (rc = download(m, sshn_u, s[F_SSHN_U], 0, jpi, 1, jpj))) return rc;
    if (sshn_v && (rc = download(m, sshn_v, s[F_SSHN_V], 1, jpi, 0, jpj))) return rc;
    // after a fused step the new ssh is the now-state's sshn
    if (ssha && (rc = download(m, ssha, m->a_current ? m->ssha : s[F_SSHN], 1, jpi, 1, jpj))) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return NEMO2D_OK;
}

int nemo2d_retrieve_grid(nemo2d_model* m, int which, double* out) {
    if (!m || !out) return fail(NEMO2D_ERR_ARG, "nemo2d_retrieve_grid: bad argument");
    if (!m->grid_set) return fail(NEMO2D_ERR_STATE, "nemo2d_retrieve_grid: no grid set");
    if (which < 0 || which >= NS + 2)
        return fail(NEMO2D_ERR_ARG, "nemo2d_retrieve_grid: which must be 0.." + std::to_string(NS + 1));
    const int jpi = m->L.jpi, jpj = m->L.jpj;
    // host layouts of the statics in S_* order, then the Flather coefficients
    static const char kind[NS + 2] = {'T', 'T', 'U', 'U', 'V', 'V', 'T', 'U', 'V',
                                      'T', 'U', 'V', 'U', 'V', 'U', 'V'};
    const double* src = which < NS ? m->st[which] : which == NS ? m->flu : m->flv;
    const char k = kind[which];
    int rc = k == 'T' ? download(m, out, src, 1, jpi, 1, jpj)
             : k == 'U' ? download(m, out, src, 0, jpi, 1, jpj)
                        : download(m, out, src, 1, jpi, 0, jpj);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
    return NEMO2D_OK;
}

int nemo2d_retrieve_classes(nemo2d_model* m, unsigned short* out) {
    if (!m || !out) return fail(NEMO2D_ERR_ARG, "nemo2d_retrieve_classes: bad argument");
    if (!m->grid_set) return fail(NEMO2D_ERR_STATE, "nemo2d_retrieve_classes: no grid set");
    const int jpi = m->L.jpi, jpj = m->L.jpj;
    const size_t w = (size_t)(jpi + 2) * sizeof(uint16_t);
    HIPCHK(hipMemcpy2DAsync(out, w, m->flags + m->L.at(0, 0), m->L.pitch * sizeof(uint16_t), w, jpj + 2,
                            hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return NEMO2D_OK;
}

int nemo2d_classify(int jpi, int jpj, const int* pt, const double* hu, const double* hv,
                    unsigned short* out) {
    if (jpi < 1 || jpj < 1 || !pt || !hu || !hv || !out)
        return fail(NEMO2D_ERR_ARG, "nemo2d_classify: bad argument");
    const Layout L = Layout::of(jpi, jpj);
    Classes cls;
    std::string err;
    const int rc = classify(L, pt, hu, hv, cls, err);
    if (rc) return fail(rc, "nemo2d_classify: " + err);
    for (int j = 0; j <= jpj + 1; ++j)
        for (int i = 0; i <= jpi + 1; ++i) out[i + (long)j * (jpi + 2)] = cls.fl[L.at(i, j)];
    return NEMO2D_OK;
}

int nemo2d_timer_start(nemo2d_model* m) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_timer_start: null model");
    HIPCHK(hipEventRecord(m->ev0, m->stream));
    return NEMO2D_OK;
}

int nemo2d_timer_stop(nemo2d_model* m, double* ms) {
    if (!m || !ms) return fail(NEMO2D_ERR_ARG, "nemo2d_timer_stop: bad argument");
    HIPCHK(hipEventRecord(m->ev1, m->stream));
    HIPCHK(hipEventSynchronize(m->ev1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, m->ev0, m->ev1));
    *ms = (double)t;
    return NEMO2D_OK;
}

int nemo2d_create_mixed(nemo2d_model** out, int jpi, int jpj, double dx, double dy,
                        double dep_const, int nit000, int nitend, int irecord, double rdt,
                        double cbfr, double visc, int compensated) {
    int rc = nemo2d_create(out, jpi, jpj, dx, dy, dep_const, nit000, nitend, irecord, rdt, cbfr, visc);
    if (rc) return rc;
    nemo2d_model* m = *out;
    m->mixed = true;
    m->compensated = compensated != 0;
    const size_t n = (size_t)m->L.size();
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < NF; ++f) {
            const size_t es = (f == F_SSHN && m->compensated) ? sizeof(double) : sizeof(float);
            if (hipMalloc(&m->X[k][f], n * es) != hipSuccess ||
                hipMemsetAsync(m->X[k][f], 0, n * es, m->stream) != hipSuccess ||
                hipStreamSynchronize(m->stream) != hipSuccess) {
                nemo2d_destroy(m);
                *out = nullptr;
                return fail(NEMO2D_ERR_HIP, "nemo2d_create_mixed: allocation failed");
            }
        }
    return NEMO2D_OK;
}

int nemo2d_diagnostics(nemo2d_model* m, double out[3]) {
    READY(m, "nemo2d_diagnostics");
    if (!out) return fail(NEMO2D_ERR_ARG, "nemo2d_diagnostics: null output");
    const Layout& L = m->L;
    if (!m->diag) {
        int rc = dalloc(&m->diag, 3, m->stream);
        if (rc) return rc;
    }
    dim3 g = grid2d(L.jpi + 1, L.jpj + 1);
    const double* e12t = m->uniform ? nullptr : m->st[S_E12T];
    if (m->mixed) {
        void** x = m->X[m->cur];
        if (m->compensated)
            k_diag_partials<double, float, float><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                (const double*)x[F_SSHN], (const float*)x[F_UN], (const float*)x[F_VN], m->hf_t, 0.0,
                nullptr, m->scal.e12t, L, m->partials);
        else
            k_diag_partials<float, float, float><<<g, dim3(TBX, TBY), 0, m->stream>>>(
                (const float*)x[F_SSHN], (const float*)x[F_UN], (const float*)x[F_VN], m->hf_t, 0.0,
                nullptr, m->scal.e12t, L, m->partials);
    } else {
        double** x = m->S[m->cur];
        k_diag_partials<double, double, double><<<g, dim3(TBX, TBY), 0, m->stream>>>(
            x[F_SSHN], x[F_UN], x[F_VN], m->uniform ? nullptr : m->st[S_HT], m->scal.ht, e12t,
            m->scal.e12t, L, m->partials);
    }
    LAUNCHCHK("diag_partials");
    k_checksum_finish<<<1, RED_THREADS, 0, m->stream>>>(m->partials, (int)(g.x * g.y), m->diag);
    LAUNCHCHK("diag_finish");
    HIPCHK(hipMemcpyAsync(m->h_pinned, m->diag, 3 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    for (int k = 0; k < 3; ++k) out[k] = m->h_pinned[k];
    return NEMO2D_OK;
}

int nemo2d_stage_timers(nemo2d_model* m, int on) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_stage_timers: null model");
    HIPCHK(hipStreamSynchronize(m->stream));
    m->tcalls.clear();
    m->tev_used = 0;
    for (int k = 0; k < TR_N; ++k) {
        m->topen[k] = -1;
        m->tcount[k] = 0;
        m->tsum[k] = m->tsq[k] = 0.0;
    }
    m->timing = on != 0;
    return NEMO2D_OK;
}

int nemo2d_stage_times(nemo2d_model* m, int* counts, double* total_s, double* stderr_s) {
    if (!m || !counts || !total_s || !stderr_s) return fail(NEMO2D_ERR_ARG, "nemo2d_stage_times: bad argument");
    timers_fold(m);
    for (int k = 0; k < TR_N; ++k) {
        const long n = m->tcount[k];
        counts[k] = (int)n;
        total_s[k] = m->tsum[k] * 1e-3;
        double se = 0.0;
        if (n > 1) {
            const double mean = m->tsum[k] / n;
            const double var = (m->tsq[k] - n * mean * mean) / (n - 1);
            se = var > 0.0 ? sqrt(var / n) * 1e-3 : 0.0;
        }
        stderr_s[k] = se;
    }
    return NEMO2D_OK;
}

int nemo2d_set_output_grid(nemo2d_model* m, const double* xt, const double* yt, const double* ht) {
    if (!m || !xt || !yt || !ht) return fail(NEMO2D_ERR_ARG, "nemo2d_set_output_grid: bad argument");
    const size_t n = (size_t)m->L.jpi * m->L.jpj;
    m->out_xt.assign(xt, xt + n);
    m->out_yt.assign(yt, yt + n);
    m->out_ht.assign(ht, ht + n);
    return NEMO2D_OK;
}

int nemo2d_snapshot_begin(nemo2d_model* m, int istp, const char* dir) {
    READY(m, "nemo2d_snapshot_begin");
    if (m->mixed) return fail(NEMO2D_ERR_STATE, "nemo2d_snapshot_begin: FP64 model only");
    if (m->out_xt.empty()) return fail(NEMO2D_ERR_STATE, "nemo2d_snapshot_begin: set_output_grid first");
    if (istp < 0 || istp > 99999) return fail(NEMO2D_ERR_ARG, "nemo2d_snapshot_begin: istp out of range");
    int rc = snap_join(m);  // the previous file first (its buffers are reused)
    if (rc) return rc;
    const size_t n = (size_t)m->L.size();
    if (!m->copy_stream) {
        HIPCHK(hipStreamCreateWithFlags(&m->copy_stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&m->ev_state, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&m->ev_copy, hipEventDisableTiming));
        for (int k = 0; k < 3; ++k) HIPCHK(hipHostMalloc((void**)&m->h_snap[k], n * sizeof(double), 0));
    }
    // copies of the now-state, after the work queued so far
    const int src = m->cur;
    double* const* s = m->S[src];
    HIPCHK(hipEventRecord(m->ev_state, m->stream));
    HIPCHK(hipStreamWaitEvent(m->copy_stream, m->ev_state, 0));
    HIPCHK(hipMemcpyAsync(m->h_snap[0], s[F_SSHN], n * sizeof(double), hipMemcpyDeviceToHost, m->copy_stream));
    HIPCHK(hipMemcpyAsync(m->h_snap[1], s[F_UN], n * sizeof(double), hipMemcpyDeviceToHost, m->copy_stream));
    HIPCHK(hipMemcpyAsync(m->h_snap[2], s[F_VN], n * sizeof(double), hipMemcpyDeviceToHost, m->copy_stream));
    HIPCHK(hipEventRecord(m->ev_copy, m->copy_stream));
    m->snap_src = src;
    char name[32];
    snprintf(name, sizeof name, "go2d_%05d.dat", istp);
    std::string path = std::string(dir && *dir ? dir : ".") + "/" + name;
    m->snap_thread = std::thread([m, path] {
        if (hipEventSynchronize(m->ev_copy) != hipSuccess) {
            m->snap_rc = NEMO2D_ERR_HIP;
            m->snap_err = "nemo2d_snapshot: copy failed";
            return;
        }
        const Layout& L = m->L;
        const int jpi = L.jpi, jpj = L.jpj;
        const Arr2 xt{m->out_xt.data(), -1L - jpi, jpi}, yt{m->out_yt.data(), -1L - jpi, jpi};
        const Arr2 ht{m->out_ht.data(), -1L - jpi, jpi};
        const Arr2 ss{m->h_snap[0], L.base, L.pitch}, un{m->h_snap[1], L.base, L.pitch},
            vn{m->h_snap[2], L.base, L.pitch};
        std::string err;
        const int rc = write_go2d(path.c_str(), xt, yt, ht, ss, un, vn, jpi, jpj, err);
        if (rc) {
            m->snap_err = "nemo2d_snapshot: " + err;
            m->snap_rc = rc;
        }
    });
    return NEMO2D_OK;
}

int nemo2d_snapshot_wait(nemo2d_model* m) {
    if (!m) return fail(NEMO2D_ERR_ARG, "nemo2d_snapshot_wait: null model");
    return snap_join(m);
}

int nemo2d_destroy(nemo2d_model* m) {
    if (!m) return NEMO2D_OK;
    if (m->snap_thread.joinable()) m->snap_thread.join();
    if (m->copy_stream) (void)hipStreamSynchronize(m->copy_stream);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (int k = 0; k < 3; ++k)
        if (m->h_snap[k]) (void)hipHostFree(m->h_snap[k]);
    for (hipEvent_t e : m->tev) (void)hipEventDestroy(e);
    if (m->ev_state) (void)hipEventDestroy(m->ev_state);
    if (m->ev_copy) (void)hipEventDestroy(m->ev_copy);
    if (m->copy_stream) (void)hipStreamDestroy(m->copy_stream);
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < NF; ++f) dfree(m->S[k][f]);
    for (int s = 0; s < NS; ++s) dfree(m->st[s]);
    dfree(m->ssha); dfree(m->ua); dfree(m->va);
    dfree(m->flags); dfree(m->flu); dfree(m->flv); dfree(m->rc_plain); dfree(m->rc_plain_short);
    dfree(m->flu_ptr); dfree(m->flu_pts); dfree(m->flv_ptr); dfree(m->flv_pts);
    dfree(m->partials); dfree(m->cs_now); dfree(m->chk); dfree(m->cs_log); dfree(m->cs_part); dfree(m->d_pend);
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < NF; ++f) dfree(m->X[k][f]);
    dfree(m->hf_t); dfree(m->hf_u); dfree(m->hf_v); dfree(m->diag);
    if (m->h_pinned) (void)hipHostFree(m->h_pinned);
    if (m->stage) (void)hipFree(m->stage);
    if (m->ev_free) (void)hipEventDestroy(m->ev_free);
    if (m->ev_tgt) (void)hipEventDestroy(m->ev_tgt);
    if (m->ev_feat) (void)hipEventDestroy(m->ev_feat);
    if (m->ev_incr) (void)hipEventDestroy(m->ev_incr);
    if (m->ev_read) (void)hipEventDestroy(m->ev_read);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return NEMO2D_OK;
}

}  // extern "C"
