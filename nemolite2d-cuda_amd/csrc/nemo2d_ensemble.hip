// Batched ensembles of NEMOLite2D basins of uneven size (no counterpart in
// the reference).
//
// Every member is the serial model's basin (land ring, open south edge,
// f-plane at 50 N, original/nemolite2d.f90:180-279) with its own size,
// constant depth and tidal amplitude/period.  All members live in pooled
// device arrays, each in its own padded layout.  One step of the whole
// ensemble is one kernel over a list of (member, tile) work items: blocks of
// cells computed one per thread with the single-model path's per-cell body,
// or one wave's column strip marched with the single-model path's marching
// body (by a resident grid of blocks walking the items, or one item per wave
// on a grid the hardware dispatches); either way an FP64 member is bitwise
// the same as that member run alone.  Per-member checksums are reduced in a
// fixed order by a second kernel.  The step index and the now-state buffer
// live in a device control word that the second kernel's last block moves
// on, so one step is captured into a HIP graph once and replayed for every
// step.  The tile shape is chosen by timing the candidates
// (nemo2d_ensemble_tune).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "nemo2d.h"
#include "nemo2d_core.h"
#include "nemo2d_internal.h"
#include "nemo2d_step.h"
#include "nemo2d_march.h"
#include "nemo2d_rc.h"

using namespace nemo2d;

namespace {

#define EHIP(expr)                                                                           \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return set_error(NEMO2D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int EB = 256;  // threads per block of the ensemble step

// Step control in device memory: the step index (relative to the run's
// first step: the tide table's column and the checksum log's slot) and which
// of the two state buffers holds the now-state.  The step kernels read it and
// the last block of k_ens_finish to end moves it on (done counts the blocks
// that have ended), so one captured step can be replayed.
struct EnsCtl {
    int step, cur;
    unsigned done, pad_;
};
// The two now-state buffers (sshn, sshn_u, sshn_v, un, vn each).
struct EnsBufs {
    void* x[2][5];
};

template <class T>
__device__ __forceinline__ T* pick(const EnsBufs& b, int set, int f) {
    return (T*)(set ? b.x[1][f] : b.x[0][f]);
}

struct EnsMember {
    Layout L;      // member layout; offsets relative to the member's base
    long off;      // element offset of the member in the pooled arrays
    Scal s;        // FP64 uniform statics
    ScalF sf;      // FP32 uniform statics (with depths)
    int tile0, ntiles;
};

// Candidate tilings: TX columns x (EB/TX)*ROWS rows per work item of a
// block, one point per thread; the last three are column-marching: a work
// item is one wave's strip of GOUT columns over EMJ rows (shared face
// terms), or of GROUT columns read as sshn, un, vn only (sshn_u/sshn_v formed
// from sshn, plain tiles class-free).  The sshn-only items run either on a
// resident grid of blocks whose waves take the items in a cost-balanced
// order, or (the last tiling) one item per wave on a grid of one block per
// four items that the hardware dispatches as slots free up, the items in
// member, segment, strip order so that a block's four waves march four
// neighbouring strips of one segment together (their halo columns shared
// in the CU's L1) and each XCD gets a contiguous run of segments.
struct Tiling {
    int tx, rows;
};
constexpr int EMJ = 12;  // rows per marching work item
constexpr Tiling TILINGS[] = {{64, 1}, {64, 2}, {64, 4}, {128, 1}, {128, 2}, {256, 1}, {32, 1}, {32, 2},
                              {GOUT, EMJ}, {GROUT, EMJ}, {GROUT, EMJ}};
constexpr int NTILINGS = sizeof(TILINGS) / sizeof(TILINGS[0]);
constexpr int MARCH_TILING = NTILINGS - 3;
constexpr int MARCH_RC_TILING = NTILINGS - 2;   // resident grid, cost-balanced items
constexpr int MARCH_RCD_TILING = NTILINGS - 1;  // dispatched, items in grid order
constexpr int ERC_LA = 2;  // rows loaded ahead by the sshn-only marching item
constexpr bool is_rc(int t) { return t == MARCH_RC_TILING || t == MARCH_RCD_TILING; }
constexpr bool is_march(int t) { return t == MARCH_TILING || is_rc(t); }

template <int TX, int ROWS>
__global__ void __launch_bounds__(EB)
k_ens_step64(const EnsMember* __restrict__ mem, const int4* __restrict__ tiles, int ntiles,
             EnsBufs bufs, const EnsCtl* __restrict__ ctl, const uint16_t* __restrict__ flp,
             const double* __restrict__ flup, const double* __restrict__ flvp, Params p,
             const double* __restrict__ tide_tab, int tide_stride, double* __restrict__ partials) {
    constexpr int TY = EB / TX;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int cur = ctl->cur, step_idx = ctl->step;
    const NState n0{pick<const double>(bufs, cur, 0), pick<const double>(bufs, cur, 1),
                    pick<const double>(bufs, cur, 2), pick<const double>(bufs, cur, 3),
                    pick<const double>(bufs, cur, 4)};
    const AOut a0{pick<double>(bufs, 1 - cur, 0), pick<double>(bufs, 1 - cur, 1), pick<double>(bufs, 1 - cur, 2),
                  pick<double>(bufs, 1 - cur, 3), pick<double>(bufs, 1 - cur, 4)};
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        const EnsMember& M = mem[tl.x];
        const long off = M.off;
        const NState n{n0.sshn + off, n0.sshn_u + off, n0.sshn_v + off, n0.un + off, n0.vn + off};
        const AOut a{a0.sshn + off, a0.sshn_u + off, a0.sshn_v + off, a0.un + off, a0.vn + off};
        const StUniform st{M.s};
        const Layout L = M.L;
        const double tide = tide_tab[(long)tl.x * tide_stride + step_idx];
        DD ct = {0.0, 0.0}, cu = {0.0, 0.0}, cv = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
            cell_step<StUniform, true>(st, n, a, flp + off, flup + off, flvp + off, L, p, tide,
                                       tl.y * TX + tx, (tl.z * ROWS + r) * TY + ty, ct, cu, cv);
        block_reduce3(ct, cu, cv, partials + 6 * (long)t);
        __syncthreads();  // the reduction's shared memory is reused by the next tile
    }
}

template <int TX, int ROWS>
__global__ void __launch_bounds__(EB)
k_ens_step32(const EnsMember* __restrict__ mem, const int4* __restrict__ tiles, int ntiles,
             EnsBufs bufs, const EnsCtl* __restrict__ ctl, const uint16_t* __restrict__ flp,
             const double* __restrict__ flup, const double* __restrict__ flvp, ParamsT<float> p,
             const double* __restrict__ tide_tab, int tide_stride, double* __restrict__ partials) {
    constexpr int TY = EB / TX;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int cur = ctl->cur, step_idx = ctl->step;
    const MixedState<true> n0{pick<const double>(bufs, cur, 0), pick<const float>(bufs, cur, 1),
                              pick<const float>(bufs, cur, 2), pick<const float>(bufs, cur, 3),
                              pick<const float>(bufs, cur, 4)};
    const MixedOut<true> a0{pick<double>(bufs, 1 - cur, 0), pick<float>(bufs, 1 - cur, 1),
                            pick<float>(bufs, 1 - cur, 2), pick<float>(bufs, 1 - cur, 3),
                            pick<float>(bufs, 1 - cur, 4)};
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tiles[t];
        const EnsMember& M = mem[tl.x];
        const long off = M.off;
        const MixedState<true> n{n0.sshn + off, n0.sshn_u + off, n0.sshn_v + off, n0.un + off,
                                 n0.vn + off};
        const MixedOut<true> a{a0.sshn + off, a0.sshn_u + off, a0.sshn_v + off, a0.un + off,
                               a0.vn + off};
        const ScalF sf = M.sf;
        const Layout L = M.L;
        const double tide = tide_tab[(long)tl.x * tide_stride + step_idx];
        DD ct = {0.0, 0.0}, cu = {0.0, 0.0}, cv = {0.0, 0.0};
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
            cell_step_mixed<true, true>(sf, n, a, nullptr, nullptr, nullptr, flp + off, flup + off,
                                        flvp + off, L, p, tide, tl.y * TX + tx,
                                        (tl.z * ROWS + r) * TY + ty, ct, cu, cv);
        block_reduce3(ct, cu, cv, partials + 6 * (long)t);
        __syncthreads();
    }
}

// Column-marching step of the whole ensemble: each wave takes work items
// (member, strip, segment) in turn and writes the item's checksum partial.
template <class R, class SS>
__global__ void __launch_bounds__(EB)
k_ens_march(const EnsMember* __restrict__ mem, const int4* __restrict__ items, int nitems,
            EnsBufs bufs, const EnsCtl* __restrict__ ctl, const uint16_t* __restrict__ flp,
            const double* __restrict__ flvp, ParamsT<R> p, const double* __restrict__ tide_tab,
            int tide_stride, double* __restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int cur = ctl->cur, step_idx = ctl->step;
    const StateG<R, SS, false> n0{pick<const SS>(bufs, cur, 0), pick<const R>(bufs, cur, 1),
                                  pick<const R>(bufs, cur, 2), pick<const R>(bufs, cur, 3),
                                  pick<const R>(bufs, cur, 4), nullptr, nullptr, nullptr};
    const OutG<R, SS> a0{pick<SS>(bufs, 1 - cur, 0), pick<R>(bufs, 1 - cur, 1), pick<R>(bufs, 1 - cur, 2),
                         pick<R>(bufs, 1 - cur, 3), pick<R>(bufs, 1 - cur, 4)};
    // the wave index made wave-uniform (an SGPR): the work item, the member's
    // record and its metrics are then scalar loads held in SGPRs
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int w = blockIdx.x * (EB / 64) + wv; w < nitems; w += gridDim.x * (EB / 64)) {
        const int4 it = items[w];
        if (it.x < 0) continue;  // an empty slot of the item order
        const EnsMember& M = mem[it.x];
        const long off = M.off;
        MetG<R> mt;
        if constexpr (sizeof(R) == sizeof(double)) {
            const Scal& d = M.s;
            mt = MetG<R>{d.e1t, d.e2t, d.e1u, d.e2u, d.e1v, d.e2v, d.e12t, d.e12u, d.e12v,
                         d.fcu, d.fcv, d.ht, d.hu, d.hv};
        } else {
            const ScalF& d = M.sf;
            mt = MetG<R>{d.e1t, d.e2t, d.e1u, d.e2u, d.e1v, d.e2v, d.e12t, d.e12u, d.e12v,
                         d.fcu, d.fcv, d.ht, d.hu, d.hv};
        }
        const StateG<R, SS, false> n{n0.sshn + off, n0.sshn_u + off, n0.sshn_v + off, n0.un + off,
                                     n0.vn + off, nullptr, nullptr, nullptr};
        const OutG<R, SS> a{a0.sshn + off, a0.sshn_u + off, a0.sshn_v + off, a0.un + off, a0.vn + off};
        const double tide = tide_tab[(long)it.x * tide_stride + step_idx];
        CarryG<R, SS> k;
        k.ct = k.cu = k.cv = 0.0;
        march_wave_g<R, SS, false, EMJ>(mt, p, n, a, flp + off, flvp + off, M.L, tide, it.y, it.z, lane, k);
        wave_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)(it.w >> 1));
    }
}

// The same step reading sshn, un, vn only: a work item (member, strip,
// segment, plain) marches one wave's strip of GROUT columns; the new state's
// sshn_u/sshn_v are not written (k_ens_interp fills them in when a kernel
// that reads them comes next).  DISP: one item per wave, block b taking the
// items of block xcd_tile(b) (the dispatched tiling).
template <class R, class SS, bool DISP>
__global__ void __launch_bounds__(EB)
k_ens_march_rc(const EnsMember* __restrict__ mem, const int4* __restrict__ items, int nitems,
               EnsBufs bufs, const EnsCtl* __restrict__ ctl, const uint16_t* __restrict__ flp,
               const double* __restrict__ flvp, ParamsT<R> p, const double* __restrict__ tide_tab,
               int tide_stride, double* __restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int cur = ctl->cur, step_idx = ctl->step;
    const StateG<R, SS, false> n0{pick<const SS>(bufs, cur, 0), pick<const R>(bufs, cur, 1),
                                  pick<const R>(bufs, cur, 2), pick<const R>(bufs, cur, 3),
                                  pick<const R>(bufs, cur, 4), nullptr, nullptr, nullptr};
    const OutG<R, SS> a0{pick<SS>(bufs, 1 - cur, 0), pick<R>(bufs, 1 - cur, 1), pick<R>(bufs, 1 - cur, 2),
                         pick<R>(bufs, 1 - cur, 3), pick<R>(bufs, 1 - cur, 4)};
    // the wave index made wave-uniform (an SGPR): the work item, the member's
    // record and its metrics are then scalar loads held in SGPRs
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b0 = DISP ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    for (int w = b0 * (EB / 64) + wv; w < nitems; w += gridDim.x * (EB / 64)) {
        const int4 it = items[w];
        if (it.x < 0) continue;  // an empty slot of the item order
        const EnsMember& M = mem[it.x];
        const long off = M.off;
        CtxGR<R, SS, false> x;
        if constexpr (sizeof(R) == sizeof(double)) {
            const Scal& d = M.s;
            x.m = MetG<R>{d.e1t, d.e2t, d.e1u, d.e2u, d.e1v, d.e2v, d.e12t, d.e12u, d.e12v,
                          d.fcu, d.fcv, d.ht, d.hu, d.hv};
        } else {
            const ScalF& d = M.sf;
            x.m = MetG<R>{d.e1t, d.e2t, d.e1u, d.e2u, d.e1v, d.e2v, d.e12t, d.e12u, d.e12v,
                          d.fcu, d.fcv, d.ht, d.hu, d.hv};
        }
        x.p = p;
        x.n = StateG<R, SS, false>{n0.sshn + off, n0.sshn_u + off, n0.sshn_v + off, n0.un + off,
                                   n0.vn + off, nullptr, nullptr, nullptr};
        x.a = OutG<R, SS>{a0.sshn + off, a0.sshn_u + off, a0.sshn_v + off, a0.un + off, a0.vn + off};
        x.fl = flp + off;
        x.flv = flvp + off;
        x.L = M.L;
        x.tide = tide_tab[(long)it.x * tide_stride + step_idx];
        const Layout& L = M.L;
        x.i = it.y * GROUT + lane - 1;
        x.il = x.i < L.jpi + 1 ? x.i : L.jpi + 1;
        x.lk = (unsigned)x.il;
        x.own = lane >= 1 && lane <= GROUT && x.i <= L.jpi;
        x.tcol = x.i >= 1 && x.i <= L.jpi;
        x.ucol = x.i >= 0 && x.i <= L.jpi;
        x.vcol = x.i >= 1 && x.i <= L.jpi;
        const int jb = it.z * EMJ;
        const int je = jb + EMJ - 1 < L.jpj ? jb + EMJ - 1 : L.jpj;
        CarryGR<R> k;
        k.ct = k.cu = k.cv = 0.0;
        if (it.w & 1) rcg_march<R, SS, false, ERC_LA, true>(x, k, jb, je);
        else rcg_march<R, SS, false, ERC_LA, false>(x, k, jb, je);
        wave_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)(it.w >> 1));
    }
}

// The FP64 step of the same tiling on the single-grid path's own segment
// march (rc_march): with DM 2 the divisions of plain items go by the
// member's precomputed reciprocals under the per-row guard, which falls
// back to IEEE divisions on a row where any lane's quotient could differ,
// so every member stays bitwise the single run.
static_assert(ROUT == GROUT && RC_NJ == EMJ && RC_LA == ERC_LA,
              "the work items and their plain flags are made for the marching tiling");
template <int DM, bool DISP>
__global__ void __launch_bounds__(EB, 3)  // 3 waves per SIMD, as the single-grid kernel
k_ens_march_rc64(const EnsMember* __restrict__ mem, const int4* __restrict__ items, int nitems,
                 EnsBufs bufs, const EnsCtl* __restrict__ ctl, const uint16_t* __restrict__ flp,
                 const double* __restrict__ flvp, Params p, const double* __restrict__ tide_tab,
                 int tide_stride, double* __restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const int cur = ctl->cur, step_idx = ctl->step;
    const double* const* nb = (const double* const*)bufs.x[cur];
    double* const* ab = (double* const*)bufs.x[1 - cur];
    const NState n0{nb[0], nb[1], nb[2], nb[3], nb[4]};
    const AOut a0{ab[0], ab[1], ab[2], ab[3], ab[4]};
    // the wave index made wave-uniform (an SGPR): the work item, the member's
    // record and its statics are then scalar loads held in SGPRs
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b0 = DISP ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    for (int w = b0 * (EB / 64) + wv; w < nitems; w += gridDim.x * (EB / 64)) {
        const int4 it = items[w];
        if (it.x < 0) continue;  // an empty slot of the item order
        const EnsMember& M = mem[it.x];
        const long off = M.off;
        RcCtx x;
        x.s = M.s;
        x.p = p;
        x.n = NState{n0.sshn + off, n0.sshn_u + off, n0.sshn_v + off, n0.un + off, n0.vn + off};
        x.a = AOut{a0.sshn + off, a0.sshn_u + off, a0.sshn_v + off, a0.un + off, a0.vn + off};
        x.fl = flp + off;
        x.flv = flvp + off;
        x.L = M.L;
        x.tide = tide_tab[(long)it.x * tide_stride + step_idx];
        const Layout& L = x.L;
        // lane 0 is column i0 - 1 (>= -1, inside the left padding)
        x.i = it.y * ROUT + lane - 1;
        x.il = x.i < L.jpi + 1 ? x.i : L.jpi + 1;
        x.lb8 = (unsigned)(8 * x.il);  // wraps for -1; sums stay in range
        x.own = lane >= 1 && lane <= ROUT && x.i <= L.jpi;
        x.tcol = x.i >= 1 && x.i <= L.jpi;
        x.ucol = x.i >= 0 && x.i <= L.jpi;
        x.vcol = x.i >= 1 && x.i <= L.jpi;
        const int jb = it.z * RC_NJ;
        const int je = jb + RC_NJ - 1 < L.jpj ? jb + RC_NJ - 1 : L.jpj;
        RcCarry k;
        k.ct = k.cu = k.cv = 0.0;
        // (items touching a boundary divide by IEEE division: their general
        // code with the guard's bookkeeping would not fit the register budget)
        // (a member whose divisors all pass the one-correction test divides
        // with one correction: the branch is on the item, uniform per wave)
        if (it.w & 1) {
            if (DM == 2 && M.s.one_corr) rc_march<3, true>(x, k, jb, je);
            else rc_march<DM, true>(x, k, jb, je);
        } else {
            rc_march<0, false>(x, k, jb, je);
        }
        wave_reduce3(DD{k.ct, 0.0}, DD{k.cu, 0.0}, DD{k.cv, 0.0}, partials + 6 * (long)(it.w >> 1));
    }
}

// sshn_u, sshn_v of every member from its sshn by next's rules (the points
// next leaves alone keep their values), in the state's arithmetic.
template <class R, class SS>
__global__ void __launch_bounds__(256)
k_ens_interp(const EnsMember* __restrict__ mem, const SS* __restrict__ sshn, R* __restrict__ sshn_u,
             R* __restrict__ sshn_v, const uint16_t* __restrict__ fl) {
    const EnsMember& M = mem[blockIdx.y];
    const Layout& L = M.L;
    R e12t, e12u, e12v;
    if constexpr (sizeof(R) == sizeof(double)) {
        e12t = M.s.e12t; e12u = M.s.e12u; e12v = M.s.e12v;
    } else {
        e12t = M.sf.e12t; e12u = M.sf.e12u; e12v = M.sf.e12v;
    }
    const long npts = (long)(L.jpi + 1) * (L.jpj + 1);
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < npts; q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % (L.jpi + 1)), j = (int)(q / (L.jpi + 1));
        const long o = M.off + L.at(i, j), P = L.pitch;
        const unsigned f = fl[o];
        if (j >= 1) {
            const unsigned c = sucls(f);
            if (c != NX_SKIP) sshn_u[o] = next_g<R>(c, R(sshn[o]), R(sshn[o + 1]), R(0), e12t, e12u);
        }
        if (i >= 1) {
            const unsigned c = svcls(f);
            if (c != NX_SKIP) sshn_v[o] = next_g<R>(c, R(sshn[o]), R(sshn[o + P]), R(0), e12t, e12v);
        }
    }
}

// Per-member checksums: block m reduces its members' tile partials in order
// into the log slot of the control word's step index.  Then the step is
// done: the block that ends last (every block read the control word before
// it counted itself in done) gives the next step the next index and the
// other buffer, and clears done for it.
__global__ void __launch_bounds__(256)
k_ens_finish(const EnsMember* __restrict__ mem, const double* __restrict__ partials,
             double* __restrict__ log, int cap, EnsCtl* __restrict__ ctl) {
    const EnsMember& M = mem[blockIdx.x];
    const int step = ctl->step, cur = ctl->cur;
    finish_partials(partials + 6 * (long)M.tile0, M.ntiles, log + 3 * ((long)blockIdx.x * cap + step));
    __syncthreads();  // every thread of the block has read the control word
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctl->done, 1u) == gridDim.x - 1) {
            ctl->step = step + 1;
            ctl->cur = 1 - cur;
            ctl->done = 0u;
            __threadfence();
        }
    }
}

}  // namespace

struct nemo2d_ensemble {
    int nmem = 0;
    bool fp32 = false;
    hipStream_t stream = nullptr;
    Params prm{};
    std::vector<int> jpi, jpj;
    std::vector<double> amp, omega;
    std::vector<EnsMember> hmem;
    EnsMember* dmem = nullptr;
    long total = 0;  // pooled elements
    void* X[2][5] = {};
    int cur = 0;
    uint16_t* fl = nullptr;
    double *flu = nullptr, *flv = nullptr;
    int tiling = 0;
    bool uv_stale = false;       // sshn_u/sshn_v of the current state not written
    bool fast_div = true;        // every member's divisors allow division by reciprocals
    std::vector<uint16_t> hfl;   // class words (pooled), for the plain work items
    int4* tiles = nullptr;
    int ntiles = 0, ntiles_cap = 0;
    double* partials = nullptr;
    int grid = 0;     // blocks of the per-cell tilings
    int cus = 256;
    int mgrid[2][2] = {};  // blocks of the marching kernels [fp32][rc]: as many as stay resident
    double* tide_tab = nullptr;  // [member][tide_stride]: tide of the run's steps
    int tide_stride = 0;
    double* log = nullptr;
    int log_cap = 0, log_n = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    EnsCtl* ctl = nullptr;       // device step control
    EnsCtl h_ctl{};              // its value at the start of a run
    // GRAPH_STEPS captured steps (step kernel, checksum finish and advance
    // each), replayed for every run of that many steps; a run's remaining
    // steps are launched one by one.  Rebuilt when the tiling or a buffer it
    // names changes (gen counts reallocations)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_tiling = -1;
    long gen = 0, g_gen = -1;
};

namespace {

// Blocks of a persistent marching kernel: every block resident at once (a
// block that waits for a slot would start its fixed share of the items late).
template <class K>
int resident_grid(const nemo2d_ensemble* e, K kern) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, EB, 0) != hipSuccess || nb < 1) nb = 2;
    return nb * e->cus;
}

// Waves of the marching kernel a tiling runs (its resident grid).
int march_waves(nemo2d_ensemble* e, int tiling) {
    const bool rc = tiling == MARCH_RC_TILING;
    if (tiling == MARCH_RCD_TILING) return 0;  // dispatched: the grid follows the items
    int& grid = e->mgrid[e->fp32 ? 1 : 0][rc ? 1 : 0];
    if (!grid) {
        if (e->fp32)
            grid = rc ? resident_grid(e, k_ens_march_rc<float, double, false>) : resident_grid(e, k_ens_march<float, double>);
        else
            grid = rc ? (e->fast_div ? resident_grid(e, k_ens_march_rc64<2, false>) : resident_grid(e, k_ens_march_rc64<0, false>))
                      : resident_grid(e, k_ens_march<double, double>);
    }
    return grid * (EB / 64);
}

// Order of the marching work items: wave w of W runs items w, w+W, w+2W, ...
// of the array.  Blocks go to the eight XCDs in turn (block b to XCD b % 8,
// wave w in block w / 4), and each XCD reads through its own L2, so the
// items, made in member, segment, strip order, are cut into eight runs of
// equal estimated cost (rows, and 1.4x for items that run the general
// code), one per XCD: neighbouring items, which read each other's halo
// rows and columns, then go through the same L2.  Within a run the items
// are sorted by cost (stably: items of one cost keep their order) and dealt
// to the XCD's waves in a snake (round t: waves in order for even t,
// reversed for odd t), which evens out the waves' totals (the kernel ends
// with its slowest wave).  Slots left over in a run's last round hold empty
// items (member -1).  Each item keeps its own checksum slot (w = plain |
// slot << 1), so a member's partials stay contiguous.
constexpr int NXCD = 8;
void order_items(std::vector<int4>& tl, const std::vector<float>& cost, int W) {
    const int n = (int)tl.size();
    constexpr int WB = EB / 64;  // waves per block
    std::vector<std::vector<int>> xw(NXCD);
    for (int w = 0; w < W; ++w) xw[(w / WB) % NXCD].push_back(w);
    // cuts of equal cost, each XCD's share in proportion to its waves
    double total = 0.0;
    for (int q = 0; q < n; ++q) total += cost[q];
    std::vector<int> cut(NXCD + 1, n);
    cut[0] = 0;
    {
        double acc = 0.0, share = 0.0;
        int x = 0, q = 0;
        for (x = 0; x < NXCD - 1; ++x) {
            share += total * (double)xw[x].size() / (double)W;
            while (q < n && acc + 0.5 * cost[q] < share) acc += cost[q++];
            cut[x + 1] = q;
        }
    }
    int rounds = 0;
    for (int x = 0; x < NXCD; ++x)
        if (!xw[x].empty()) {
            const int len = cut[x + 1] - cut[x];
            rounds = std::max(rounds, (len + (int)xw[x].size() - 1) / (int)xw[x].size());
        }
    std::vector<int4> out((size_t)rounds * W, make_int4(-1, 0, 0, 0));
    for (int x = 0; x < NXCD; ++x) {
        const int wx = (int)xw[x].size();
        if (!wx) continue;
        std::vector<int> idx;
        for (int q = cut[x]; q < cut[x + 1]; ++q) idx.push_back(q);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        for (int t0 = 0, r = 0; t0 < (int)idx.size(); t0 += wx, ++r) {
            const int m = std::min(wx, (int)idx.size() - t0);
            const bool rev = r & 1;
            for (int q = 0; q < m; ++q) out[(size_t)r * W + xw[x][rev ? wx - 1 - q : q]] = tl[idx[t0 + q]];
        }
    }
    tl.swap(out);
}

int build_tiles(nemo2d_ensemble* e, int tiling) {
    const Tiling T = TILINGS[tiling];
    // rows per work item (a marching item is one wave's strip)
    const int th = is_march(tiling) ? T.rows : (EB / T.tx) * T.rows;
    std::vector<int4> tl;
    std::vector<float> cost;
    for (int k = 0; k < e->nmem; ++k) {
        EnsMember& M = e->hmem[k];
        const int nbx = (M.L.jpi + 1 + T.tx - 1) / T.tx, nby = (M.L.jpj + 1 + th - 1) / th;
        M.tile0 = (int)tl.size();
        M.ntiles = nbx * nby;
        for (int by = 0; by < nby; ++by)
            for (int bx = 0; bx < nbx; ++bx) {
                int plain = 0;
                if (is_rc(tiling)) {
                    // every class word the wave touches (lanes at columns
                    // bx*GROUT-1 .. +63, rows jb-1 .. je+1+ERC_LA) plain
                    const int i0 = bx * GROUT - 1, i1 = bx * GROUT + 62;
                    const int jb = by * EMJ, je = std::min(jb + EMJ - 1, M.L.jpj);
                    const int j0 = jb - 1, j1 = je + 1 + ERC_LA;
                    if (i0 >= 0 && i1 <= M.L.jpi + 1 && j0 >= 0 && j1 <= M.L.jpj + 1) {
                        plain = 1;
                        for (int j = j0; j <= j1 && plain; ++j)
                            for (int i = i0; i <= i1 && plain; ++i)
                                plain = e->hfl[(size_t)(M.off + M.L.at(i, j))] == G_PLAIN;
                    }
                }
                tl.push_back(make_int4(k, bx, by, plain | ((int)tl.size() << 1)));
                if (is_march(tiling)) {
                    const int rows = std::min(by * EMJ + EMJ - 1, M.L.jpj) - by * EMJ + 1;
                    cost.push_back((plain ? 1.0f : 1.4f) * ((float)rows + 1.5f));
                }
            }
    }
    if (tiling == MARCH_TILING || tiling == MARCH_RC_TILING) order_items(tl, cost, march_waves(e, tiling));
    e->ntiles = (int)tl.size();
    if (e->ntiles > e->ntiles_cap) {
        if (e->tiles) (void)hipFree(e->tiles);
        if (e->partials) (void)hipFree(e->partials);
        e->tiles = nullptr;
        e->partials = nullptr;
        EHIP(hipMalloc((void**)&e->tiles, e->ntiles * sizeof(int4)));
        EHIP(hipMalloc((void**)&e->partials, 6 * (size_t)e->ntiles * sizeof(double)));
        e->ntiles_cap = e->ntiles;
        ++e->gen;
    }
    // uploads on the ensemble's (non-blocking) stream, finished before the
    // host buffers go out of scope
    EHIP(hipMemcpyAsync(e->tiles, tl.data(), tl.size() * sizeof(int4), hipMemcpyHostToDevice, e->stream));
    EHIP(hipMemcpyAsync(e->dmem, e->hmem.data(), e->hmem.size() * sizeof(EnsMember), hipMemcpyHostToDevice,
                        e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    e->tiling = tiling;
    return NEMO2D_OK;
}

EnsBufs bufs_of(const nemo2d_ensemble* e) {
    EnsBufs b;
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < 5; ++f) b.x[k][f] = e->X[k][f];
    return b;
}

template <int TX, int ROWS>
void launch_step_t(nemo2d_ensemble* e) {
    if (!e->fp32) {
        k_ens_step64<TX, ROWS><<<e->grid, EB, 0, e->stream>>>(e->dmem, e->tiles, e->ntiles, bufs_of(e), e->ctl,
                                                              e->fl, e->flu, e->flv, e->prm, e->tide_tab,
                                                              e->tide_stride, e->partials);
    } else {
        ParamsT<float> pf{(float)e->prm.rdt, (float)e->prm.cbfr, (float)e->prm.visc,
                          (float)e->prm.fric};
        k_ens_step32<TX, ROWS><<<e->grid, EB, 0, e->stream>>>(e->dmem, e->tiles, e->ntiles, bufs_of(e), e->ctl,
                                                              e->fl, e->flu, e->flv, pf, e->tide_tab,
                                                              e->tide_stride, e->partials);
    }
}

template <bool RC, bool DISP = false>
void launch_march(nemo2d_ensemble* e) {
    // dispatched: one block per four items, a multiple of 8 blocks (xcd_tile)
    const int grid = DISP ? ((e->ntiles + EB / 64 - 1) / (EB / 64) + 7) / 8 * 8
                          : march_waves(e, RC ? MARCH_RC_TILING : MARCH_TILING) / (EB / 64);
    if (!e->fp32) {
        if constexpr (RC) {
            if (e->fast_div)
                k_ens_march_rc64<2, DISP><<<grid, EB, 0, e->stream>>>(e->dmem, e->tiles, e->ntiles, bufs_of(e),
                                                                       e->ctl, e->fl, e->flv, e->prm, e->tide_tab,
                                                                       e->tide_stride, e->partials);
            else
                k_ens_march_rc64<0, DISP><<<grid, EB, 0, e->stream>>>(e->dmem, e->tiles, e->ntiles, bufs_of(e),
                                                                       e->ctl, e->fl, e->flv, e->prm, e->tide_tab,
                                                                       e->tide_stride, e->partials);
        } else
            k_ens_march<double, double><<<grid, EB, 0, e->stream>>>(
                e->dmem, e->tiles, e->ntiles, bufs_of(e), e->ctl, e->fl, e->flv, e->prm, e->tide_tab,
                e->tide_stride, e->partials);
    } else {
        ParamsT<float> pf{(float)e->prm.rdt, (float)e->prm.cbfr, (float)e->prm.visc,
                          (float)e->prm.fric};
        if constexpr (RC)
            k_ens_march_rc<float, double, DISP><<<grid, EB, 0, e->stream>>>(
                e->dmem, e->tiles, e->ntiles, bufs_of(e), e->ctl, e->fl, e->flv, pf, e->tide_tab,
                e->tide_stride, e->partials);
        else
            k_ens_march<float, double><<<grid, EB, 0, e->stream>>>(
                e->dmem, e->tiles, e->ntiles, bufs_of(e), e->ctl, e->fl, e->flv, pf, e->tide_tab,
                e->tide_stride, e->partials);
    }
}

// fill in the current state's sshn_u/sshn_v (after sshn-only steps)
void launch_interp(nemo2d_ensemble* e) {
    const int k = e->cur;
    const dim3 g(64, e->nmem);
    if (!e->fp32) {
        double** x = (double**)e->X[k];
        k_ens_interp<double, double><<<g, 256, 0, e->stream>>>(e->dmem, x[0], x[1], x[2], e->fl);
    } else {
        void** x = e->X[k];
        k_ens_interp<float, double><<<g, 256, 0, e->stream>>>(e->dmem, (const double*)x[0], (float*)x[1],
                                                              (float*)x[2], e->fl);
    }
}

// One step as the device control word names it: the step kernel of the
// current tiling and the per-member checksum finish, which also advances the
// control word.  Host state (e->cur) is not touched: the caller accounts
// for the steps it issued.
int launch_step(nemo2d_ensemble* e) {
    switch (e->tiling) {
        case MARCH_TILING: launch_march<false>(e); break;
        case MARCH_RC_TILING: launch_march<true>(e); break;
        case MARCH_RCD_TILING: launch_march<true, true>(e); break;
        case 0: launch_step_t<64, 1>(e); break;
        case 1: launch_step_t<64, 2>(e); break;
        case 2: launch_step_t<64, 4>(e); break;
        case 3: launch_step_t<128, 1>(e); break;
        case 4: launch_step_t<128, 2>(e); break;
        case 5: launch_step_t<256, 1>(e); break;
        case 6: launch_step_t<32, 1>(e); break;
        default: launch_step_t<32, 2>(e); break;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string("ensemble step: ") + hipGetErrorString(err));
    k_ens_finish<<<e->nmem, 256, 0, e->stream>>>(e->dmem, e->partials, e->log, e->log_cap, e->ctl);
    err = hipGetLastError();
    if (err != hipSuccess) return set_error(NEMO2D_ERR_HIP, std::string("ensemble finish: ") + hipGetErrorString(err));
    return NEMO2D_OK;
}

// The captured steps for the current tiling and buffers (captured once,
// replayed until one of them changes).  Each replay is one host call, and
// consecutive graph launches leave a gap on the device: GRAPH_STEPS steps
// per graph make it one gap per GRAPH_STEPS steps.
constexpr int GRAPH_STEPS = 8;
int ensure_graph(nemo2d_ensemble* e) {
    if (e->gexec && e->g_tiling == e->tiling && e->g_gen == e->gen) return NEMO2D_OK;
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
    EHIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int s = 0; s < GRAPH_STEPS && !rc; ++s) rc = launch_step(e);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : set_error(NEMO2D_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
    }
    e->graph = g;
    EHIP(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
    e->g_tiling = e->tiling;
    e->g_gen = e->gen;
    return NEMO2D_OK;
}

// Tide values of every member for steps nit000..nitend (host C library),
// in a table whose row length only grows (the captured step names it).
int upload_tides(nemo2d_ensemble* e, int nit000, int nsteps) {
    if (nsteps > e->tide_stride) {
        EHIP(hipStreamSynchronize(e->stream));
        if (e->tide_tab) (void)hipFree(e->tide_tab);
        e->tide_tab = nullptr;
        EHIP(hipMalloc((void**)&e->tide_tab, (size_t)e->nmem * nsteps * sizeof(double)));
        e->tide_stride = nsteps;
        ++e->gen;
    }
    const int ts = e->tide_stride;
    std::vector<double> tab((size_t)e->nmem * ts, 0.0);
    for (int k = 0; k < e->nmem; ++k)
        for (int s = 0; s < nsteps; ++s)
            tab[(size_t)k * ts + s] = tide_eval(e->amp[k], e->omega[k], (double)(nit000 + s) * e->prm.rdt);
    EHIP(hipMemcpyAsync(e->tide_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                        e->stream));
    EHIP(hipStreamSynchronize(e->stream));
    return NEMO2D_OK;
}

int ensure_log(nemo2d_ensemble* e, int nsteps) {
    if (nsteps > e->log_cap) {
        EHIP(hipStreamSynchronize(e->stream));
        if (e->log) (void)hipFree(e->log);
        e->log = nullptr;
        EHIP(hipMalloc((void**)&e->log, 3 * (size_t)e->nmem * nsteps * sizeof(double)));
        e->log_cap = nsteps;
        ++e->gen;
    }
    EHIP(hipMemsetAsync(e->log, 0, 3 * (size_t)e->nmem * e->log_cap * sizeof(double), e->stream));
    e->log_n = 0;
    return NEMO2D_OK;
}

}  // namespace

extern "C" {

int nemo2d_ensemble_reset(nemo2d_ensemble* e) {
    if (!e) return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_reset: null ensemble");
    // rest state: every field zero (sshn = 0 makes sshn_u = sshn_v = 0)
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < 5; ++f) {
            const size_t es = (e->fp32 && f > 0) ? sizeof(float) : sizeof(double);
            EHIP(hipMemsetAsync(e->X[k][f], 0, (size_t)e->total * es, e->stream));
        }
    EHIP(hipStreamSynchronize(e->stream));
    e->cur = 0;
    e->uv_stale = false;  // the rest state's interpolations are zero, as stored
    return NEMO2D_OK;
}

int nemo2d_ensemble_create(nemo2d_ensemble** out, int nmem, const int* jpi, const int* jpj,
                           const double* dep, const double* amp, const double* period_h, double dx,
                           double dy, double rdt, double cbfr, double visc, int precision) {
    if (!out || nmem < 1 || !jpi || !jpj || !dep || !amp || !period_h)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_create: bad argument");
    if (precision != 64 && precision != 32)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_create: precision must be 64 or 32");
    if (!(rdt > 0.0) || !(dx > 0.0) || !(dy > 0.0))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_create: dx, dy, rdt must be > 0");
    *out = nullptr;
    nemo2d_ensemble* e = new nemo2d_ensemble();
    e->nmem = nmem;
    e->fp32 = precision == 32;
    e->prm = Params{rdt, cbfr, visc, 1.0 + cbfr * rdt};
    auto bail = [&](int rc) {
        nemo2d_ensemble_destroy(e);
        return rc;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess)
        return bail(set_error(NEMO2D_ERR_HIP, "nemo2d_ensemble_create: stream/event creation failed"));
    // member layouts in the pools
    long off = 0;
    const double d2r = PI / 180.0;
    const double fc = 2.0 * OMEGA * sin(50.0 * d2r);
    for (int k = 0; k < nmem; ++k) {
        if (jpi[k] < 1 || jpj[k] < 1 || !(dep[k] > 0.0) || !(period_h[k] > 0.0))
            return bail(set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_create: bad member " + std::to_string(k)));
        EnsMember M{};
        M.L = Layout::of(jpi[k], jpj[k]);
        M.off = off;
        off += (M.L.size() + 63) / 64 * 64;
        const double e12 = dx * dy;
        M.s = Scal{dx, dy, dx, dy, dx, dy, e12, e12, e12, dep[k], dep[k], dep[k], fc, fc, {}, 0};
        set_reciprocals(M.s, e->prm.fric);
        e->fast_div = e->fast_div && M.s.fast_div;
        M.sf = ScalF{(float)dx, (float)dy, (float)dx, (float)dy, (float)dx, (float)dy, (float)e12,
                     (float)e12, (float)e12, (float)fc, (float)fc, (float)dep[k], (float)dep[k],
                     (float)dep[k]};
        e->hmem.push_back(M);
        e->jpi.push_back(jpi[k]);
        e->jpj.push_back(jpj[k]);
        e->amp.push_back(amp[k]);
        e->omega.push_back(tide_omega_of(period_h[k]));
    }
    e->total = off;
    if (hipMalloc((void**)&e->dmem, nmem * sizeof(EnsMember)) != hipSuccess ||
        hipMalloc((void**)&e->ctl, sizeof(EnsCtl)) != hipSuccess)
        return bail(set_error(NEMO2D_ERR_HIP, "nemo2d_ensemble_create: allocation failed"));
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < 5; ++f) {
            const size_t es = (e->fp32 && f > 0) ? sizeof(float) : sizeof(double);
            if (hipMalloc(&e->X[k][f], (size_t)off * es) != hipSuccess)
                return bail(set_error(NEMO2D_ERR_HIP, "nemo2d_ensemble_create: state allocation failed"));
        }
    if (hipMalloc((void**)&e->fl, (size_t)off * sizeof(uint16_t)) != hipSuccess ||
        hipMalloc((void**)&e->flu, (size_t)off * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&e->flv, (size_t)off * sizeof(double)) != hipSuccess)
        return bail(set_error(NEMO2D_ERR_HIP, "nemo2d_ensemble_create: allocation failed"));
    // classes of each member's mask (the reference grid)
    std::vector<uint16_t> hfl((size_t)off, 0);
    std::vector<double> hflu((size_t)off, 0.0), hflv((size_t)off, 0.0);
    for (int k = 0; k < nmem; ++k) {
        const EnsMember& M = e->hmem[k];
        const int ni = jpi[k], nj = jpj[k];
        std::vector<int> pt((size_t)(ni + 2) * (nj + 2), 1);
        for (int j = 0; j <= nj + 1; ++j) {
            pt[(size_t)j * (ni + 2)] = 0;
            pt[(size_t)j * (ni + 2) + ni + 1] = 0;
        }
        for (int i = 0; i <= ni + 1; ++i) pt[(size_t)(nj + 1) * (ni + 2) + i] = 0;
        for (int i = 1; i <= ni; ++i) pt[i] = -1;
        std::vector<double> hu((size_t)(ni + 1) * nj, dep[k]), hv((size_t)ni * (nj + 1), dep[k]);
        Classes c;
        std::string err;
        int rc = classify(M.L, pt.data(), hu.data(), hv.data(), c, err);
        if (rc) return bail(set_error(rc, "nemo2d_ensemble_create: member " + std::to_string(k) + ": " + err));
        std::copy(c.fl.begin(), c.fl.end(), hfl.begin() + M.off);
        std::copy(c.flu.begin(), c.flu.end(), hflu.begin() + M.off);
        std::copy(c.flv.begin(), c.flv.end(), hflv.begin() + M.off);
    }
    if (hipMemcpyAsync(e->fl, hfl.data(), hfl.size() * sizeof(uint16_t), hipMemcpyHostToDevice, e->stream) !=
            hipSuccess ||
        hipMemcpyAsync(e->flu, hflu.data(), hflu.size() * sizeof(double), hipMemcpyHostToDevice, e->stream) !=
            hipSuccess ||
        hipMemcpyAsync(e->flv, hflv.data(), hflv.size() * sizeof(double), hipMemcpyHostToDevice, e->stream) !=
            hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
        return bail(set_error(NEMO2D_ERR_HIP, "nemo2d_ensemble_create: upload failed"));
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
    }
    e->grid = 4 * cus;
    e->cus = cus;
    int rc;
    e->hfl = std::move(hfl);
    if ((rc = build_tiles(e, MARCH_RCD_TILING)) || (rc = nemo2d_ensemble_reset(e))) return bail(rc);
    *out = e;
    return NEMO2D_OK;
}

int nemo2d_ensemble_set_tiling(nemo2d_ensemble* e, int choice) {
    if (!e) return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_set_tiling: null ensemble");
    if (choice < 0 || choice >= NTILINGS)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_set_tiling: choice out of range");
    return build_tiles(e, choice);
}

int nemo2d_ensemble_tiling(int choice, int* tx, int* ty) {
    if (choice < 0 || choice >= NTILINGS || !tx || !ty)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_tiling: bad argument");
    *tx = TILINGS[choice].tx;
    *ty = is_march(choice) ? TILINGS[choice].rows : (EB / TILINGS[choice].tx) * TILINGS[choice].rows;
    return NEMO2D_OK;
}

int nemo2d_ensemble_run(nemo2d_ensemble* e, int nit000, int nitend, int use_graph, double* ms) {
    if (!e || nitend < nit000) return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_run: bad argument");
    const int nsteps = nitend - nit000 + 1;
    int rc;
    if ((rc = upload_tides(e, nit000, nsteps)) || (rc = ensure_log(e, nsteps))) return rc;
    if (e->uv_stale && !is_rc(e->tiling)) {  // this tiling reads sshn_u/sshn_v
        launch_interp(e);
        EHIP(hipGetLastError());
        e->uv_stale = false;
    }
    e->h_ctl = EnsCtl{0, e->cur, 0u, 0u};
    EHIP(hipMemcpyAsync(e->ctl, &e->h_ctl, sizeof(EnsCtl), hipMemcpyHostToDevice, e->stream));
    if (use_graph && (rc = ensure_graph(e))) return rc;
    EHIP(hipEventRecord(e->ev0, e->stream));
    // (graph replays while GRAPH_STEPS steps remain, then single launches;
    // each step moves the device control word on, whichever way it came)
    for (int s = 0; s < nsteps && !rc;) {
        if (use_graph && nsteps - s >= GRAPH_STEPS) {
            const hipError_t le = hipGraphLaunch(e->gexec, e->stream);
            if (le != hipSuccess) rc = set_error(NEMO2D_ERR_HIP, std::string("graph launch: ") + hipGetErrorString(le));
            s += GRAPH_STEPS;
        } else {
            rc = launch_step(e);
            ++s;
        }
    }
    EHIP(hipEventRecord(e->ev1, e->stream));
    EHIP(hipEventSynchronize(e->ev1));
    if (rc) return rc;
    if (nsteps & 1) e->cur = 1 - e->cur;
    if (is_rc(e->tiling)) e->uv_stale = true;
    float t = 0.f;
    EHIP(hipEventElapsedTime(&t, e->ev0, e->ev1));
    if (ms) *ms = (double)t;
    e->log_n = nsteps;
    return NEMO2D_OK;
}

int nemo2d_ensemble_tune(nemo2d_ensemble* e, int nsteps, int* choice, double* best_ms) {
    if (!e || nsteps < 1) return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_tune: bad argument");
    int best = 0;
    double bt = 1e300;
    for (int c = 0; c < NTILINGS; ++c) {
        int rc;
        double t = 0.0;
        if ((rc = build_tiles(e, c)) || (rc = nemo2d_ensemble_reset(e)) ||
            (rc = nemo2d_ensemble_run(e, 1, 2, 0, nullptr)) ||   // warm-up
            (rc = nemo2d_ensemble_run(e, 3, 2 + nsteps, 0, &t)))
            return rc;
        if (t < bt) {
            bt = t;
            best = c;
        }
    }
    int rc;
    if ((rc = build_tiles(e, best)) || (rc = nemo2d_ensemble_reset(e))) return rc;
    if (choice) *choice = best;
    if (best_ms) *best_ms = bt / nsteps;
    return NEMO2D_OK;
}

int nemo2d_ensemble_checksum_log(nemo2d_ensemble* e, int member, double* out, int max_steps,
                                 int* nsteps) {
    if (!e || member < 0 || member >= e->nmem || (!out && max_steps > 0))
        return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_checksum_log: bad argument");
    EHIP(hipStreamSynchronize(e->stream));
    const int n = std::min(e->log_n, max_steps);
    if (n > 0)
        EHIP(hipMemcpy(out, e->log + 3 * (size_t)member * e->log_cap, 3 * (size_t)n * sizeof(double),
                       hipMemcpyDeviceToHost));
    if (nsteps) *nsteps = e->log_n;
    return NEMO2D_OK;
}

int nemo2d_ensemble_retrieve(nemo2d_ensemble* e, int member, double* sshn, double* un, double* vn) {
    if (!e || member < 0 || member >= e->nmem)
        return set_error(NEMO2D_ERR_ARG, "nemo2d_ensemble_retrieve: bad argument");
    EHIP(hipStreamSynchronize(e->stream));
    const EnsMember& M = e->hmem[member];
    const Layout& L = M.L;
    const int ni = L.jpi, nj = L.jpj;
    struct {
        double* dst;
        int f, lb1, ub1, lb2, ub2;
    } w[3] = {{sshn, 0, 1, ni, 1, nj}, {un, 3, 0, ni, 1, nj}, {vn, 4, 1, ni, 0, nj}};
    for (auto& q : w) {
        if (!q.dst) continue;
        const long n1 = q.ub1 - q.lb1 + 1, n2 = q.ub2 - q.lb2 + 1;
        const bool dbl = !e->fp32 || q.f == 0;
        const size_t es = dbl ? sizeof(double) : sizeof(float);
        std::vector<char> buf((size_t)n1 * n2 * es);
        const char* src = (const char*)e->X[e->cur][q.f] + (M.off + L.at(q.lb1, q.lb2)) * es;
        EHIP(hipMemcpy2D(buf.data(), n1 * es, src, L.pitch * es, n1 * es, n2, hipMemcpyDeviceToHost));
        for (long k = 0; k < n1 * n2; ++k)
            q.dst[k] = dbl ? ((const double*)buf.data())[k] : (double)((const float*)buf.data())[k];
    }
    return NEMO2D_OK;
}

int nemo2d_ensemble_destroy(nemo2d_ensemble* e) {
    if (!e) return NEMO2D_OK;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (int k = 0; k < 2; ++k)
        for (int f = 0; f < 5; ++f)
            if (e->X[k][f]) (void)hipFree(e->X[k][f]);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    void* ptrs[] = {e->dmem, e->fl, e->flu, e->flv, e->tiles, e->partials, e->tide_tab, e->log, e->ctl};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return NEMO2D_OK;
}

}  // extern "C"
