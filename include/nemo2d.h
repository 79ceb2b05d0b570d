/*
 * nemo2d -- C-ABI of the MI355X (gfx950) NEMOLite2D time-stepper.
 *
 * Drop-in for the reference CUDA port's extern "C" boundary
 * (cuda/cuda_nemolite2d.cu:234-270, bound from Fortran in
 * cuda/nemolite2d.f90:62-127).  Differences from that boundary, all in the
 * direction of a safer C-ABI:
 *   - an opaque handle instead of process-global device state
 *     (cuda/cuda_nemolite2d.cu:272-280), so several models can coexist;
 *   - every entry point returns an int status (0 = success) instead of
 *     throwing through Fortran frames (cuda/cuda_utils.cuh:9-24);
 *     nemo2d_last_error() gives the message;
 *   - the Fortran driver owns the grid, mask and initial state and uploads
 *     them (the reference recomputes the grid on the device,
 *     cuda/cuda_nemolite2d.cu:647-771);
 *   - the reference's cuda_momentum_ / cuda_boundary_conditions_ are split
 *     into the nine stage calls of the serial model's step
 *     (original/nemolite2d.f90:375-769).
 *
 * Host arrays use the reference's Fortran layout: column-major, first index
 * fastest, with these shapes (jpi, jpj = interior T-grid size, i.e. the
 * namelist's jpiglo-2, jpjglo-2, original/nemolite2d.f90:133-137):
 *   T  (1:jpi, 1:jpj)      e1t e2t e12t ht sshn ssha
 *   U  (0:jpi, 1:jpj)      e1u e2u e12u gphiu hu sshn_u un ua
 *   V  (1:jpi, 0:jpj)      e1v e2v e12v gphiv hv sshn_v vn va
 *   pt (0:jpi+1, 0:jpj+1)  int: 1 water, 0 land, -1 open boundary
 * All arithmetic is IEEE double; results are bitwise those of the serial
 * model built without fast-math (same operation order, no FMA contraction,
 * sin/sqrt evaluated on the host with the C library).
 *
 * Threading: one host thread per model; all device work of a model is
 * issued on one HIP stream in call order.  Calls that return data
 * (checksums, retrieve) synchronise that stream.
 */
#ifndef NEMO2D_H
#define NEMO2D_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nemo2d_model nemo2d_model;

/* status codes */
#define NEMO2D_OK 0
#define NEMO2D_ERR_ARG 1      /* bad argument (null pointer, size, range)  */
#define NEMO2D_ERR_HIP 2      /* HIP runtime error                        */
#define NEMO2D_ERR_MASK 3     /* mask the serial model would index outside */
#define NEMO2D_ERR_STATE 4    /* call out of order (no grid / no state)   */
#define NEMO2D_ERR_NONFINITE 5 /* a checksum is NaN or Inf                 */

/* Message for the last failing call on this thread (never NULL). */
const char *nemo2d_last_error(void);

/* Library version string. */
const char *nemo2d_version(void);

/* Create a model on the current HIP device.  Replaces
 * cuda_setup_model_params_ (cuda/cuda_nemolite2d.cu:236-246, all scalars by
 * value as there) plus the allocation half of cuda_initialise_grid_
 * (:248).  jpi, jpj are the interior sizes.  The grid must then be given
 * with nemo2d_set_grid and the state with nemo2d_set_state. */
int nemo2d_create(nemo2d_model **model, int jpi, int jpj, double dx, double dy,
                  double dep_const, int nit000, int nitend, int irecord,
                  double rdt, double cbfr, double visc);

/* Upload the grid: mask, metrics, latitudes and depths, in the host layout
 * above.  Replaces the grid half of cuda_initialise_grid_
 * (cuda/cuda_nemolite2d.cu:282-416), which rebuilt them on the device.
 * Static coefficients (Coriolis 2*omega*sin(phi*d2r), Flather sqrt(g/h))
 * are evaluated here on the host.  Uniform metrics and depths are detected
 * and passed to the kernels as scalars. */
int nemo2d_set_grid(nemo2d_model *m, const int *pt,
                    const double *e1t, const double *e2t,
                    const double *e1u, const double *e2u,
                    const double *e1v, const double *e2v,
                    const double *e12t, const double *e12u, const double *e12v,
                    const double *gphiu, const double *gphiv,
                    const double *ht, const double *hu, const double *hv);

/* Upload the now-state (sshn T, sshn_u U, sshn_v V, un U, vn V).  The
 * after-state ua/va starts at zero, as after the serial model's allocation
 * (original/nemolite2d.f90:310-318). */
int nemo2d_set_state(nemo2d_model *m, const double *sshn, const double *sshn_u,
                     const double *sshn_v, const double *un, const double *vn);

/* --- the nine stages of one step, in the serial model's order ----------
 * Each replaces one block of original/nemolite2d.f90; together, called in
 * this order, they are one step.  cuda_continuity_
 * (cuda/cuda_nemolite2d.cu:249), cuda_momentum_ (:250, empty there) and
 * cuda_boundary_conditions_ (:251) are what they replace at the C-ABI. */
int nemo2d_continuity(nemo2d_model *m);              /* :390-410 */
int nemo2d_momentum_u(nemo2d_model *m);              /* :430-517 */
int nemo2d_momentum_v(nemo2d_model *m);              /* :519-609 */
int nemo2d_bc_ssh(nemo2d_model *m, double rtime);    /* :628-645 */
int nemo2d_bc_solid_u(nemo2d_model *m);              /* :648-652 */
int nemo2d_bc_solid_v(nemo2d_model *m);              /* :656-660 */
int nemo2d_bc_flather_u(nemo2d_model *m);            /* :669-680 */
int nemo2d_bc_flather_v(nemo2d_model *m);            /* :684-695 */
int nemo2d_next(nemo2d_model *m);                    /* :704-769; replaces cuda_next_ (:252) */

/* One whole step (original/nemolite2d.f90:375-383, rtime = istp*rdt) as one
 * fused pass over HBM plus a boundary epilogue.  Bitwise equal to the nine
 * stage calls above. */
int nemo2d_step(nemo2d_model *m, int istp);

/* Tidal forcing of the clamped-ssh boundary: ssh = amp*sin(2*3.14159/
 * (period_hours*3600) * rtime).  The default is the serial model's
 * (amp 0.2 m, period 12.42 h, original/nemolite2d.f90:629-630); other values
 * have no counterpart in the reference (ensemble members). */
int nemo2d_set_tide(nemo2d_model *m, double amp, double period_hours);

/* The nine stage calls of a step (nemo2d_continuity .. nemo2d_bc_flather_v,
 * in the serial model's order) are recorded rather than run, and
 * nemo2d_next then runs them with next as one fused step (the same results
 * bit for bit, the same checksum logging); a call out of that order, or any
 * other entry point (retrieve, checksums, snapshot, ...), first runs what was
 * recorded, stage by stage.  on = 0 runs every stage call at once (a kernel
 * per stage, as before this deferral); default on.  No counterpart in the
 * reference. */
int nemo2d_set_stage_fusion(nemo2d_model *m, int on);

/* Choose the kernel of nemo2d_step (results are identical for every
 * choice): 0 default, by grid size (uniform statics: 2 up to 230,000 T cells
 * when the Flather points are all on row 0, else 1; 13 up to 3,000,000;
 * then 10.  Per-point depths with uniform metrics: 1 up to 100,000, 12 up
 * to 10,000,000, then 11.
 * Per-point metrics: 1).  1 one thread per cell (Flather points by an
 * epilogue), 2 the same with Flather points recomputed inline, 3 five-array
 * column-marching with 16-row segments (the kernel for arrays beyond 4 GiB,
 * where the others' 32-bit offsets do not reach), 4 column-marching with
 * face terms shared between neighbouring points (12-row segments), 5 the
 * same reading per-point depths, 6 the same with each lane loading only its
 * own column (neighbours by DPP lane shifts), 7 as 6 dividing by
 * precomputed reciprocals with two exact FMA corrections, 8 and 9 as 6 and
 * 7 with the block's waves kept in step row by row, 10 reading sshn, un and
 * vn only (sshn_u/sshn_v recomputed from sshn as next forms them), loads two
 * rows ahead, reciprocal divisions under one guard per row (the row is done
 * again with IEEE divisions when any numerator is out of range), class-free
 * code on tiles of plain interior points, 11 as 5 reading sshn, un, vn and
 * the depths only, 12 and 13 as 11 and 10 with 8-row segments (more waves
 * on mid-size grids).  Marching (3..13) needs uniform metrics, 3, 4, 6..10
 * and 13 also uniform depths, 5 and 10..13 Flather points only on row 0 with
 * the inner point north; 10..13 need sshn_u/sshn_v to be sshn's
 * interpolations
 * (true after any step; nemo2d_set_state checks the given arrays on the
 * device, and when they are not, the first step after it runs 9 or 5).
 * Other grids use 1 (or 2).  The single-precision form marches where it
 * applies unless 1 is chosen.  No counterpart in the reference: a tuning
 * hook.  nemo2d_last_variant gives the kernel the last nemo2d_step (or fused
 * nemo2d_next) ran. */
int nemo2d_set_variant(nemo2d_model *m, int variant);
int nemo2d_last_variant(nemo2d_model *m, int *variant);

/* Steps nit000..nitend back to back (nemo2d_step each), recording the
 * per-step checksums on the device when checksum recording is on. */
int nemo2d_run(nemo2d_model *m, int nit000, int nitend);

/* Per-step checksums.  When recording is on (capacity > 0), every
 * nemo2d_step / nemo2d_next stores sum|sshn|, sum|un|, sum|vn| over
 * (1:jpi,1:jpj) of the new now-state into a device log (slot istp-nit000
 * for nemo2d_step; the next free slot for nemo2d_next).  capacity = 0 turns
 * recording off. */
int nemo2d_record_checksums(nemo2d_model *m, int capacity);
/* Copies up to max_steps*3 logged values (step-major: ssh, u, v) to out and
 * sets *nsteps to how many steps are logged.  Synchronises. */
int nemo2d_checksum_log(nemo2d_model *m, double *out, int max_steps, int *nsteps);

/* sum|sshn|, sum|un|, sum|vn| over (1:jpi,1:jpj) of the current now-state
 * (the field_checksum of original/nemolite2d.f90:82-85 and
 * shared/dl_esm_inf/finite_difference/src/field_mod.f90:935-956, with a
 * compensated, order-fixed reduction).  Synchronises.  Returns
 * NEMO2D_ERR_NONFINITE (after filling out) when any sum is not finite. */
int nemo2d_checksums(nemo2d_model *m, double out[3]);

/* Copy the state back to host arrays (host layout).  Replaces
 * cuda_retrieve_results_ (cuda/cuda_nemolite2d.cu:259-265): sshn T, un U,
 * vn V, ua U, va V.  Any pointer may be NULL to skip that array. */
int nemo2d_retrieve_results(nemo2d_model *m, double *sshn, double *un, double *vn,
                            double *ua, double *va);
/* Copy sshn_u (U), sshn_v (V) and ssha (T) back; any may be NULL. */
int nemo2d_retrieve_aux(nemo2d_model *m, double *sshn_u, double *sshn_v, double *ssha);

/* Copy one grid array back as the device holds it, in the host layout of
 * nemo2d_set_grid.  which: 0..11 = e1t e2t e1u e2u e1v e2v e12t e12u e12v
 * ht hu hv; 12, 13 = the Coriolis parameter 2*omega*sin(phi*d2r) at U, V
 * points; 14, 15 = the Flather coefficient sqrt(g/h) at the Flather U, V
 * points (0 elsewhere).  Synchronises.  Replaces
 * cuda_retrieve_grid_constants_ (cuda/cuda_nemolite2d.cu:253-257,
 * :566-576), which returns xt, yt and ht; xt and yt stay with the driver
 * here (nemo2d_set_output_grid). */
int nemo2d_retrieve_grid(nemo2d_model *m, int which, double *out);

/* The per-point class words (bit fields of the U, V and T roles each point
 * plays in the step's stages) that nemo2d_set_grid derived from the mask, as
 * the device holds them, in the mask's layout (0:jpi+1, 0:jpj+1).
 * Synchronises.  nemo2d_classify derives the same words on the host from a
 * mask and the depths hu (U layout), hv (V layout), without a device.  No
 * counterpart in the reference, whose kernels test the mask itself
 * (cuda/cuda_nemolite2d.cu:815-902). */
int nemo2d_retrieve_classes(nemo2d_model *m, unsigned short *out);
int nemo2d_classify(int jpi, int jpj, const int *pt, const double *hu, const double *hv,
                    unsigned short *out);

/* Device time in milliseconds of the work issued on the model's stream
 * between nemo2d_timer_start and nemo2d_timer_stop (HIP events on that
 * stream).  timer_stop synchronises. */
int nemo2d_timer_start(nemo2d_model *m);
int nemo2d_timer_stop(nemo2d_model *m, double *ms);

/* --- single-precision form (no counterpart in the reference) ------------
 * Same model, FP32 state (sshn_u, sshn_v, un, vn) and FP32 stencil
 * arithmetic; the new ssh is accumulated as sshn + increment in FP64 with
 * sshn kept in double when compensated != 0, in FP32 otherwise.  Needs
 * uniform metrics and latitude (depths may vary per point).  Only the fused
 * step (nemo2d_step / nemo2d_run) is available; the stage calls return
 * NEMO2D_ERR_STATE.  Results are not bitwise those of the serial model:
 * their error against the FP64 path is measured by the tests and bench. */
int nemo2d_create_mixed(nemo2d_model **model, int jpi, int jpj, double dx, double dy,
                        double dep_const, int nit000, int nitend, int irecord,
                        double rdt, double cbfr, double visc, int compensated);

/* FP64 diagnostics of the current now-state over the T points (no
 * counterpart in the reference): out[0] = volume sum(e12t*ssh),
 * out[1] = energy sum(e12t*(0.5*(ht+ssh)*(ubar^2+vbar^2) + 0.5*g*ssh^2))
 * with ubar, vbar the means of each cell's two faces, out[2] = wet area
 * sum(e12t).  Compensated, order-fixed reduction.  Synchronises. */
int nemo2d_diagnostics(nemo2d_model *m, double out[3]);

/* --- ensembles of basins (no counterpart in the reference) ---------------
 * nmem members, each the serial model's basin (land ring, open south edge,
 * f-plane at 50 N, jphgr_msh=1 grid of original/nemolite2d.f90:180-279)
 * with its own interior size jpi[k] x jpj[k], constant depth dep[k] (m) and
 * tide amp[k] (m) / period_h[k] (hours); common dx, dy, rdt, cbfr, visc.
 * precision 64: each member is bitwise the single-model FP64 path
 * (nemo2d_create + nemo2d_set_tide on the same basin); precision 32: the
 * single-precision form with FP64 ssh accumulation.  All members start at
 * rest.  One kernel over (member, tile) work items advances every member
 * by one step. */
typedef struct nemo2d_ensemble nemo2d_ensemble;
int nemo2d_ensemble_create(nemo2d_ensemble **e, int nmem, const int *jpi, const int *jpj,
                           const double *dep, const double *amp, const double *period_h,
                           double dx, double dy, double rdt, double cbfr, double visc,
                           int precision);
/* Back to rest. */
int nemo2d_ensemble_reset(nemo2d_ensemble *e);
/* Autotuner: time every candidate tiling over nsteps steps, keep the fastest
 * (index in *choice, ms per step in *best_ms) and reset to rest. */
int nemo2d_ensemble_tune(nemo2d_ensemble *e, int nsteps, int *choice, double *best_ms);
int nemo2d_ensemble_set_tiling(nemo2d_ensemble *e, int choice);
/* Work-item shape (columns x rows of cells) of tiling `choice`: choices
 * 0..7 are blocks of 256 cells, one per thread; choice 8 is one wave's
 * strip of 63 columns by 12 rows, column-marching with shared face terms;
 * choices 9 and 10 the same over strips of 61 columns reading sshn, un, vn
 * only (sshn_u/sshn_v formed from sshn as next forms them, and filled in
 * before a step of another choice), class-free on plain work items: 9 on a
 * resident grid of blocks taking the items in a cost-balanced order, 10 (the
 * default) one wave per item, neighbouring strips of a segment in one block
 * and each XCD a contiguous run of items. */
int nemo2d_ensemble_tiling(int choice, int *tx, int *ty);
/* Steps nit000..nitend; use_graph != 0 replays a HIP graph of 8 captured
 * steps (each the step kernel and the checksum finish, whose last block
 * advances a device control word holding the step index and the now-state
 * buffer) while 8 or more steps remain, and launches the rest one by one;
 * the graph is captured on first use and kept until the tiling or a buffer
 * it names changes.  *ms = device time of the steps.  Per-member, per-step
 * checksums are logged. */
int nemo2d_ensemble_run(nemo2d_ensemble *e, int nit000, int nitend, int use_graph, double *ms);
int nemo2d_ensemble_checksum_log(nemo2d_ensemble *e, int member, double *out, int max_steps,
                                 int *nsteps);
/* sshn (T), un (U), vn (V) of one member, host layout. */
int nemo2d_ensemble_retrieve(nemo2d_ensemble *e, int member, double *sshn, double *un, double *vn);
int nemo2d_ensemble_destroy(nemo2d_ensemble *e);

/* --- learned-closure network kernels (no counterpart in the reference) ---
 * Device pointers, NHWC activations with the channel stride rounded up to
 * 8 (BF16 stored as 16-bit integers), 1 <= channels <= 32, stream a
 * hipStream_t (NULL = default stream).  Convolutions are 'same'-padded 3x3,
 * implicit-GEMM on BF16 matrix cores with FP32 accumulation; an image has
 * at most 2^24 pixels (h * w; larger: NEMO2D_ERR_ARG).
 * conv_fwd writes the pre-activation b + conv(x) as FP32 into z and/or as
 *   BF16 into zb, and BF16 gelu(it) (use_gelu) or it into act; any of the
 *   three may be NULL (not all).
 *   wf: forward weights [9][32][pad8(cin)], wb: data-gradient weights
 *   [9][32][pad8(cout)] (both from nemo2d_nn_pack_weights of FP32
 *   W[9][cin][cout]). */
int nemo2d_nn_conv_fwd(const void *x, int n, int h, int w, int cin, const void *wf,
                       const float *bias, int cout, float *z, void *zb, void *act, int use_gelu,
                       void *stream);
int nemo2d_nn_conv_bwd_data(const void *dz, int n, int h, int w, int cout, const void *wb,
                            int cin, float *dx, void *stream);
/* The data gradient with the layer below's GELU backward fused in: dzp
 * (BF16, [P][pad8(cin)]) = BF16(dX * GELU'(zp)), zp the layer below's
 * pre-activation [P][pad8(cin)], BF16 bits if zp_bf16 else FP32; padded
 * channels get 0. */
int nemo2d_nn_conv_bwd_data_gelu(const void *dz, int n, int h, int w, int cout, const void *wb,
                                 int cin, const void *zp, int zp_bf16, void *dzp, void *stream);
/* Work buffer size (floats) for nemo2d_nn_conv_bwd_weight over `pixels`
 * (per-block partial sums of the weight and bias gradients; the block count
 * is capped, so the size does not grow with the batch). */
long nemo2d_nn_wgrad_work_floats(long pixels);
int nemo2d_nn_conv_bwd_weight(const void *x, const void *dz, int n, int h, int w, int cin,
                              int cout, float *dw, float *db, float *work, void *stream);
/* nemo2d_nn_conv_bwd_weight without a work buffer and without allocating:
 * the block partials go to a fixed 18.9 MB device array of the library
 * (one per GPU, held until the process ends), so the call makes no
 * hipMalloc/hipFree and needs no workspace from the caller.  No counterpart
 * in the reference: for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid it).  Any batch up to
 * 2147483647 pixels of images of at most 2^24 pixels (larger:
 * NEMO2D_ERR_ARG); the results are bitwise those
 * of nemo2d_nn_conv_bwd_weight.  Calls that overlap on one GPU share the
 * array and race. */
int nemo2d_nn_conv_bwd_weight_noalloc(const void *x, const void *dz, int n, int h, int w, int cin,
                                      int cout, float *dw, float *db, void *stream);
/* dz = dact * GELU'(z), BF16. */
int nemo2d_nn_gelu_bwd(const float *dact, const float *z, long count, void *dz, void *stream);
/* Weighted MSE over channels < ch (1 <= ch <= 8) of [pixels][8] FP32
 * tensors: gradient dy (BF16 [pixels][8], 0 in channels >= ch) and per-block
 * loss parts (nblocks floats, to be summed). */
int nemo2d_nn_mse(const float *y, const float *t, long pixels, int ch, const float *wch,
                  void *dy, float *part, int nblocks, void *stream);
int nemo2d_nn_adam(float *p, const float *g, float *m, float *v, long n, float lr, float b1,
                   float b2, float eps, int step, void *stream);
int nemo2d_nn_pack_weights(const float *w, int cin, int cout, void *wf, void *wb, void *stream);
int nemo2d_nn_to_bf16(const float *a, long n, void *b, void *stream);

/* Learned closure on the model's device state (FP64 model; no counterpart
 * in the reference).  Pixel p = (j-1)*jpi + (i-1) stands for T cell (i, j)
 * with its east face U(i, j) and north face V(i, j).
 * features: x (device, BF16 bits [jpj*jpi][8]) gets, per pixel, BF16 of
 *   FP32((value - mean[c]) / scale[c]) for value = sshn, un, vn, ht of the
 *   cell (c = 0..3); channels 4..7 are 0.  Asynchronous: the kernel writing x
 *   waits for the work queued so far on the default stream (the network,
 *   which may still read x), and work queued later on the default stream
 *   waits for x; the host does not wait.
 * apply: adds (double)z[p][c] * scale[c] (z device FP32 [jpj*jpi][zstride])
 *   to sshn (every cell, c = 0), un on the east faces with i < jpi (c = 1)
 *   and vn on the north faces with j < jpj (c = 2), then re-derives sshn_u
 *   and sshn_v from the new sshn by the next-stage rules.  Asynchronous:
 *   the model's stream waits for the work queued so far on the default
 *   stream (where z is produced), and work queued later on the default
 *   stream waits for the read of z; the host does not wait. */
int nemo2d_closure_features(nemo2d_model *m, const double *mean, const double *scale, void *x);
int nemo2d_closure_apply(nemo2d_model *m, const float *z, int zstride, const double *scale);
/* restrict: the coarse model's now-state becomes the restriction of the fine
 *   model's (fine grid exactly 2x the coarse one per side; coarse cell (I, J)
 *   covers fine cells 2I-1..2I x 2J-1..2J): sshn the mean of the 4 fine
 *   cells ((a+b)+c)+d times 0.25, un on a coarse U face the fine-depth-
 *   weighted mean (h0*u0 + h1*u1)/(h0 + h1) of the 2 fine U faces on it, vn
 *   likewise; sshn_u, sshn_v from the new sshn by the next-stage rules
 *   (original/nemolite2d.f90:735-762), 0 where next leaves a point alone.
 *   As nemo2d_set_state, ua/va restart at zero.  Synchronous.
 * target: y (device FP32 [jpj*jpi][8]) gets, per pixel, FP32((truth - pred)
 *   / scale[c]) for the cell's sshn (c = 0), east-face un (1) and north-face
 *   vn (2), and 0 in channels 3..7 (two models of one size).  Asynchronous,
 *   on pred's stream: the kernel waits for truth's stream and for the work
 *   queued so far on the default stream (which may still read y); work
 *   queued later on the default stream waits for y, and on truth's stream
 *   for the read of truth's state; the host does not wait. */
int nemo2d_closure_restrict(nemo2d_model *fine, nemo2d_model *coarse);
int nemo2d_closure_target(nemo2d_model *truth, nemo2d_model *pred, const double *scale, float *y);

/* Device memory for the closure network's tensors, on the current HIP
 * device (no counterpart in the reference).  Copies and zeroing are
 * synchronous with the host and ordered after work on the default stream. */
int nemo2d_dev_alloc(long bytes, void **out);
int nemo2d_dev_free(void *p);
int nemo2d_dev_upload(void *dst, const void *src, long bytes);
int nemo2d_dev_download(void *dst, const void *src, long bytes);
int nemo2d_dev_zero(void *p, long bytes);
/* Wait for all work on the current device (a timing aid). */
int nemo2d_dev_synchronize(void);

/* Device timers of the serial model's timed regions (dl_timer's
 * 'Continuity', 'Momentum', 'BCs', 'Next', original/nemolite2d.f90:391-769,
 * shared/dl_timer/src/dl_timer.f90:713-773) for the nine-call step, plus
 * the fused step as one region: on != 0 starts recording HIP events around
 * each region on the model's stream (and clears the sums); stage_times
 * fills, for regions 0..4 = Continuity, Momentum (momentum_u..momentum_v),
 * BCs (bc_ssh..bc_flather_v), Next, fused step, the call count, the total
 * seconds and the standard error of the mean per-call time.  A stage call
 * recorded for the fused pass (nemo2d_set_stage_fusion) counts with no
 * time: its work is timed in the Next region that runs it.  No
 * counterpart in the reference's C-ABI (its driver times on the host). */
int nemo2d_stage_timers(nemo2d_model *m, int on);
int nemo2d_stage_times(nemo2d_model *m, int *counts, double *total_s, double *stderr_s);

/* --- snapshot files (go2d_NNNNN.dat), original/nemolite2d.f90:773-799 ---
 * The reference's `output` retrieves the state and writes the file before
 * stepping on (cuda/nemolite2d.f90:496-500 retrieves synchronously).  Here
 * the copy and the file run behind the steps:
 * set_output_grid: T-point coordinates and depth (host T arrays), copied once.
 * snapshot_begin: queues device-to-host copies of sshn, un, vn on a second
 *   stream after the work queued so far, and a host thread that writes
 *   dir/go2d_<istp %05d>.dat in the reference's format once they land.
 *   Later steps are queued at once (a step that would overwrite the buffer
 *   being copied waits for the copy on the device).  A previous snapshot is
 *   finished first.  FP64 model only.
 * snapshot_wait: waits for the file; returns its write status.
 * format_go2d: the file's text for host arrays in the reference's layout
 *   (xt, yt, ht, sshn T; un U; vn V) into buf (cap bytes; buf NULL: *len
 *   = the size only).  Host code; needs no device. */
int nemo2d_set_output_grid(nemo2d_model *m, const double *xt, const double *yt, const double *ht);
int nemo2d_snapshot_begin(nemo2d_model *m, int istp, const char *dir);
int nemo2d_snapshot_wait(nemo2d_model *m);
int nemo2d_format_go2d(int jpi, int jpj, const double *xt, const double *yt, const double *ht,
                       const double *sshn, const double *un, const double *vn, char *buf, long cap,
                       long *len);

/* Practical bandwidth ceiling for the fused step's traffic: a kernel that
 * reads five FP64 arrays of n elements and writes five (elementwise, no
 * stencil) is timed over `repeats` launches on the current device; *gbps
 * gets 80 * n bytes per launch over the mean launch time.  Allocates and
 * frees 10 * n doubles.  No counterpart in the reference (a measurement
 * aid). */
int nemo2d_stream_bandwidth(long n, int repeats, double *gbps);

/* Free the model's device memory and stream.  Replaces cuda_finalise_
 * (cuda/cuda_nemolite2d.cu:267), without resetting the device. */
int nemo2d_destroy(nemo2d_model *m);

#ifdef __cplusplus
}
#endif
#endif /* NEMO2D_H */
