/*
 * c3d_hip.h -- C-ABI of the MI355X (gfx950) implementation of the C3D-2015 per-timestep
 * hot path: the M_N dual Newton closure, the kinetic and P_N moment fluxes, the SSP-RK2
 * stage update and the periodic ghost fill.
 *
 * Two layers:
 *   1. Drop-in replacements for the reference's accelerator seam, src/cuda/cuda_headers.h,
 *      with identical argument lists and host-pointer, synchronous semantics
 *      (solveFluxKinetic_hip, solveH_hip, solveHInit_hip).
 *   2. A device-resident API (c3d_*): an explicit context holding the quadrature/basis
 *      tables in HBM, device pointers in, device pointers out, one stream argument,
 *      int status codes instead of abort().
 *
 * Layouts (as the reference's, src/global.h:15-19):
 *   moments/alphas per cell contiguous:  idx(i,j,k,m) = m + nm*(k + (nz+4)*(j + (ny+4)*i))
 *   for grids (2 ghost layers per side), or cell*nm + m for closure batches;
 *   basis tables M[q*nm + m], M2[q*nm2 + k]; Hessians row-major nm x nm.
 * All arithmetic is IEEE FP64 unless a function says otherwise.
 *
 * Streams are passed as void* (a hipStream_t; NULL = the null stream).
 */
#ifndef C3D_HIP_H
#define C3D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define C3D_OK 0
#define C3D_ERR_ARG 1          /* bad argument (sizes, NULL pointers, unsupported order) */
#define C3D_ERR_HIP 2          /* a HIP runtime call failed */
#define C3D_ERR_U0 3           /* a cell had u0 <= 0 (reference aborts: optBatch.cpp:319-323) */
#define C3D_ERR_LADDER 4       /* a cell exhausted the regularisation ladder
                                  (reference dumps and aborts: optBatch.cpp:233-238) */
#define C3D_ERR_UNSUPPORTED 5  /* option combination not implemented */
#define C3D_ERR_BUDGET 6       /* cells still unconverged after the pass budget
                                  (c3d_set_pass_budget); their u, alpha are left as given */
#define C3D_ERR_SCRATCH 7      /* c3d_closure_noalloc: the cells do not fit its fixed scratch */

/* per-cell outcome flags (c3d_cell_stats.flags) */
#define C3D_FLAG_ISO_RESTART 1 /* isotropic restart fired (optBatch.cpp:208-218) */
#define C3D_FLAG_U0_NONPOS 2
#define C3D_FLAG_LADDER 4
#define C3D_FLAG_BUDGET 8      /* cell left unsolved by the pass budget */
#define C3D_FLAG_FP64_FALLBACK 16 /* mixed mode: the cell was solved by the FP64 path from
                                     its input (the mixed iteration declined it) */
#define C3D_FLAG_FP64_FINISH 64   /* mixed mode: near its optimum (||g|| <= switch_tol) the
                                     cell was too ill-conditioned for the split-FP16 Hessian
                                     and FP64 Hessians finished it from its iterate */

/* arithmetic modes of a context (c3d_set_precision) */
#define C3D_PREC_FP64 0          /* FP64; quadrature sums on the matrix cores (default) */
#define C3D_PREC_FP64_REFORDER 1 /* FP64 in the reference's operation order with a
                                    deterministic exp/log: bit-reproducible on a CPU */
#define C3D_PREC_MIXED 2         /* closure: the Hessian on the FP16 matrix cores as three
                                    split products (about 22 bits, FP32 accumulation; for
                                    M_1..M_3 the cells' weights times an image of the basis
                                    pair products) under a log-sum-exp shift; exponents, f,
                                    the gradient (on the FP64
                                    matrix cores), Cholesky, the condition indicator and the
                                    stopping test in FP64; ill-conditioned cells continue on
                                    FP64 Hessians from their iterate near the optimum, or
                                    restart from their input on the FP64 path before it */

typedef struct c3d_ctx_s *c3d_ctx;

/* Closure options: OPTIONS (src/opt/opt.h:14-23) plus the semantics switch.
 *   unbatched = 0: Optimization::opt semantics (optBatch.cpp:143-282): absolute tol,
 *                  tolGamma from the deck, the custom dpotrs (utils.cpp:71-86).
 *   unbatched = 1: opt()/optScaled semantics (opt.cpp:139-355): tol relative to
 *                  ||u/u0||_2 (OPTIONS.tolRel), tolGamma from the options (update() passes
 *                  1.1, update.cpp:57; c3d_update does the same), LAPACK dpotrs order.
 *   useGaunt = 1 (dense beta, or its CSR with useGauntSparse = 1; fobj.cpp:59-95) asks for H
 *                  as sum_k beta_ijk G_k.  The library checks once per context that beta
 *                  reproduces the basis products m_i m_j at every ordinate (to 1e-10 of their
 *                  largest value), where that H equals the direct form's, and then evaluates
 *                  the direct form; otherwise, and in C3D_PREC_FP64_REFORDER (whose contract
 *                  is the reference's bits), it returns C3D_ERR_UNSUPPORTED. */
typedef struct {
    int32_t maxIter;        /* MN_MAX_ITER */
    int32_t useGaunt;       /* MN_USE_GAUNT_COEF */
    int32_t useGauntSparse; /* MN_USE_GAUNT_SPARSE */
    int32_t unbatched;      /* MN_BATCHED false */
    double tol;             /* MN_TOL */
    double tolGamma;        /* MN_TOL_GAMMA */
    double condHMax;        /* MN_COND_H_MAX */
} c3d_closure_opts;

/* Per-cell outcome: OUTS (src/opt/opt.h:26-33) with the batched counters. 40 bytes. */
typedef struct {
    double r;           /* regularisation r of the final rung */
    double err;         /* final ||grad h||_2 on the u0-scaled problem */
    double gamma;       /* final modcfl value */
    int32_t totalIter;  /* Newton iterations including restarts and rungs */
    int32_t iterGamma;  /* iterations that met tol but not the gamma test */
    int32_t regIndex;   /* ladder rung 0..4 */
    int32_t flags;      /* C3D_FLAG_* */
} c3d_cell_stats;

/* Sums over a range of c3d_cell_stats (c3d_stats_summary): what update.cpp:145-174 folds
 * into its run statistics. */
typedef struct {
    int64_t cells;
    int64_t sumIter, sumIterGamma; /* totalIter, iterGamma summed */
    int64_t maxIter, maxIterGamma; /* their maxima (0 for no cells) */
    int64_t rung[5];               /* cells with r within 1e-10 of 0, 1e-3, 1e-2, 1e-1, 0.5 */
} c3d_stats_sum;

/* ======================================================================
 * 1. Drop-in replacements for src/cuda/cuda_headers.h (host pointers, synchronous)
 * ====================================================================== */

/* Replaces solveFluxKinetic_cuda (src/cuda/cuda_headers.h:10-13; called at
 * src/flux_kinetic.cpp:84-92).  nx,ny,nz include the ghost layers.  Writes flux on the
 * interior cells and zero on the ghosts, like the reference kernel; theta = 2 as in
 * src/cuda/cuda_flux_kinetic.cu:82 (use c3d_flux_kinetic for the deck's THETA). */
void solveFluxKinetic_hip(int nx, int ny, int nz, int nm, int nq, double dx, double dy,
                          double dz, double *alpha, double *flux, double *xi, double *eta,
                          double *mu, double *w, double *p, int isPn);

/* Replaces solveHInit_cuda (src/cuda/cuda_headers.h:19-21; called at
 * src/opt/optBatch.cpp:364-365).  Uploads w, p = M, p2 = M2 and the Gaunt tables. */
void solveHInit_hip(int nm, int nm2, int nq, int batchSize, double *w, double *p, double *p2,
                    double *beta, double *betaVal, int *betaColInd, int *betaRowPtr,
                    int useGaunt);

/* Replaces solveH_cuda (src/cuda/cuda_headers.h:15-17; called at
 * src/opt/optBatch.cpp:393): hs[b] = H(alphas[b]) for b < batchSize, row-major nm x nm,
 * lower triangle mirrored as solveHMn does (src/opt/fobj.cpp:34-102). */
void solveH_hip(int nm, int nm2, int nq, int batchSize, double *alphas, double *hs,
                int useGaunt, int useGauntSparse);

/* Status (C3D_OK or C3D_ERR_*) of the most recent call of one of the three functions above,
 * which return void as the reference's do and print their failures.  Their device staging
 * buffers are kept between calls and grown on demand (no allocation per call at a steady
 * size).  No counterpart in the reference (whose CUDA errors are printed and ignored,
 * src/cuda/cuda_solveH.cu:13-20). */
int c3d_seam_status(void);

/* ======================================================================
 * 2. Device-resident API
 * ====================================================================== */

/* Context with tables built by the library from the moment order N and the quadrature
 * order Q (init.cpp:423-548; Gaunt coefficients init.cpp:287-337; P_N operators
 * init.cpp:199-234).  N in 1..4.  Binds to the current HIP device. */
int c3d_create(c3d_ctx *ctx, int moment_order, int quad_order);

/* Same, with caller-supplied tables (as solveHInit_cuda receives them).  M2/beta/CSR may
 * be NULL when the Gaunt form is not used; nnz = betaRowPtr[nm*nm]. */
int c3d_create_from_tables(c3d_ctx *ctx, int moment_order, int quad_order, const double *w,
                           const double *xi, const double *eta, const double *mu,
                           const double *M, const double *M2, const double *beta, int nnz,
                           const double *betaVal, const int32_t *betaColInd,
                           const int32_t *betaRowPtr);

int c3d_destroy(c3d_ctx ctx);

/* The library's own tables on the host, without a device (for inspection and tests):
 * w, xi, eta, mu [nq]; M [nq][nm]; M2 [nq][nm2]; beta [nm*nm][nm2]; pnOps [6][nm][nm];
 * any pointer may be NULL.  *nnz receives the Gaunt CSR size; betaVal/betaColInd
 * (capacity nnzCap) and betaRowPtr [nm*nm+1] are filled when non-NULL and large enough. */
int c3d_tables_host(int moment_order, int quad_order, double *w, double *xi, double *eta,
                    double *mu, double *M, double *M2, double *beta, double *pnOps, int *nnz,
                    int nnzCap, double *betaVal, int32_t *betaColInd, int32_t *betaRowPtr);

/* nm, nm2, nq and the Gaunt nnz of a context. */
int c3d_sizes(c3d_ctx ctx, int *nm, int *nm2, int *nq, int *nnz);

/* Copy the context's tables to host arrays (any may be NULL).  pnOps: 6 x nm x nm in the
 * order XiPlus, XiMinus, EtaPlus, EtaMinus, MuPlus, MuMinus. */
int c3d_get_tables(c3d_ctx ctx, double *w, double *xi, double *eta, double *mu, double *M,
                   double *M2, double *beta, double *betaVal, int32_t *betaColInd,
                   int32_t *betaRowPtr, double *pnOps);

/* M_N closure over ncells independent cells, FP64 (replaces the closure of
 * update.cpp:105-178 / Optimization::opt, optBatch.cpp:350-472, and of
 * update.cpp:43-99 / opt(), opt.cpp:330-355).  u, alpha: device, AoS [ncells][nm],
 * updated in place exactly as the reference updates them (alpha solved; u rescaled,
 * regularised where a rung fired).  stats: device [ncells] or NULL.  Work is queued on
 * `stream` as passes over a shrinking worklist of unconverged cells; the call waits on the
 * stream every 8 passes to test for an empty worklist.  Cell errors are reported by
 * c3d_sync.  Scratch: about (8*nm + nm*nm + 12)*8 bytes per cell, owned by the context. */
int c3d_closure(c3d_ctx ctx, int64_t ncells, double *u, double *alpha,
                const c3d_closure_opts *opts, c3d_cell_stats *stats, void *stream);

/* c3d_closure without any device allocation, for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  Same arguments, semantics and results as c3d_closure (the same kernels;
 * bit-identical outputs), but all of its scratch lives in one fixed device array of
 * C3D_NOALLOC_SCRATCH_BYTES that every device the library runs on holds from its first
 * context until the process ends; its address is looked up at context creation.  Limits:
 *   bytes(ncells) = ncells*((8 nm + 12 + nm^2)*8 + 44) + 32, plus, with the warm start on,
 *   256-byte alignment + ncells*(100 + 8 nm);  it must not exceed the array, which gives at
 *   most about 256k (nm 4), 98k (nm 9), 41.7k (nm 16), 19.9k (nm 25) cells, or 204k, 87k,
 *   39k, 19k with the warm start.  Larger batches return C3D_ERR_SCRATCH and do nothing.
 * No profiling events are recorded (c3d_get_closure_profile reports passes only).  A Gaunt
 * request is served only after c3d_closure or c3d_hessian has checked the context's beta
 * (else C3D_ERR_UNSUPPORTED).  Every call on one device shares that device's array and
 * overwrites it (each call initialises what it reads): calls that overlap on one device --
 * from other threads, streams or contexts -- race and must be serialised by the caller. */
#define C3D_NOALLOC_SCRATCH_BYTES (128u << 20)
int c3d_closure_noalloc(c3d_ctx ctx, int64_t ncells, double *u, double *alpha,
                        const c3d_closure_opts *opts, c3d_cell_stats *stats, void *stream);

/* Select the arithmetic of the closure and of the kinetic-flux ansatz (C3D_PREC_*).
 * C3D_PREC_FP64_REFORDER evaluates H(alpha) entry by entry in ordinate order
 * (solveHMn's loop, fobj.cpp:42-69) and uses an exp/log built from + - * / (Cody-Waite
 * reduction, Taylor and atanh series), so that a CPU evaluation of the same method gives
 * identical bits, decisions and outputs; it is several times slower than C3D_PREC_FP64.
 * No counterpart in the reference (whose results depend on the host libm). */
int c3d_set_precision(c3d_ctx ctx, int mode);

/* Parameters of C3D_PREC_MIXED.  The mixed phase makes the FP64 stopping test itself (its
 * gradient is summed in FP64), so most cells converge in it.  A cell whose gradient norm is at
 * most switch_tol (default 5e-4) continues from its iterate on FP64 Hessians
 * (C3D_FLAG_FP64_FINISH) when its condition indicator (max_i L_ii / min_i L_ii)^2 -- a lower
 * bound on cond_2(H) that stands in for dpocon in this phase -- exceeds 1e5 (see
 * c3d_set_launch), its factorisation fails, the indicator exceeds cond_max (default 1e8), its
 * line search fails or it has taken max_iter iterations (default 40).  The same events farther
 * from the optimum restart the cell from its input on the FP64 path, which then takes exactly
 * the FP64 mode's decisions (regularisation included; C3D_FLAG_FP64_FALLBACK).  No counterpart
 * in the reference. */
int c3d_set_mixed(c3d_ctx ctx, double cond_max, int max_iter, double switch_tol);

/* Kernel-time accounting of c3d_closure (HIP events on the closure's stream around every
 * launch), for throughput and roofline reporting. */
typedef struct {
    int64_t cell_sweeps;       /* cell evaluations of H(alpha) by the FP64 sweep kernel */
    int64_t tail_cell_sweeps;  /* cell evaluations of H(alpha) inside the fused tail kernel */
    int64_t mixed_cell_sweeps; /* cell evaluations of H(alpha) by the mixed-phase sweep kernel */
    int32_t passes;            /* worklist passes (a tail launch counts its pass limit) */
    int32_t tail_launches;
    double sweep_ms;           /* summed duration of the FP64 sweep kernels */
    double mixed_sweep_ms;     /* summed duration of the mixed-phase sweep kernels */
    double step_ms;            /* summed duration of the Newton-step kernels */
    double tail_ms;            /* summed duration of the fused tail kernels */
    double warm_ms;            /* warm start: features, network, selection */
} c3d_closure_profile;

int c3d_set_profiling(c3d_ctx ctx, int enable);
/* Profile of the most recent c3d_closure call (waits for its stream). */
int c3d_get_closure_profile(c3d_ctx ctx, c3d_closure_profile *out);

/* Upper bound on the sweep passes one c3d_closure call may run (default 100000).  A pass
 * evaluates every unconverged cell once (a Newton iteration or one line-search trial; a
 * round of the fused kernel's four-wave form settles up to 32, see c3d_set_tail_wide); the
 * reference has no such bound (its worst case is 1400 iterations x ~1100 backtracking
 * steps per cell).  No counterpart in the reference. */
int c3d_set_pass_budget(c3d_ctx ctx, int max_passes);

/* Once at most max_cells cells remain unconverged, c3d_closure switches from one launch pair
 * per pass (sweep kernel over all listed cells, then step kernel) to a fused kernel in which
 * one block runs up to passes_per_launch sweep+step rounds for its cell.  Results are
 * identical either way (same operations in the same order per cell).  max_cells = 0
 * disables the fused kernel.  Defaults 8192 and 32.  No counterpart in the reference. */
int c3d_set_tail(c3d_ctx ctx, int max_cells, int passes_per_launch);

/* Once at most max_cells cells remain, the fused kernel runs four wavefronts per cell that
 * evaluate f at the next steps of the backtracking line search (t, t/2, t/4, ...): one step
 * per wavefront in a search's first round, then 8 (nm <= 9), 2 (nm = 16) or 1 (nm = 25) each
 * (in the mixed phase 4, or 2 at nm = 25: the rows of the exponent products), so a round
 * settles up to 4-32 trials of the sequential search, with the same accept/reject decisions
 * and results.  0 keeps one wavefront per cell.  Default 512.  No counterpart in the
 * reference. */
int c3d_set_tail_wide(c3d_ctx ctx, int max_cells);

/* Launch and hand-over parameters of the closure kernels, for tuning per (N, Q); results do not
 * depend on the launch shapes.  No counterpart in the reference.
 *   C3D_LAUNCH_MIXED_WAVES        2 or 3, accepted and without effect: the mixed sweep (the
 *                                 column sweep) has one launch shape per order (eight waves
 *                                 per block for M_1..M_3, four for M_4)
 *   C3D_LAUNCH_MIXED_FINISH_COND  10 * log10 of the condition indicator above which a mixed
 *                                 cell with ||g|| <= switch_tol continues on FP64 Hessians
 *                                 (default 50, i.e. 1e5); changes which cells take that path
 *   C3D_LAUNCH_ASYNC              0 (default): the closure adapts its schedule to the number of
 *                                 unconverged cells, which it reads back on the host now and
 *                                 then (the call returns once the cells are done).  1..10000:
 *                                 the asynchronous schedule, with no host synchronisation: that
 *                                 many bulk passes, each gated on the device (a pass sweeps only
 *                                 while more than the tail threshold of c3d_set_tail are
 *                                 listed), then one fused-kernel launch of passes_per_launch
 *                                 rounds per cell and one four-wave launch for the rest of the
 *                                 pass budget.  The call returns once the work is queued on the
 *                                 stream, and c3d_closure, c3d_update and c3d_closure_noalloc can
 *                                 then be captured in a HIP graph (after one eager call has sized
 *                                 the buffers).  Per-cell results are the same in both
 *                                 schedules; cells still listed after the bulk passes go to the
 *                                 fused kernel whatever their number. */
#define C3D_LAUNCH_MIXED_WAVES 1
#define C3D_LAUNCH_MIXED_FINISH_COND 2
#define C3D_LAUNCH_ASYNC 3
int c3d_set_launch(c3d_ctx ctx, int param, int value);

/* H(alpha) for ncells cells (solveHMn semantics), device pointers, hs [ncells][nm][nm].
 * Under C3D_PREC_MIXED the sums are the mixed closure's (split FP16 products with FP32
 * accumulation under a per-cell shift; relative error ~1e-6).  The
 * Gaunt form is served as c3d_closure_opts describes. */
int c3d_hessian(c3d_ctx ctx, int64_t ncells, const double *alpha, double *hs, int useGaunt,
                int useGauntSparse, void *stream);

/* The dense part of one Newton iteration for n given symmetric nm x nm matrices (device,
 * row-major [n][nm][nm]) and gradients g [n][nm]: Cholesky in LAPACK dpotf2 order (info as
 * dpotrf, optBatch.cpp:199), rcond as dpocon with the max-row-sum norm (optBatch.cpp:201-204),
 * and d = -H^{-1} g by utils.cpp:71-86 (lapackOrder = 0) or LAPACK dpotrs (lapackOrder = 1,
 * opt.cpp:285).  L [n][nm][nm] receives the factor (lower, row-major, zero above).  d is
 * left equal to g and rcond = 0 where info != 0.  nm in {4, 9, 16, 25}. */
int c3d_dense_solve(int nm, int64_t n, const double *H, const double *g, double *d,
                    double *rcond, int32_t *info, double *L, int lapackOrder, void *stream);

/* Wait for `stream`; return C3D_OK, or C3D_ERR_U0 / C3D_ERR_LADDER / C3D_ERR_BUDGET if a
 * closure since the last c3d_sync hit that condition (the per-cell flags say which cells). */
int c3d_sync(c3d_ctx ctx, void *stream);

/* Periodic ghost fill of the 2 ghost layers on the 6 faces (serial branch of
 * communicateBoundaries, comm.cpp:277-323).  u: device grid of the ctx's nm. */
int c3d_ghost_fill(c3d_ctx ctx, int nx, int ny, int nz, double *u, void *stream);

/* Kinetic flux (flux_kinetic.cpp:66-198): flux on interior cells from alpha on all cells;
 * isPn selects the linear P_N ansatz (alpha then holds u).  Needs ansatz scratch: the
 * context allocates (nx+4)(ny+4)(nz+4)*nq doubles on first use and keeps them. */
int c3d_flux_kinetic(c3d_ctx ctx, int nx, int ny, int nz, const double *alpha, double *flux,
                     double dx, double dy, double dz, double theta, int isPn, void *stream);

/* P_N moment flux (flux_moment.cpp:30-104) with the context's half-range operators. */
int c3d_flux_moment(c3d_ctx ctx, int nx, int ny, int nz, const double *u, double *flux,
                    double dx, double dy, double dz, void *stream);

/* Euler stage with isotropic scattering (update.cpp:21-37) on interior cells. */
int c3d_euler(c3d_ctx ctx, int nx, int ny, int nz, double *u, const double *flux,
              const double *sigmaT, const double *sigmaS, double dt, void *stream);

/* uOld <- u (kind 0, update.cpp:206-212) or u <- (u + uOld)/2 (kind 1, update.cpp:268-274)
 * on interior cells. */
int c3d_rk2_combine(c3d_ctx ctx, int nx, int ny, int nz, double *u, double *uOld, int kind,
                    void *stream);

/* One full SSP-RK2 step, update(dt) (update.cpp:185-275): two stages of
 * [ghost fill, closure over all cells (M_N), flux, Euler], then the average.
 * solver: 1 = M_N, 0 = P_N.  pnUseKinetic: P_N flux via the kinetic scheme.
 * Scratch (uOld, flux, ansatz) is owned by the context.  stats may be NULL or a device
 * array of 2*(nx+4)(ny+4)(nz+4) entries (stage 0 then stage 1). */
int c3d_update(c3d_ctx ctx, int nx, int ny, int nz, double *u, double *alpha,
               const double *sigmaT, const double *sigmaS, double dx, double dy, double dz,
               double theta, double dt, int solver, int pnUseKinetic,
               const c3d_closure_opts *opts, c3d_cell_stats *stats, void *stream);

/* ======================================================================
 * 3. Learned warm start (no counterpart in the reference)
 * ======================================================================
 * A multilayer perceptron 25 -> 128 -> 128 -> 25 with GELU activations maps the
 * standardised moments x_k = (u_k/u0 - x_mean_k)/x_std_k (k < nm, zero beyond) of a cell to
 * a standardised initial alpha of the u0-scaled dual problem,
 * alpha_k = alpha_mean_k + alpha_std_k * y_k.  Parameters (FP32, C3D_MLP_NPARAM):
 * W1[25][128] b1[128] W2[128][128] b2[128] W3[128][25] b3[25], z_j = b_j + sum_i x_i W[i][j].
 * With the warm start enabled, c3d_closure evaluates the network for every cell (FP16
 * weights on the matrix cores) and starts Newton from the proposal only if it has a smaller
 * ||grad h||_2 than the caller's alpha, both gradients summed in FP32 (the test only ranks the
 * two candidates; a proposal whose exponentials overflow FP32 loses); otherwise from the
 * caller's alpha.  Newton itself, and every stopping test, stays as the precision mode says. */
#define C3D_MLP_IN 25
#define C3D_MLP_HID 128
#define C3D_MLP_NPARAM 23065
#define C3D_FLAG_WARM 32 /* the closure started from the network's proposal */

typedef struct {
    float x_mean[C3D_MLP_IN], x_std[C3D_MLP_IN];
    float alpha_mean[C3D_MLP_IN], alpha_std[C3D_MLP_IN];
} c3d_mlp_norm;

/* Train the context's network on n (u, alpha) pairs (device, [n][nm], e.g. a closure's
 * output): normalisation from the first 65536 pairs, Glorot-uniform initialisation from
 * `seed`, then mse_steps Adam steps on the mean squared error of the standardised alpha
 * (step size lr, cosine decay) followed by dual_steps Adam steps on the mean dual gap
 * f(alpha_net) - f(alpha_pair) of the u0-scaled problems (step size 0.3 lr, cosine decay;
 * its gradient with respect to alpha is the moment residual, evaluated in FP64 on the
 * context's quadrature).  Minibatches of `batch` pairs are drawn by a counter RNG.
 * loss_first / loss_last (host, may be NULL) receive the first loss of the first phase run
 * and the last loss of the last one.  Synchronous. */
int c3d_mlp_train(c3d_ctx ctx, int64_t n, const double *u, const double *alpha, int mse_steps,
                  int dual_steps, int batch, float lr, uint64_t seed, float *loss_first,
                  float *loss_last, void *stream);
/* Copy the trained parameters (host, C3D_MLP_NPARAM floats) and normalisation out / in. */
int c3d_mlp_get(c3d_ctx ctx, float *params, c3d_mlp_norm *norm);
int c3d_mlp_set(c3d_ctx ctx, const float *params, const c3d_mlp_norm *norm);
/* Enable (1) or disable (0) the warm start in c3d_closure (needs a trained or set network). */
int c3d_set_warmstart(c3d_ctx ctx, int enable);
/* Building blocks, device pointers (for checking the kernels):
 * features X [n][25] (and targets T [n][25] if alpha and T are not NULL) of (u, alpha);
 * Y [n][25] = net(X), FP32 (fp16 = 0) or the FP16 matrix-core inference path (fp16 = 1,
 * then with the context's own network; Y is written as alpha_mean + alpha_std*net in FP64
 * to the double array Yd [n][nm] instead);
 * loss (device float) and grad (device, C3D_MLP_NPARAM) of the mean squared error over
 * outputs k < nm. */
int c3d_mlp_features(c3d_ctx ctx, int64_t n, const double *u, const double *alpha,
                     const c3d_mlp_norm *norm, float *X, float *T, void *stream);
int c3d_mlp_forward(c3d_ctx ctx, int64_t n, const float *params, const float *X, float *Y,
                    void *stream);
int c3d_mlp_infer(c3d_ctx ctx, int64_t n, const float *X, double *Yd, void *stream);
int c3d_mlp_loss_grad(c3d_ctx ctx, int64_t n, const float *params, const float *X,
                      const float *T, float *grad, float *loss, void *stream);
/* The dual-gap loss of the second training phase without the constant f*: loss = mean over
 * rows of f(alpha(net(X_i))) for the u0-scaled moments of u [n][nm] (device), alpha =
 * alpha_mean + alpha_std * net (norm, host), and its gradient (device, C3D_MLP_NPARAM). */
int c3d_mlp_dual_loss_grad(c3d_ctx ctx, int64_t n, const float *params, const float *X,
                           const double *u, const c3d_mlp_norm *norm, float *grad, float *loss,
                           void *stream);

/* Initial state of a grid on the device (init.cpp:480-506 and 262-268): u [ng][nm], alpha
 * [ng][nm] (may be NULL), sigmaT/sigmaS [ng] = crossSection, over all (nx+4)(ny+4)(nz+4)
 * cells with centres a + (i - 2 + 1/2) d.  ic 0: the reference's normalised Gaussian with
 * u0 = max(floor, 2 sqrt(pi) G); ic 1: rho = max(exp(-|x|^2/(2 sigma^2)), floor) times
 * (1 + noise U[-1,1]) from the counter RNG (seed), u0 = 2 sqrt(pi) rho; ic 2: the smooth
 * state (init.cpp:497-502).  a, d: 3 doubles (host). */
int c3d_init_grid(c3d_ctx ctx, int nx, int ny, int nz, const double *a, const double *d, int ic,
                  double sigma, double floor_value, double noise, uint64_t seed,
                  double cross_section, double *u, double *alpha, double *sigmaT,
                  double *sigmaS, void *stream);

/* Sums of ncells per-cell outcomes (device, as c3d_closure/c3d_update write them) into
 * *out (host); blocks on `stream`.  No counterpart in the reference: it replaces the
 * per-cell host fold of update.cpp:145-174 when the grid stays on the device (the host
 * drop-in's resident mode), so the statistics need no per-cell copy. */
int c3d_stats_summary(c3d_ctx ctx, int64_t ncells, const c3d_cell_stats *stats,
                      c3d_stats_sum *out, void *stream);

/* Synthetic closure batch on the device: per cell s ~ U[smin,smax], alpha*_k ~ N(0,s^2),
 * u = sum_q w_q m_q exp(alpha*.m_q), alpha0 from init.cpp:265.  SplitMix64 counter RNG
 * keyed by (seed, cell). alphaStar may be NULL. */
int c3d_make_closure_batch(c3d_ctx ctx, int64_t ncells, uint64_t seed, double smin,
                           double smax, double *u, double *alpha0, double *alphaStar,
                           void *stream);

#ifdef __cplusplus
}
#endif
#endif
