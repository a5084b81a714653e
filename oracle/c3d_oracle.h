/*
 * c3d_oracle.h -- plain FP64 CPU restatement of the C3D-2015 per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker or as the timed
 * CPU baseline.  The product (c3d-2015_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference repository root).  Functions whose parity is pinned only by properties
 * (no reference build, no golden vector) say so in their comment.
 *
 * Layouts follow the reference:
 *   - moments and alphas per cell are contiguous (IU layout, src/global.h:16):
 *       idx(i,j,k,m) = m + nm*(k + NZg*(j + NYg*i)),  NXg = nx+4 etc.
 *   - basis table M[q][m] (IM layout, src/global.h:18), M2[q][k] (IM2, :19)
 *   - Hessian h[i][j] row-major nm x nm (src/opt/fobj.cpp:62-69)
 *   - Cholesky factor column-major lower, L(i,j) at L[i + n*j] (LAPACK 'L')
 */
#ifndef C3D_ORACLE_H
#define C3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NUM_REG 5
#define ORC_NUM_GHOST 2

/* per-cell outcome flags */
#define ORC_FLAG_ISO_RESTART 1   /* isotropic restart fired (optBatch.cpp:208-218) */
#define ORC_FLAG_U0_NONPOS   2   /* u0 <= 0: the reference aborts (optBatch.cpp:319-323) */
#define ORC_FLAG_LADDER      4   /* ladder exhausted: the reference aborts (optBatch.cpp:233-238) */

/* closure options: OPTIONS (src/opt/opt.h:14-23) + the semantics switch */
typedef struct {
    int32_t maxIter;          /* MN_MAX_ITER */
    int32_t useGaunt;         /* MN_USE_GAUNT_COEF */
    int32_t useGauntSparse;   /* MN_USE_GAUNT_SPARSE */
    int32_t unbatched;        /* 0: optBatch.cpp semantics (MN_BATCHED true), 1: opt.cpp semantics */
    double  tol;              /* MN_TOL (absolute in batched mode, relative in unbatched) */
    double  tolGamma;         /* MN_TOL_GAMMA (batched) / OPTIONS.tolGamma (unbatched; update.cpp:57 passes 1.1) */
    double  condHMax;         /* MN_COND_H_MAX */
} orc_opts;

/* per-cell outcome: OUTS (src/opt/opt.h:26-33) + the batched-class counters */
typedef struct {
    double  r;          /* regularisation amount */
    double  err;        /* last ||g||_2 */
    double  gamma;      /* last modcfl value */
    int32_t totalIter;  /* Newton iterations including restarts/regularisations */
    int32_t iterGamma;  /* iterations past tolerance but failing the gamma test */
    int32_t regIndex;   /* ladder rung 0..4 */
    int32_t flags;      /* ORC_FLAG_* */
} orc_stats;

/* tables consumed by the kernels (init.cpp:423-548, 287-337, 199-234) */
typedef struct {
    int32_t N, Q, nm, nm2, nq, nnz;
    double *w, *xi, *eta, *mu;       /* [nq] */
    double *M;                       /* [nq][nm]  */
    double *M2;                      /* [nq][nm2] */
    double *beta;                    /* [nm*nm][nm2] dense Gaunt */
    double *betaVal;                 /* [nnz] */
    int32_t *betaColInd;             /* [nnz] */
    int32_t *betaRowPtr;             /* [nm*nm+1] */
} orc_tables;

/* ---- exp/log on the path: 0 = C library (reference), 1 = deterministic method ---- */
void   orc_set_exp_mode(int mode);
double orc_exp_det(double x);
double orc_log_det(double x);

/* ---- tables ---- */
void   orc_gauss_legendre(int n, double *x, double *w);
double orc_sph_plm(int l, int m, double x);
double orc_real_sh(int m, int n, double mu, double phi);
void   orc_quadrature(int N, int Q, double *w, double *xi, double *eta, double *mu,
                      double *M, double *M2);
/* least-squares Gaunt beta (init.cpp:287-337); returns nnz, or 0 with beta zero when
 * nq < nm2 (underdetermined) */
int    orc_gaunt(int nm, int nm2, int nq, const double *M2, double *beta,
                 double *betaVal, int32_t *betaColInd, int32_t *betaRowPtr);
void   orc_pn_operators(int nm, int nq, const double *w, const double *xi, const double *eta,
                        const double *mu, const double *M, double *ops);

/* ---- dense linear algebra (LAPACK / BLAS semantics) ---- */
int    orc_dpotrf_l(int n, double *A);
double orc_dpocon_l(int n, const double *L, double anorm, int *info);
void   orc_dpotrs_custom(int n, const double *L, double *d);
void   orc_dpotrs_lapack(int n, const double *L, double *d);
double orc_dnrm2(int n, const double *x);
double orc_norm1(int n, const double *A);
double orc_modcfl(int n, const double *d);

/* ---- dual objective ---- */
void   orc_solveH(int nm, int nm2, int nq, const double *alpha, const orc_tables *t,
                  int useGaunt, int useGauntSparse, double *h, double *gExt);
double orc_fobj(int nm, int nq, const double *alpha, const double *u, const double *w,
                const double *M, double *g);
double orc_fobj_useH(int nm, const double *alpha, const double *u, double m0, double *g,
                     const double *h);
double orc_linesearch(int nm, int nq, double *alpha, const double *d, double f,
                      const double *g, const double *u, const double *w, const double *M);

/* ---- closure ---- */
int    orc_closure_cell(int nm, int nm2, int nq, double *u, double *alpha,
                        const orc_opts *o, const orc_tables *t, orc_stats *s);
int    orc_closure(int64_t ncells, int nm, int nm2, int nq, double *u, double *alpha,
                   const orc_opts *o, const orc_tables *t, orc_stats *stats, int nthreads);

/* ---- grid kernels (ghost-inclusive AoS layout) ---- */
void   orc_ghost_fill(int nx, int ny, int nz, int nm, double *u);
void   orc_flux_kinetic(int nx, int ny, int nz, int nm, int nq, const double *alpha,
                        double *flux, const orc_tables *t, double dx, double dy, double dz,
                        double theta, int isPn, double *ansatz, int nthreads);
void   orc_flux_moment(int nx, int ny, int nz, int nm, const double *u, double *flux,
                       const double *ops, double dx, double dy, double dz, int nthreads);
void   orc_euler(int nx, int ny, int nz, int nm, double *u, const double *flux,
                 const double *sigmaT, const double *sigmaS, double dt);
int    orc_update(int nx, int ny, int nz, int nm, int nm2, int nq, int solverMn,
                  int pnUseKinetic, double *u, double *alpha, double *uOld, double *flux,
                  double *ansatz, const double *sigmaT, const double *sigmaS,
                  const orc_tables *t, const double *ops, const orc_opts *o,
                  double dx, double dy, double dz, double theta, double dt,
                  orc_stats *stats, int nthreads);

/* ---- initial conditions and synthetic inputs ---- */
uint64_t orc_splitmix64(uint64_t x);
double   orc_uniform(uint64_t seed, uint64_t idx, uint64_t stream);
double   orc_normal(uint64_t seed, uint64_t idx, uint64_t stream);
void     orc_init_grid(int nx, int ny, int nz, int nm, double ax, double ay, double az,
                       double dx, double dy, double dz, int icKind, double sigma,
                       double floorv, double noise, uint64_t seed, double crossSection,
                       double *u, double *sigmaT, double *sigmaS);
void     orc_init_alpha(int64_t ncells, int nm, const double *u, double *alpha);
void     orc_closure_batch_input(int64_t ncells, int nm, int nq, const double *w,
                                 const double *M, uint64_t seed, double smin, double smax,
                                 double *alphaStar, double *u, double *alpha0);

#ifdef __cplusplus
}
#endif
#endif
