/*
 * c3d_host.h -- host-level drop-in for C3D-2015's own entry points (C++), implemented in
 * libc3d_host.so over the C-ABI of libc3d_hip.so.
 *
 * These are the functions and types the reference's update() loop, its init() and its
 * deck reader use, with the reference's names, signatures and argument meaning, so that
 * a build of the reference links against libc3d_host.so instead of compiling
 * src/opt/opt.cpp, src/opt/optBatch.cpp, src/flux_kinetic.cpp, src/flux_moment.cpp,
 * src/update.cpp, src/comm.cpp (serial branch) and src/input_deck_reader.cpp:
 *
 *   OPTIONS, OUTS, opt()                   src/opt/opt.h:14-43, opt.cpp:330-355
 *   class Optimization (batched closure)   src/opt/opt.h:63-91, optBatch.cpp:289-472
 *   solveFluxKinetic / solveFluxMoment     src/global.h:37-38
 *   update(dt) / communicateBoundaries()   src/global.h:35-36, update.cpp:185-275,
 *                                          comm.cpp:277-323
 *   class InputDeckReader                  src/input_deck_reader.h:12-29
 *
 * State the reference keeps in globals (src/global.h:55-140) is read from globals of the
 * same names; libc3d_host.so defines them as weak symbols, so a program that defines its
 * own (the reference's global.cpp) supplies them.  Arrays stay host arrays in the
 * reference's layouts; each call stages them through device memory of a context built
 * from the reference's own tables (g_w, g_m, g_m2, g_xi, g_eta, g_mu).
 *
 * Error behaviour follows the reference: a cell with u0 <= 0 or an exhausted
 * regularisation ladder prints the reference's message and aborts the process
 * (optBatch.cpp:233-238, 319-323; utils.cpp:26-33).  HIP errors print and abort too (the
 * reference prints and continues with garbage, cuda_solveH.cu:13-20).
 */
#ifndef C3D_HOST_H
#define C3D_HOST_H

#include <stdint.h>

#include <list>
#include <stack>

/* ---------------------------------------------------------------- closure, one cell */
struct OPTIONS {
    int maxIter;
    double tolAbs; /* set by the caller, not used (update.cpp:55) */
    double tolRel;
    double tolGamma;
    double condHMax;
    bool useGaunt;
    bool useGauntSparse;
};

struct OUTS {
    double r;
    int iter;
    int iterGamma;
    double gamma;
    double normG;
};

/* One cell with opt()/optScaled semantics (relative tolerance), host pointers; u and alpha
 * updated in place as the reference does. */
void opt(int nm, int nm2, int nq, double *u, double *alpha, OPTIONS options, double *w,
         double *p, double *p2, OUTS *outs);

/* ---------------------------------------------------------------- closure, batched */
class Optimization {
public:
    Optimization(int i, int j, int k, int n, double *u, double *alpha);
    void preProcess(int n);
    void postProcess(int n);

    /* Solves every Optimization on c_stack1 (and c_stack2) with the batched semantics of
     * optBatch.cpp:143-282 and the globals g_mnMaxIter, g_mnTol, g_mnTolGamma,
     * g_mnCondHMax; finished objects go to c_stack3 with c_u, c_alpha post-processed and
     * the counters set. */
    static void opt(int nm, int nm2, int nq);

    static std::stack<Optimization *> *c_stack1;
    static std::stack<Optimization *> *c_stack2;
    static std::stack<Optimization *> *c_stack3;

    double *c_u;
    double *c_alpha;
    double *c_h;
    double *c_alpha0;
    double *c_uOriginal;
    double c_t;
    double c_r;
    double c_err;
    int c_iter;
    int c_iterGamma;
    int c_totalIter;
    int c_i, c_j, c_k;
    int c_regIndex;
    double c_u0;
    bool c_preProcessed;
};

/* ---------------------------------------------------------------- grid entry points */
void solveFluxKinetic(double *u, double *flux, double *alpha, double dx, double dy, double dz);
void solveFluxMoment(double *u, double *flux, double dx, double dy, double dz);
void update(double dt);
void communicateBoundaries();

/* ---------------------------------------------------------------- input deck */
class InputDeckReader {
public:
    static const int VALUE_CHAR_LEN = 128;

    void readInputDeck(const char *filename);
    bool getValue(const char *key, char value[VALUE_CHAR_LEN]);
    bool getValue(const char *key, int *value);
    bool getValue(const char *key, double *value);
    bool getValue(const char *key, bool *value);
    void print();

private:
    struct KeyValue {
        char key[VALUE_CHAR_LEN];
        char value[VALUE_CHAR_LEN];
    };
    std::list<KeyValue> keyValueList;
};

/* ---------------------------------------------------------------- globals read */
enum SOLVER_TYPE { SOLVER_PN, SOLVER_MN };

struct OPTIMIZATION_STATS {
    double r;
    int64_t iter;
    int64_t iterGamma;
    double gamma;
    double normG;
};

extern int g_node;
extern SOLVER_TYPE g_solver;
extern int g_nx, g_ny, g_nz;
extern double g_dx, g_dy, g_dz;
extern int g_gX[4], g_gY[4], g_gZ[4];
extern int g_numQuadPoints, g_numMoments, g_numMoments2;
extern int g_pnUseKineticFlux;
extern double *g_m, *g_m2, *g_w, *g_xi, *g_eta, *g_mu;
extern double *g_u, *g_sigmaT, *g_sigmaS;
extern double g_theta;
extern int g_numUpdates;
extern double *g_updateTimes;
extern double *g_mnAlpha;
extern int g_mnMaxIter;
extern double g_mnTol, g_mnCondHMax, g_mnTolGamma;
extern bool g_mnUseGaunt, g_mnUseGauntSparse, g_mnBatched;
extern int g_mnMaxIterHist, g_mnMaxGammaIter;
extern int *g_mnHistReg;
extern double g_mnIterMean, g_mnIterGammaMean;
extern int g_mnNumDualSolves;
extern OPTIMIZATION_STATS *g_optStats;
extern bool g_outputOptStats;
extern double *g_beta, *g_betaVal; /* Gaunt tables (src/opt/opt.h:56-59), read when */
extern int *g_betaColInd, *g_betaRowPtr; /* g_mnUseGaunt is true                    */

/* ---------------------------------------------------------------- C helpers (tests, FFI) */
extern "C" {
/* Read `path` with InputDeckReader and copy the (lower-cased) value of `key` into value
 * (VALUE_CHAR_LEN bytes); returns 1 if the key is present, 0 otherwise. */
int c3d_host_deck_value(const char *path, const char *key, char *value);
/* The precision mode (C3D_PREC_*) of the contexts behind the host entry points. */
int c3d_host_set_precision(int mode);
/* Resident mode of update(dt) (on = 1): the grid (g_u, g_mnAlpha, g_sigmaT/S) is uploaded
 * on the next call and stays on the device across calls; g_u and g_mnAlpha are refreshed
 * only by c3d_host_sync_to_host() (call it before reading them, e.g. before output), and
 * host-side changes to them are not seen until c3d_host_set_resident is called again.  The
 * run statistics are folded from device-side sums (c3d_stats_summary): counts and maxima as
 * in the per-cell fold, the two means up to rounding.  on = 0 syncs and returns to staging
 * the grid through host memory on every call (the default).  No counterpart in the
 * reference, whose update() works on host arrays. */
int c3d_host_set_resident(int on);
int c3d_host_sync_to_host(void);
/* The batched closure as update.cpp:105-178 drives it, for n cells (host arrays [n][nm],
 * tables from the globals g_w, g_m, ...): one Optimization per cell on c_stack1,
 * Optimization::opt, results collected from c_stack3 into u, alpha and (if not NULL)
 * totalIter[n], regIndex[n]. */
int c3d_host_optimize_cells(int64_t n, int nm, int nm2, int nq, double *u, double *alpha,
                            int *totalIter, int *regIndex);
}

#endif
