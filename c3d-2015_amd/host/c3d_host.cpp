// c3d_host.cpp -- the reference's host-level entry points (include/c3d_host.h) over the
// C-ABI of libc3d_hip.so.  Host arrays in the reference's layouts are staged through device
// buffers of one context built from the reference's own tables; the device kernels do the
// work.
#include <hip/hip_runtime.h>

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/c3d_hip.h"
#include "../../include/c3d_host.h"

// ---------------------------------------------------------------- globals (weak: a program
// that defines the reference's globals, src/global.cpp, supplies them instead)
#define C3D_WEAK __attribute__((weak))
C3D_WEAK int g_node = 0;
C3D_WEAK SOLVER_TYPE g_solver = SOLVER_MN;
C3D_WEAK int g_nx = 0, g_ny = 0, g_nz = 0;
C3D_WEAK double g_dx = 0, g_dy = 0, g_dz = 0;
C3D_WEAK int g_gX[4] = {0, 0, 0, 0}, g_gY[4] = {0, 0, 0, 0}, g_gZ[4] = {0, 0, 0, 0};
C3D_WEAK int g_numQuadPoints = 0, g_numMoments = 0, g_numMoments2 = 0;
C3D_WEAK int g_pnUseKineticFlux = 0;
C3D_WEAK double *g_m = nullptr, *g_m2 = nullptr, *g_w = nullptr, *g_xi = nullptr,
                *g_eta = nullptr, *g_mu = nullptr;
C3D_WEAK double *g_u = nullptr, *g_sigmaT = nullptr, *g_sigmaS = nullptr;
C3D_WEAK double g_theta = 2.0;
C3D_WEAK int g_numUpdates = 0;
C3D_WEAK double *g_updateTimes = nullptr;
C3D_WEAK double *g_mnAlpha = nullptr;
C3D_WEAK int g_mnMaxIter = 100;
C3D_WEAK double g_mnTol = 1e-8, g_mnCondHMax = 1e10, g_mnTolGamma = 1.1;
C3D_WEAK bool g_mnUseGaunt = false, g_mnUseGauntSparse = false, g_mnBatched = true;
C3D_WEAK int g_mnMaxIterHist = 0, g_mnMaxGammaIter = 0;
C3D_WEAK int *g_mnHistReg = nullptr;
C3D_WEAK double g_mnIterMean = 0.0, g_mnIterGammaMean = 0.0;
C3D_WEAK int g_mnNumDualSolves = 0;
C3D_WEAK OPTIMIZATION_STATS *g_optStats = nullptr;
C3D_WEAK bool g_outputOptStats = false;
// Gaunt tables (src/opt/fobj.cpp:12-15 defines them; init.cpp fills them)
C3D_WEAK double *g_beta = nullptr;
C3D_WEAK double *g_betaVal = nullptr;
C3D_WEAK int *g_betaColInd = nullptr;
C3D_WEAK int *g_betaRowPtr = nullptr;

std::stack<Optimization *> *Optimization::c_stack1 = nullptr;
std::stack<Optimization *> *Optimization::c_stack2 = nullptr;
std::stack<Optimization *> *Optimization::c_stack3 = nullptr;

// Scratch initialisers the reference calls after setting its thread count (init.cpp:270-271);
// the device path keeps no per-thread scratch.
void fobj_init(int) {}
void linesearch_init(int) {}

namespace {

constexpr double kReg[5] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};  // opt.h:37-38

[[noreturn]] void die(const char *what, int rc)
{
    std::fprintf(stderr, "c3d_host: %s failed (%d)\n", what, rc);
    std::abort();
}

void check(const char *what, int rc)
{
    if (rc != C3D_OK)
        die(what, rc);
}

void hcheck(const char *what, hipError_t e)
{
    if (e != hipSuccess)
        die(what, (int)e);
}

struct DevArray {
    void *p = nullptr;
    size_t bytes = 0;
    void *get(size_t b)
    {
        if (b > bytes) {
            if (p)
                hcheck("hipFree", hipFree(p));
            hcheck("hipMalloc", hipMalloc(&p, b));
            bytes = b;
        }
        return p;
    }
};

// The context behind the host entry points, built from the reference's tables.
struct HostState {
    std::mutex mtx;
    c3d_ctx ctx = nullptr;
    int nm = 0, nq = 0;
    const double *w = nullptr, *m = nullptr, *xi = nullptr, *beta = nullptr;
    int precision = C3D_PREC_FP64;
    DevArray u, alpha, flux, stats, sig;
    // resident mode of update(dt): the grid it advances stays in these between calls
    bool resident = false;
    bool devValid = false;  // ru/ralpha/rsig hold the state of the host arrays below
    bool devAhead = false;  // ... and are newer than them
    const double *hostU = nullptr, *hostA = nullptr;
    size_t gridBytes = 0;
    DevArray ru, ralpha, rsig, rstats;
};
HostState &state()
{
    static HostState s;
    return s;
}

int order_of(int nm)
{
    const int n = (int)std::lround(std::sqrt((double)nm)) - 1;
    return (n + 1) * (n + 1) == nm ? n : -1;
}

int quad_of(int nq)
{
    const int q = (int)std::lround(std::sqrt(nq / 2.0));
    return 2 * q * q == nq ? q : -1;
}

// (Re)build the context when the tables change; caller holds the mutex.  With gaunt, the
// context also takes the reference's Gaunt tables (g_beta and its CSR, which initMn fills
// when MN_USE_GAUNT_COEF is true: init.cpp:287-337), against which the library checks that
// the Gaunt form equals the direct one before it evaluates the latter.
c3d_ctx context(int nm, int nq, const double *w, const double *xi, const double *eta,
                const double *mu, const double *M, const double *M2, bool gaunt)
{
    HostState &S = state();
    const double *beta = gaunt ? g_beta : nullptr;
    if (S.ctx && S.nm == nm && S.nq == nq && S.w == w && S.m == M && S.xi == xi &&
        S.beta == beta)
        return S.ctx;
    if (S.ctx)
        check("c3d_destroy", c3d_destroy(S.ctx));
    S.ctx = nullptr;
    const int N = order_of(nm), Q = quad_of(nq);
    if (N < 1 || N > 4 || Q < 1)
        die("moment/quadrature order", nm * 1000 + nq);
    const double *xiKey = xi;
    std::vector<double> zero;
    if (!xi || !eta || !mu) {
        zero.assign(nq, 0.0);
        xi = eta = mu = zero.data();
    }
    const bool csr = beta && g_betaVal && g_betaColInd && g_betaRowPtr;
    check("c3d_create_from_tables",
          c3d_create_from_tables(&S.ctx, N, Q, w, xi, eta, mu, M, M2, beta,
                                 csr ? g_betaRowPtr[nm * nm] : 0, csr ? g_betaVal : nullptr,
                                 csr ? g_betaColInd : nullptr, csr ? g_betaRowPtr : nullptr));
    check("c3d_set_precision", c3d_set_precision(S.ctx, S.precision));
    S.nm = nm;
    S.nq = nq;
    S.w = w;
    S.m = M;
    S.xi = xiKey;
    S.beta = beta;
    return S.ctx;
}

c3d_ctx grid_context()
{
    return context(g_numMoments, g_numQuadPoints, g_w, g_xi, g_eta, g_mu, g_m, g_m2,
                   g_mnUseGaunt);
}

size_t grid_cells()
{
    return (size_t)(g_gX[3] - g_gX[0] + 1) * (g_gY[3] - g_gY[0] + 1) * (g_gZ[3] - g_gZ[0] + 1);
}

void h2d(void *d, const void *h, size_t b) { hcheck("hipMemcpy", hipMemcpy(d, h, b, hipMemcpyHostToDevice)); }
void d2h(void *h, const void *d, size_t b) { hcheck("hipMemcpy", hipMemcpy(h, d, b, hipMemcpyDeviceToHost)); }

// The reference aborts on these per-cell outcomes (optBatch.cpp:233-238, 319-323).
void closure_errors(int rc)
{
    if (rc == C3D_ERR_U0) {
        std::printf("u[0] < 0\n");
        std::fflush(stdout);
        std::abort();
    }
    if (rc == C3D_ERR_LADDER) {
        std::printf("Max regularization achieved.\n");
        std::fflush(stdout);
        std::abort();
    }
    check("c3d_closure", rc);
}

// The same fold over a whole stage from its device-side sums (resident mode): maxima and rung
// counts as the per-cell fold gives them; the two running means take the stage's sums in one
// step, (N mean + sum) / (N + n), which agrees with the per-cell recurrence up to rounding.
void fold_summary(const c3d_stats_sum &t)
{
    if (t.cells == 0)
        return;
    if (t.maxIter > g_mnMaxIterHist)
        g_mnMaxIterHist = (int)t.maxIter;
    if (t.maxIterGamma > g_mnMaxGammaIter)
        g_mnMaxGammaIter = (int)t.maxIterGamma;
    if (g_mnHistReg)
        for (int r = 0; r < 5; ++r)
            g_mnHistReg[r] += (int)t.rung[r];
    const double n0 = g_mnNumDualSolves, n1 = n0 + (double)t.cells;
    g_mnIterMean = (n0 * g_mnIterMean + (double)t.sumIter) / n1;
    g_mnIterGammaMean = (n0 * g_mnIterGammaMean + (double)t.sumIterGamma) / n1;
    g_mnNumDualSolves += (int)t.cells;
}

// update.cpp:145-174: fold one cell's outcome into the run statistics
void fold_stats(const c3d_cell_stats &s)
{
    if (s.totalIter > g_mnMaxIterHist)
        g_mnMaxIterHist = s.totalIter;
    if (s.iterGamma > g_mnMaxGammaIter)
        g_mnMaxGammaIter = s.iterGamma;
    if (g_mnHistReg)
        for (int r = 0; r < 5; ++r)
            if (std::fabs(s.r - kReg[r]) < 1e-10)
                g_mnHistReg[r]++;
    g_mnIterMean = (g_mnNumDualSolves * g_mnIterMean + s.totalIter) / (g_mnNumDualSolves + 1.0);
    g_mnIterGammaMean =
        (g_mnNumDualSolves * g_mnIterGammaMean + s.iterGamma) / (g_mnNumDualSolves + 1.0);
    g_mnNumDualSolves++;
}

}  // namespace

// ---------------------------------------------------------------- opt(), one cell
void opt(int nm, int nm2, int nq, double *u, double *alpha, OPTIONS options, double *w,
         double *p, double *p2, OUTS *outs)
{
    (void)nm2;
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    c3d_ctx ctx = context(nm, nq, w, nullptr, nullptr, nullptr, p, p2, options.useGaunt);
    c3d_closure_opts o{options.maxIter, options.useGaunt ? 1 : 0, options.useGauntSparse ? 1 : 0,
                       1, options.tolRel, options.tolGamma, options.condHMax};
    double *du = (double *)S.u.get(nm * sizeof(double));
    double *da = (double *)S.alpha.get(nm * sizeof(double));
    c3d_cell_stats *ds = (c3d_cell_stats *)S.stats.get(sizeof(c3d_cell_stats));
    h2d(du, u, nm * sizeof(double));
    h2d(da, alpha, nm * sizeof(double));
    check("c3d_closure", c3d_closure(ctx, 1, du, da, &o, ds, nullptr));
    closure_errors(c3d_sync(ctx, nullptr));
    c3d_cell_stats st;
    d2h(u, du, nm * sizeof(double));
    d2h(alpha, da, nm * sizeof(double));
    d2h(&st, ds, sizeof(st));
    if (outs) {  // opt.cpp:315-322
        outs->iter = st.totalIter;
        outs->iterGamma = st.iterGamma < 0 ? 0 : st.iterGamma;
        outs->gamma = st.gamma;
        outs->normG = st.err;
        outs->r = st.r;
    }
}

// ---------------------------------------------------------------- Optimization
Optimization::Optimization(int i, int j, int k, int n, double *u, double *alpha)
{
    c_i = i;
    c_j = j;
    c_k = k;
    c_u = (double *)std::malloc(n * sizeof(double));
    c_alpha = (double *)std::malloc(n * sizeof(double));
    std::memcpy(c_u, u, n * sizeof(double));
    std::memcpy(c_alpha, alpha, n * sizeof(double));
    c_h = nullptr;
    c_alpha0 = nullptr;
    c_uOriginal = nullptr;
    c_t = 1.0;
    c_r = 0.0;
    c_err = 0.0;
    c_iter = 0;
    c_iterGamma = 0;
    c_totalIter = 0;
    c_regIndex = 0;
    c_u0 = c_u[0];
    c_preProcessed = false;
}

// The device closure scales u by u0 itself (its k_init); these keep the reference's
// object-level contract for callers that pre- or post-process by hand.
void Optimization::preProcess(int n)
{
    const double alphaScale = std::log(c_u0) * 2.0 * std::sqrt(M_PI);
    c_alpha[0] = c_alpha[0] - alphaScale;
    for (int k = 0; k < n; ++k)
        c_u[k] = c_u[k] / c_u0;
    if (c_u0 <= 0) {
        std::printf("u[0] < 0\n");
        std::abort();
    }
    c_preProcessed = true;
}

void Optimization::postProcess(int n)
{
    const double alphaScale = std::log(c_u0) * 2.0 * std::sqrt(M_PI);
    for (int k = 0; k < n; ++k)
        c_u[k] = c_u[k] * c_u0;
    c_alpha[0] = c_alpha[0] + alphaScale;
    c_preProcessed = false;
}

void Optimization::opt(int nm, int nm2, int nq)
{
    (void)nm2;
    if (!c_stack3)
        c_stack3 = new std::stack<Optimization *>;
    std::vector<Optimization *> cells;
    for (std::stack<Optimization *> *st : {c_stack1, c_stack2})
        while (st && !st->empty()) {
            cells.push_back(st->top());
            st->pop();
        }
    if (cells.empty())
        return;
    const size_t n = cells.size();
    std::vector<double> u(n * nm), a(n * nm);
    for (size_t c = 0; c < n; ++c) {
        Optimization *o = cells[c];
        if (o->c_preProcessed)  // the device applies the u0 scaling itself
            o->postProcess(nm);
        std::memcpy(&u[c * nm], o->c_u, nm * sizeof(double));
        std::memcpy(&a[c * nm], o->c_alpha, nm * sizeof(double));
    }
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    c3d_ctx ctx = context(nm, nq, g_w, g_xi, g_eta, g_mu, g_m, g_m2, g_mnUseGaunt);
    c3d_closure_opts o{g_mnMaxIter, g_mnUseGaunt ? 1 : 0, g_mnUseGauntSparse ? 1 : 0, 0, g_mnTol,
                       g_mnTolGamma, g_mnCondHMax};
    double *du = (double *)S.u.get(u.size() * sizeof(double));
    double *da = (double *)S.alpha.get(a.size() * sizeof(double));
    c3d_cell_stats *ds = (c3d_cell_stats *)S.stats.get(n * sizeof(c3d_cell_stats));
    h2d(du, u.data(), u.size() * sizeof(double));
    h2d(da, a.data(), a.size() * sizeof(double));
    check("c3d_closure", c3d_closure(ctx, (int64_t)n, du, da, &o, ds, nullptr));
    closure_errors(c3d_sync(ctx, nullptr));
    std::vector<c3d_cell_stats> st(n);
    d2h(u.data(), du, u.size() * sizeof(double));
    d2h(a.data(), da, a.size() * sizeof(double));
    d2h(st.data(), ds, n * sizeof(c3d_cell_stats));
    for (size_t c = 0; c < n; ++c) {
        Optimization *ob = cells[c];
        std::memcpy(ob->c_u, &u[c * nm], nm * sizeof(double));
        std::memcpy(ob->c_alpha, &a[c * nm], nm * sizeof(double));
        ob->c_r = st[c].r;
        ob->c_err = st[c].err;
        ob->c_totalIter = st[c].totalIter;
        ob->c_iterGamma = st[c].iterGamma;
        ob->c_regIndex = st[c].regIndex;
        c_stack3->push(ob);
    }
}

// ---------------------------------------------------------------- grid entry points
void solveFluxKinetic(double *u, double *flux, double *alpha, double dx, double dy, double dz)
{
    (void)u;
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    c3d_ctx ctx = grid_context();
    const size_t b = grid_cells() * g_numMoments * sizeof(double);
    double *da = (double *)S.alpha.get(b), *df = (double *)S.flux.get(b);
    h2d(da, alpha, b);
    h2d(df, flux, b);
    const int isPn = g_solver == SOLVER_PN ? 1 : 0;
    check("c3d_flux_kinetic", c3d_flux_kinetic(ctx, g_nx, g_ny, g_nz, da, df, dx, dy, dz,
                                               g_theta, isPn, nullptr));
    check("c3d_sync", c3d_sync(ctx, nullptr));
    d2h(flux, df, b);
}

void solveFluxMoment(double *u, double *flux, double dx, double dy, double dz)
{
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    c3d_ctx ctx = grid_context();
    const size_t b = grid_cells() * g_numMoments * sizeof(double);
    double *du = (double *)S.u.get(b), *df = (double *)S.flux.get(b);
    h2d(du, u, b);
    h2d(df, flux, b);
    check("c3d_flux_moment", c3d_flux_moment(ctx, g_nx, g_ny, g_nz, du, df, dx, dy, dz, nullptr));
    check("c3d_sync", c3d_sync(ctx, nullptr));
    d2h(flux, df, b);
}

void communicateBoundaries()
{
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    c3d_ctx ctx = grid_context();
    const size_t b = grid_cells() * g_numMoments * sizeof(double);
    double *du = (double *)S.u.get(b);
    h2d(du, g_u, b);
    check("c3d_ghost_fill", c3d_ghost_fill(ctx, g_nx, g_ny, g_nz, du, nullptr));
    check("c3d_sync", c3d_sync(ctx, nullptr));
    d2h(g_u, du, b);
}

// g_optStats row of one stage (update.cpp:147-156): the cells of the line j = gY[3]/2,
// k = gZ[3]/2 across the interior in x.
void record_line(const c3d_cell_stats &s, int i, int stage)
{
    const int idx = (i - g_gX[1]) + g_nx * (2 * g_numUpdates + stage);
    g_optStats[idx].iter = s.totalIter;
    g_optStats[idx].iterGamma = s.iterGamma;
    g_optStats[idx].r = s.r;
    // the batched path stores iterGamma as gamma (update.cpp:153)
    g_optStats[idx].gamma = g_mnBatched ? s.iterGamma : s.gamma;
    g_optStats[idx].normG = s.err;
}

// update.cpp:185-275 on the device: uOld copy, two stages of [ghost fill, closure over all
// cells, flux, Euler step], RK2 average; statistics folded as update.cpp:145-174 does.  In
// resident mode (c3d_host_set_resident) the grid stays on the device between calls: it is
// uploaded on the first call, the host arrays are refreshed by c3d_host_sync_to_host, and
// the statistics are folded from device-side sums plus the one g_optStats line.
void update(double dt)
{
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    const auto t0 = std::chrono::steady_clock::now();
    c3d_ctx ctx = grid_context();
    const size_t ng = grid_cells();
    const size_t b = ng * g_numMoments * sizeof(double);
    const bool mn = g_solver == SOLVER_MN;
    const bool res = S.resident;
    if (res && (!S.devValid || S.hostU != g_u || S.hostA != (mn ? g_mnAlpha : nullptr) ||
                S.gridBytes != b)) {
        S.devValid = false;  // another grid (or the first call): upload it
        S.devAhead = false;
    }
    double *du = (double *)(res ? S.ru : S.u).get(b);
    double *da = (double *)(res ? S.ralpha : S.alpha).get(b);
    double *dsig = (double *)(res ? S.rsig : S.sig).get(2 * ng * sizeof(double));
    c3d_cell_stats *ds =
        (c3d_cell_stats *)(res ? S.rstats : S.stats).get(2 * ng * sizeof(c3d_cell_stats));
    if (!res || !S.devValid) {
        h2d(du, g_u, b);
        if (mn)
            h2d(da, g_mnAlpha, b);
        h2d(dsig, g_sigmaT, ng * sizeof(double));
        h2d(dsig + ng, g_sigmaS, ng * sizeof(double));
        if (res) {
            S.devValid = true;
            S.hostU = g_u;
            S.hostA = mn ? g_mnAlpha : nullptr;
            S.gridBytes = b;
        }
    }
    c3d_closure_opts o{g_mnMaxIter, g_mnUseGaunt ? 1 : 0, g_mnUseGauntSparse ? 1 : 0,
                       g_mnBatched ? 0 : 1, g_mnTol, g_mnTolGamma, g_mnCondHMax};
    check("c3d_update", c3d_update(ctx, g_nx, g_ny, g_nz, du, mn ? da : nullptr, dsig, dsig + ng,
                                   g_dx, g_dy, g_dz, g_theta, dt, mn ? 1 : 0, g_pnUseKineticFlux,
                                   &o, mn ? ds : nullptr, nullptr));
    closure_errors(c3d_sync(ctx, nullptr));
    const int NY = g_gY[3] - g_gY[0] + 1, NZ = g_gZ[3] - g_gZ[0] + 1;
    const bool line = g_outputOptStats && g_optStats;
    if (res) {
        S.devAhead = true;
        if (mn)
            for (int stage = 0; stage < 2; ++stage) {
                c3d_stats_sum t;
                check("c3d_stats_summary",
                      c3d_stats_summary(ctx, (int64_t)ng, ds + stage * ng, &t, nullptr));
                if (line) {
                    std::vector<c3d_cell_stats> row(g_nx);
                    const size_t c0 = (size_t)(g_gZ[3] / 2) +
                                      (size_t)NZ * ((size_t)(g_gY[3] / 2) + (size_t)NY * g_gX[1]);
                    hcheck("hipMemcpy2D",
                           hipMemcpy2D(row.data(), sizeof(c3d_cell_stats), ds + stage * ng + c0,
                                       (size_t)NY * NZ * sizeof(c3d_cell_stats),
                                       sizeof(c3d_cell_stats), g_nx, hipMemcpyDeviceToHost));
                    for (int i = g_gX[1]; i <= g_gX[2]; ++i)
                        record_line(row[i - g_gX[1]], i, stage);
                }
                fold_summary(t);
            }
    } else {
        d2h(g_u, du, b);
        if (mn) {
            d2h(g_mnAlpha, da, b);
            std::vector<c3d_cell_stats> st(2 * ng);
            d2h(st.data(), ds, st.size() * sizeof(c3d_cell_stats));
            for (int stage = 0; stage < 2; ++stage)
                for (int i = g_gX[0]; i <= g_gX[3]; i++)
                    for (int j = g_gY[0]; j <= g_gY[3]; j++)
                        for (int k = g_gZ[0]; k <= g_gZ[3]; k++) {
                            const size_t c = (size_t)k + NZ * ((size_t)j + (size_t)NY * i);
                            const c3d_cell_stats &s = st[stage * ng + c];
                            if (line && j == g_gY[3] / 2 && k == g_gZ[3] / 2 && i >= g_gX[1] &&
                                i <= g_gX[2])
                                record_line(s, i, stage);
                            fold_stats(s);
                        }
        }
    }
    // the whole step runs on the device: its time is booked to stage 0's flux slot
    if (g_updateTimes) {
        const double secs =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (int k = 0; k < 8; ++k)
            g_updateTimes[g_numUpdates * 8 + k] = (k == 2) ? secs : 0.0;
    }
    g_numUpdates++;
}

// ---------------------------------------------------------------- input deck
// The deck format of input_deck_reader.cpp:17-79: per line (at most 499 characters per
// read), text after '#' ignored; the key is the first token (space or tab separated), the
// value the next token (space separated only, so a tab stays part of the value); a third
// token makes the line an error (reported, dropped); a key without a value is reported and
// still kept with the value of the previous entry, as the reference's reused record does.
// Keys and values compare case-insensitively and values are returned lower-cased.
void InputDeckReader::readInputDeck(const char *filename)
{
    FILE *f = std::fopen(filename, "r");
    if (!f) {
        std::printf("Could not open file.\n");
        return;
    }
    char line[500];
    int lineNo = 0;
    KeyValue kv;
    kv.key[0] = 0;
    kv.value[0] = 0;
    while (std::fgets(line, sizeof(line), f)) {
        ++lineNo;
        if (char *nl = std::strchr(line, '\n'))
            *nl = 0;
        if (char *hash = std::strchr(line, '#'))
            *hash = 0;
        char *save = nullptr;
        char *tok = strtok_r(line, " \t", &save);
        if (!tok)
            continue;
        std::strncpy(kv.key, tok, VALUE_CHAR_LEN);
        kv.key[VALUE_CHAR_LEN - 1] = 0;
        tok = strtok_r(nullptr, " ", &save);
        if (tok) {
            std::strncpy(kv.value, tok, VALUE_CHAR_LEN);
            kv.value[VALUE_CHAR_LEN - 1] = 0;
        } else {
            std::printf("Error in input file at line: %d\n", lineNo);
        }
        if (!tok || strtok_r(nullptr, " ", &save) == nullptr)
            keyValueList.push_back(kv);
        else
            std::printf("Error in input file at line: %d\n", lineNo);
    }
    std::fclose(f);
}

static void lower_copy(const char *in, char *out)
{
    int i = 0;
    for (; in[i]; ++i)
        out[i] = (char)std::tolower((unsigned char)in[i]);
    out[i] = 0;
}

bool InputDeckReader::getValue(const char *key, char value[VALUE_CHAR_LEN])
{
    char want[VALUE_CHAR_LEN], have[VALUE_CHAR_LEN];
    std::strncpy(want, key, VALUE_CHAR_LEN - 1);
    want[VALUE_CHAR_LEN - 1] = 0;
    lower_copy(want, want);
    for (const KeyValue &kv : keyValueList) {
        lower_copy(kv.key, have);
        if (std::strcmp(have, want) == 0) {
            lower_copy(kv.value, value);
            return true;
        }
    }
    return false;
}

bool InputDeckReader::getValue(const char *key, int *value)
{
    char s[VALUE_CHAR_LEN];
    if (!getValue(key, s))
        return false;
    *value = std::atoi(s);
    return true;
}

bool InputDeckReader::getValue(const char *key, double *value)
{
    char s[VALUE_CHAR_LEN];
    if (!getValue(key, s))
        return false;
    *value = std::atof(s);
    return true;
}

bool InputDeckReader::getValue(const char *key, bool *value)
{
    char s[VALUE_CHAR_LEN];
    if (!getValue(key, s))
        return false;
    *value = std::strcmp(s, "true") == 0;
    return true;
}

void InputDeckReader::print()
{
    std::printf("--- Input Deck ---\n");
    for (const KeyValue &kv : keyValueList)
        std::printf("%s = %s\n", kv.key, kv.value);
}

// ---------------------------------------------------------------- C helpers
extern "C" int c3d_host_deck_value(const char *path, const char *key, char *value)
{
    InputDeckReader r;
    r.readInputDeck(path);
    return r.getValue(key, value) ? 1 : 0;
}

extern "C" int c3d_host_optimize_cells(int64_t n, int nm, int nm2, int nq, double *u,
                                       double *alpha, int *totalIter, int *regIndex)
{
    // what solveBatchOptimization (update.cpp:105-178) does for the cells of a grid
    if (!Optimization::c_stack1) {
        Optimization::c_stack1 = new std::stack<Optimization *>;
        Optimization::c_stack2 = new std::stack<Optimization *>;
        Optimization::c_stack3 = new std::stack<Optimization *>;
    }
    for (int64_t c = 0; c < n; ++c)
        Optimization::c_stack1->push(new Optimization((int)c, 0, 0, nm, &u[c * nm], &alpha[c * nm]));
    Optimization::opt(nm, nm2, nq);
    int64_t done = 0;
    while (!Optimization::c_stack3->empty()) {
        Optimization *o = Optimization::c_stack3->top();
        Optimization::c_stack3->pop();
        const int64_t c = o->c_i;
        std::memcpy(&alpha[c * nm], o->c_alpha, nm * sizeof(double));
        std::memcpy(&u[c * nm], o->c_u, nm * sizeof(double));
        if (totalIter)
            totalIter[c] = o->c_totalIter;
        if (regIndex)
            regIndex[c] = o->c_regIndex;
        std::free(o->c_alpha);
        std::free(o->c_u);
        delete o;
        ++done;
    }
    return done == n ? C3D_OK : C3D_ERR_ARG;
}

// Copy a resident grid back into the host arrays it was uploaded from; caller holds the
// mutex.
static void resident_to_host(HostState &S)
{
    if (!S.resident || !S.devValid || !S.devAhead)
        return;
    d2h((void *)S.hostU, S.ru.p, S.gridBytes);
    if (S.hostA)
        d2h((void *)S.hostA, S.ralpha.p, S.gridBytes);
    S.devAhead = false;
}

extern "C" int c3d_host_set_resident(int on)
{
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    resident_to_host(S);
    S.resident = on != 0;
    S.devValid = false;  // the next update() uploads the host arrays
    S.devAhead = false;
    return 0;
}

extern "C" int c3d_host_sync_to_host(void)
{
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    resident_to_host(S);
    return 0;
}

extern "C" int c3d_host_set_precision(int mode)
{
    HostState &S = state();
    std::lock_guard<std::mutex> lk(S.mtx);
    if (mode != C3D_PREC_FP64 && mode != C3D_PREC_FP64_REFORDER && mode != C3D_PREC_MIXED)
        return C3D_ERR_ARG;
    S.precision = mode;
    if (S.ctx)
        return c3d_set_precision(S.ctx, mode);
    return C3D_OK;
}
