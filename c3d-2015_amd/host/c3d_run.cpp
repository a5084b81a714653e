// c3d_run -- the reference's program on one GPU: the input deck, init(), the time loop of
// main.cpp and the output files of output.cpp, over the host-level drop-in (update(dt) of
// libc3d_host.so, which runs each SSP-RK2 step on the device).
//
//   c3d_run [--precision fp64|mixed|reforder] [--staged] [deck]   (deck: input.deck)
//
// Follows src/main.cpp:31-112 (loop, prints), src/init.cpp:344-576 with
// setupLocalMpiInfo's single-domain branch (89-125), initPn (194-238) and initMn
// (244-337), and src/output.cpp:19-222 (file names and binary layouts).  The quadrature,
// basis and Gaunt tables are the library's own (c3d_tables_host: Gauss-Legendre nodes and
// real spherical harmonics as init.cpp:423-548 builds them, without GSL).  The initial
// state is computed on the host in init.cpp's arithmetic.
//
// Differences from the reference program, each deliberate:
//   * one domain: MPIX/MPIY/MPIZ are read (the deck must hold them) and ignored;
//   * OMP_THREADS and the CUDA_* keys are read and ignored (there is no CPU path);
//   * the output/ directory is created when missing (the reference aborts);
//   * TFINAL <= 0 or OUTDT <= 0 is an input error (the reference's loop never ends);
//   * --precision selects the closure's arithmetic (default: FP64 on the matrix cores;
//     reforder: the reference's summation order, bit-identical to the oracle);
//   * the grid stays on the device between steps (the host drop-in's resident mode) and is
//     copied back before each output; the run statistics are then folded per stage from
//     device-side sums, so the two means in the .pn tail can differ from a per-cell fold in
//     the last bits.  --staged copies the grid through host memory every step instead.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/c3d_hip.h"
#include "../../include/c3d_host.h"

namespace {

constexpr int kGhost = 2;
constexpr int kNumReg = 5;  // NUM_REGULARIZATIONS, opt.h:37

struct Run {
    // deck
    std::string solver, initCond;
    double tFinal = 0, outDt = 0, cfl = 0, cross = 0, initFloor = 0, sigma = 0;
    bool outputData = false, outputTimings = false;
    int momentOrder = 0, quadOrder = 0;
    int globalN[3] = {0, 0, 0};
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    // derived
    double maxDt = 0;
    int maxNumUpdates = 0;
    // storage behind the globals
    std::vector<double> w, xi, eta, mu, M, M2, beta, betaVal, u, alpha, sigT, sigS, times;
    std::vector<int32_t> betaColInd, betaRowPtr;
    std::vector<OPTIMIZATION_STATS> stats;
    int hist[kNumReg] = {0, 0, 0, 0, 0};
};

[[noreturn]] void fail(const char *what, const char *detail)
{
    std::printf("c3d_run: %s%s%s\n", what, detail ? " " : "", detail ? detail : "");
    std::fflush(stdout);
    std::exit(1);
}

template <class T>
void need(InputDeckReader &r, const char *key, T *v)
{
    if (!r.getValue(key, v))
        fail("input error: missing or malformed key", key);
}

void need_str(InputDeckReader &r, const char *key, std::string &out)
{
    char v[InputDeckReader::VALUE_CHAR_LEN];
    if (!r.getValue(key, v))
        fail("input error: missing or malformed key", key);
    out = v;
}

// init.cpp:27-34
double gaussian(double s, double x, double y, double z)
{
    const double s2 = s * s;
    const double f1 = 1.0 / (s * std::sqrt(2.0 * M_PI));
    return f1 * f1 * f1 * std::exp(-(x * x + y * y + z * z) / (2.0 * s2));
}
// init.cpp:36-40
double smooth(double x, double y, double z)
{
    return 2.0 + std::cos(2 * M_PI * x) * std::cos(2 * M_PI * y) * std::cos(2 * M_PI * z);
}
// init.cpp:42-66: 125 Gaussians on the odd tenths of the box, plus the smooth state
double funky(const Run &R, double x, double y, double z)
{
    const double w[3] = {R.hi[0] - R.lo[0], R.hi[1] - R.lo[1], R.hi[2] - R.lo[2]};
    double s = 0.0;
    for (int i = 1; i < 10; i += 2)
        for (int j = 1; j < 10; j += 2)
            for (int k = 1; k < 10; k += 2)
                s += gaussian(R.sigma, x - (R.lo[0] + i * w[0] / 10.0),
                              y - (R.lo[1] + j * w[1] / 10.0), z - (R.lo[2] + k * w[2] / 10.0));
    return smooth(x, y, z) + s;
}

void read_deck(Run &R, InputDeckReader &r)
{
    // init.cpp:363-378, 89-100 (setupLocalMpiInfo), 196 (initPn), 250-256 (initMn)
    need_str(r, "SOLVER", R.solver);
    need(r, "TFINAL", &R.tFinal);
    need(r, "OUTDT", &R.outDt);
    need(r, "OUTPUT_DATA", &R.outputData);
    need(r, "OUTPUT_OPT_STATS", &g_outputOptStats);
    need(r, "OUTPUT_TIMINGS", &R.outputTimings);
    need(r, "MOMENT_ORDER", &R.momentOrder);
    need(r, "QUAD_ORDER", &R.quadOrder);
    need(r, "CFL_FACTOR", &R.cfl);
    need(r, "CROSS_SECTION", &R.cross);
    need(r, "INIT_FLOOR", &R.initFloor);
    need(r, "GAUSS_SIGMA", &R.sigma);
    need_str(r, "INIT_COND", R.initCond);
    int ignoredInt;
    bool ignoredBool;
    need(r, "OMP_THREADS", &ignoredInt);
    need(r, "CUDA_KINETIC", &ignoredBool);
    need(r, "CUDA_BATCHED", &ignoredBool);
    need(r, "CUDA_BATCH_SIZE", &ignoredInt);
    need(r, "THETA", &g_theta);
    const char *nk[3] = {"NX", "NY", "NZ"}, *mk[3] = {"MPIX", "MPIY", "MPIZ"};
    const char *ak[3] = {"AX", "AY", "AZ"}, *bk[3] = {"BX", "BY", "BZ"};
    for (int d = 0; d < 3; ++d) {
        need(r, nk[d], &R.globalN[d]);
        need(r, mk[d], &ignoredInt);
    }
    for (int d = 0; d < 3; ++d)
        need(r, ak[d], &R.lo[d]);
    for (int d = 0; d < 3; ++d)
        need(r, bk[d], &R.hi[d]);
    if (R.solver == "pn")
        g_solver = SOLVER_PN;
    else if (R.solver == "mn")
        g_solver = SOLVER_MN;
    else
        fail("solver type does not exist:", R.solver.c_str());
    if (R.initCond != "gaussian" && R.initCond != "smooth" && R.initCond != "funky")
        fail("initial condition not supported:", R.initCond.c_str());
    if (g_solver == SOLVER_PN) {
        int k;
        need(r, "PN_USE_KINETIC_FLUX", &k);
        g_pnUseKineticFlux = k;
    } else {
        need(r, "MN_TOL", &g_mnTol);
        need(r, "MN_COND_H_MAX", &g_mnCondHMax);
        need(r, "MN_MAX_ITER", &g_mnMaxIter);
        need(r, "MN_BATCHED", &g_mnBatched);
        need(r, "MN_TOL_GAMMA", &g_mnTolGamma);
        need(r, "MN_USE_GAUNT_COEF", &g_mnUseGaunt);
        need(r, "MN_USE_GAUNT_SPARSE", &g_mnUseGauntSparse);
    }
    if (!(R.tFinal > 0.0) || !(R.outDt > 0.0))
        fail("TFINAL and OUTDT must be positive", nullptr);
    if (R.momentOrder < 1 || R.momentOrder > 4)
        fail("MOMENT_ORDER must be 1..4 (the orders the closure kernels are built for)", nullptr);
    if (R.quadOrder < 1)
        fail("QUAD_ORDER must be positive", nullptr);
    for (int d = 0; d < 3; ++d)
        if (R.globalN[d] < 2 || !(R.hi[d] > R.lo[d]))
            fail("the grid needs at least 2 cells per axis and AX < BX (likewise y, z)",
                 nullptr);
}

void init(Run &R)
{
    const int N = R.momentOrder, Q = R.quadOrder;
    g_numQuadPoints = 2 * Q * Q;
    g_numMoments = (N + 1) * (N + 1);
    g_numMoments2 = (2 * N + 1) * (2 * N + 1);
    const int nm = g_numMoments, nm2 = g_numMoments2, nq = g_numQuadPoints;
    // one domain (init.cpp:104-124); ghost index ranges (init.cpp:443-454)
    g_nx = R.globalN[0];
    g_ny = R.globalN[1];
    g_nz = R.globalN[2];
    g_dx = (R.hi[0] - R.lo[0]) / g_nx;
    g_dy = (R.hi[1] - R.lo[1]) / g_ny;
    g_dz = (R.hi[2] - R.lo[2]) / g_nz;
    int *gs[3] = {g_gX, g_gY, g_gZ};
    const int n[3] = {g_nx, g_ny, g_nz};
    for (int d = 0; d < 3; ++d) {
        gs[d][0] = 0;
        gs[d][1] = kGhost;
        gs[d][2] = kGhost + n[d] - 1;
        gs[d][3] = 2 * kGhost + n[d] - 1;
    }
    // tables
    R.w.resize(nq);
    R.xi.resize(nq);
    R.eta.resize(nq);
    R.mu.resize(nq);
    R.M.resize((size_t)nq * nm);
    R.M2.resize((size_t)nq * nm2);
    R.beta.resize((size_t)nm * nm * nm2);
    int nnz = 0;
    if (c3d_tables_host(N, Q, R.w.data(), R.xi.data(), R.eta.data(), R.mu.data(), R.M.data(),
                        R.M2.data(), R.beta.data(), nullptr, &nnz, 0, nullptr, nullptr,
                        nullptr) != C3D_OK)
        fail("could not build the quadrature tables", nullptr);
    R.betaVal.resize(nnz > 0 ? nnz : 1);
    R.betaColInd.resize(nnz > 0 ? nnz : 1);
    R.betaRowPtr.resize((size_t)nm * nm + 1);
    if (c3d_tables_host(N, Q, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                        nullptr, &nnz, (int)R.betaVal.size(), R.betaVal.data(),
                        R.betaColInd.data(), R.betaRowPtr.data()) != C3D_OK)
        fail("could not build the Gaunt tables", nullptr);
    g_w = R.w.data();
    g_xi = R.xi.data();
    g_eta = R.eta.data();
    g_mu = R.mu.data();
    g_m = R.M.data();
    g_m2 = R.M2.data();
    g_beta = R.beta.data();
    g_betaVal = R.betaVal.data();
    g_betaColInd = R.betaColInd.data();
    g_betaRowPtr = R.betaRowPtr.data();
    // initial condition over every cell, ghosts included (init.cpp:479-506)
    const int NX = g_gX[3] + 1, NY = g_gY[3] + 1, NZ = g_gZ[3] + 1;
    const size_t ng = (size_t)NX * NY * NZ;
    R.u.assign(ng * nm, 0.0);
    R.sigT.assign(ng, R.cross);
    R.sigS.assign(ng, R.cross);
    for (int i = 0; i < NX; ++i)
        for (int j = 0; j < NY; ++j)
            for (int k = 0; k < NZ; ++k) {
                const double x = R.lo[0] + (i - g_gX[1]) * g_dx + g_dx / 2.0;
                const double y = R.lo[1] + (j - g_gY[1]) * g_dy + g_dy / 2.0;
                const double z = R.lo[2] + (k - g_gZ[1]) * g_dz + g_dz / 2.0;
                double *uc = &R.u[(((size_t)i * NY + j) * NZ + k) * nm];
                if (R.initCond == "gaussian") {
                    uc[0] = std::fmax(R.initFloor, gaussian(R.sigma, x, y, z) * 2.0 *
                                                       std::sqrt(M_PI));
                } else if (R.initCond == "smooth") {
                    uc[0] = smooth(x, y, z) * 2.0 * std::sqrt(M_PI);
                    uc[1] = 0.3 * uc[0];
                    uc[2] = 0.3 * uc[0];
                    uc[3] = 0.3 * uc[0];
                } else {
                    uc[0] = funky(R, x, y, z) * 2.0 * std::sqrt(M_PI);
                }
            }
    g_u = R.u.data();
    g_sigmaT = R.sigT.data();
    g_sigmaS = R.sigS.data();
    const double minD = std::fmin(std::fmin(g_dx, g_dy), g_dz);
    if (g_solver == SOLVER_PN) {
        R.maxDt = R.cfl * 0.5 * minD;  // init.cpp:237
    } else {
        // init.cpp:262-268, 274, 277-284
        R.alpha.assign(ng * nm, 0.0);
        for (size_t c = 0; c < ng; ++c)
            R.alpha[c * nm] = std::sqrt(4.0 * M_PI) * std::log(R.u[c * nm] / std::sqrt(4.0 * M_PI));
        g_mnAlpha = R.alpha.data();
        R.maxDt = R.cfl * (2.0 / (2.0 + g_theta)) / (1.0 / g_dx + 1.0 / g_dy + 1.0 / g_dz);
        g_mnMaxIterHist = 0;
        g_mnMaxGammaIter = 0;
        g_mnIterMean = 0.0;
        g_mnIterGammaMean = 0.0;
        g_mnNumDualSolves = 0;
    }
    g_mnHistReg = R.hist;
    // init.cpp:558-569
    g_numUpdates = 0;
    R.maxNumUpdates = (int)std::ceil(R.tFinal / R.maxDt) + 1;
    R.times.assign((size_t)8 * R.maxNumUpdates, 0.0);
    g_updateTimes = R.times.data();
    if (g_outputOptStats) {
        R.stats.assign((size_t)2 * R.maxNumUpdates * g_nx, OPTIMIZATION_STATS{});
        g_optStats = R.stats.data();
    }
}

FILE *open_out(const char *name, const char *mode)
{
    char path[512];
    std::snprintf(path, sizeof(path), "output/%s", name);
    FILE *f = std::fopen(path, mode);
    if (!f)
        fail("could not open", path);
    return f;
}

void put_i64(FILE *f, int64_t v) { std::fwrite(&v, sizeof(v), 1, f); }
void put_f64(FILE *f, double v) { std::fwrite(&v, sizeof(v), 1, f); }

// output.cpp:19-115
void output_data(const Run &R, double time)
{
    if (!R.outputData)
        return;
    c3d_host_sync_to_host();
    char name[256];
    std::snprintf(name, sizeof(name), "out_%.6f.master", time);
    FILE *f = open_out(name, "w");
    std::fprintf(f, "%d\n%d\n%d\n", R.globalN[0], R.globalN[1], R.globalN[2]);
    std::fprintf(f, "%f\n%f\n%f\n", R.lo[0], R.lo[1], R.lo[2]);
    std::fprintf(f, "%f\n%f\n%f\n", R.hi[0], R.hi[1], R.hi[2]);
    std::fprintf(f, "%d\n", g_numMoments);
    std::fprintf(f, "1\n1\n1\n");
    std::fprintf(f, "out_%.6f_%d.pn\n", time, 0);
    std::fclose(f);

    std::snprintf(name, sizeof(name), "out_%.6f_%d.pn", time, 0);
    f = open_out(name, "wb");
    put_i64(f, g_nx);
    put_i64(f, g_ny);
    put_i64(f, g_nz);
    const int NY = g_gY[3] + 1, NZ = g_gZ[3] + 1;
    for (int i = g_gX[1]; i <= g_gX[2]; i++)
        for (int j = g_gY[1]; j <= g_gY[2]; j++)
            for (int k = g_gZ[1]; k <= g_gZ[2]; k++)
                std::fwrite(&g_u[(((size_t)i * NY + j) * NZ + k) * g_numMoments],
                            sizeof(double), g_numMoments, f);
    put_i64(f, g_mnMaxIterHist);
    put_i64(f, g_mnMaxGammaIter);
    put_f64(f, g_mnIterMean);
    put_f64(f, g_mnIterGammaMean);
    put_i64(f, g_mnNumDualSolves);
    for (int r = 0; r < kNumReg; ++r)
        put_i64(f, g_mnHistReg[r]);
    std::fclose(f);
}

// output.cpp:122-180
void output_timings(const Run &R)
{
    if (!R.outputTimings)
        return;
    FILE *f = open_out("times.master", "w");
    std::fprintf(f, "%d\n1\n1\n1\ntimes_%d.dat\n", g_numUpdates, 0);
    std::fclose(f);
    f = open_out("times_0.dat", "wb");
    std::fwrite(g_updateTimes, sizeof(double), (size_t)g_numUpdates * 8, f);
    std::fclose(f);
}

// output.cpp:186-222
void output_opt_stats()
{
    if (!g_outputOptStats)
        return;
    FILE *f = open_out("stats_0.dat", "wb");
    put_i64(f, g_numUpdates);
    put_i64(f, g_nx);
    std::fwrite(g_optStats, sizeof(OPTIMIZATION_STATS), (size_t)g_numUpdates * 2 * g_nx, f);
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv)
{
    const char *deckPath = "input.deck";
    int precision = C3D_PREC_FP64;
    bool staged = false;
    for (int a = 1; a < argc; ++a) {
        if (std::strcmp(argv[a], "--precision") == 0 && a + 1 < argc) {
            const char *p = argv[++a];
            if (std::strcmp(p, "fp64") == 0)
                precision = C3D_PREC_FP64;
            else if (std::strcmp(p, "mixed") == 0)
                precision = C3D_PREC_MIXED;
            else if (std::strcmp(p, "reforder") == 0)
                precision = C3D_PREC_FP64_REFORDER;
            else
                fail("unknown precision", p);
        } else if (std::strcmp(argv[a], "--staged") == 0) {
            staged = true;
        } else if (argv[a][0] == '-') {
            std::printf("usage: c3d_run [--precision fp64|mixed|reforder] [--staged] [deck]\n");
            return 2;
        } else {
            deckPath = argv[a];
        }
    }
    struct stat sb;
    if (stat(deckPath, &sb) != 0)
        fail("no input deck at", deckPath);

    Run R;
    InputDeckReader deck;
    deck.readInputDeck(deckPath);
    read_deck(R, deck);
    deck.print();
    init(R);
    if (stat("output", &sb) != 0 && mkdir("output", 0755) != 0)
        fail("could not create", "output/");
    c3d_host_set_precision(precision);
    c3d_host_set_resident(staged ? 0 : 1);
    std::printf("Init done.\n");

    // main.cpp:50-76
    const auto t0 = std::chrono::steady_clock::now();
    double t = 0.0, dt = R.maxDt;
    output_data(R, t);
    for (double tOut = std::fmin(t + R.outDt, R.tFinal); tOut <= R.tFinal;
         tOut = std::fmin(tOut + R.outDt, R.tFinal)) {
        dt = R.maxDt;
        while (t < tOut - 1e-14 * tOut) {
            t += dt;
            if (t > tOut) {
                dt = tOut - (t - dt);
                t = tOut;
            }
            if (g_numUpdates >= R.maxNumUpdates)
                fail("more steps than init() sized the timing arrays for", nullptr);
            update(dt);
            std::printf("dt = %f   t = %f\n", dt, t);
        }
        output_data(R, t);
        if (tOut > R.tFinal - 1e-14 * R.tFinal)
            break;
    }

    // main.cpp:84-104
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Time taken: %.3lfs\n", secs);
    if (g_solver == SOLVER_MN) {
        std::printf("g_mnMaxIterHist = %d\n", g_mnMaxIterHist);
        std::printf("g_mnMaxGammaIter = %d\n", g_mnMaxGammaIter);
        std::printf("g_mnIterMean = %f\n", g_mnIterMean);
        std::printf("g_mnIterGammaMean = %f\n", g_mnIterGammaMean);
        std::printf("g_mnNumDualSolves = %d\n", g_mnNumDualSolves);
        for (int i = 0; i < kNumReg; i++)
            std::printf("  g_mnHistReg[%d] = %d\n", i, g_mnHistReg[i]);
    }
    output_timings(R);
    output_opt_stats();
    c3d_host_set_resident(0);
    return 0;
}
