// c3d_internal.h -- context layout and kernel launchers shared by the HIP sources of
// libc3d_hip.so.  Not part of the public ABI (include/c3d_hip.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/c3d_hip.h"

namespace c3d {

constexpr int kWave = 64;
constexpr int kChunk = 64;  // ordinates per LDS chunk in the closure sweeps

// The mixed closure's pair image (orders with nm <= kPairMaxNm): the basis products
// P_qp = kPairScale m_qi m_qj of the pairs p = i (i + 1) / 2 + j, j <= i (the packed lower
// triangle), as FP16 halves hi = fp16(fp32 P), lo = fp16(fp32 P - hi), in the B-operand order
// of V_MFMA_F32_16X16X32_F16: per chunk of kChunk ordinates, per block kb of 32 ordinates,
// per tile t of 16 pairs and plane (hi, lo), lane 16 g + n holds 8 consecutive halves of pair
// 16 t + n, element e at ordinate 32 kb + 16 (e >> 2) + g + 4 (e & 3) (the K order in which
// the column sweep's lanes hold their weights).  Built on the host (capi.hip), read by the
// mixed sweeps (closure_fp64.hip).
constexpr double kPairScale = 64.0;  // lifts the small products out of FP16's subnormal range
constexpr int kPairMaxNm = 25;
__host__ __device__ constexpr int pair_count(int nm) { return nm * (nm + 1) / 2; }
__host__ __device__ constexpr int pair_tiles(int nm) { return (pair_count(nm) + 15) / 16; }
__host__ __device__ constexpr int pair_chunk_halves(int nm)
{
    return (kChunk / 32) * pair_tiles(nm) * 2 * 512;
}
// element (ordinate qc of its chunk, pair p, plane) within the chunk's image
__host__ __device__ constexpr int pair_offset(int nm, int qc, int p, int plane)
{
    return ((((qc >> 5) * pair_tiles(nm) + (p >> 4)) * 2 + plane) * 64 + (qc & 3) * 16 +
            (p & 15)) * 8 +
           4 * ((qc & 31) >> 4) + ((qc >> 2) & 3);
}

// The column sweep's row image (nm <= kPairMaxNm): per chunk the kChunk rows of M, leading
// dimension nm | 1 (the pad column zero), then the chunk's kChunk weights, in FP64 -- the LDS
// layout of the sweeps, copied whole by LDS DMA ((nm | 1) + 1) / 2 KB a chunk.
__host__ __device__ constexpr int row_chunk_doubles(int nm) { return ((nm | 1) + 1) * kChunk; }

// Host-side tables (init.cpp:423-548, 287-337, 199-234), built by tables.cpp or copied from
// the caller.
struct HostTables {
    int N = 0, Q = 0, nm = 0, nm2 = 0, nq = 0, nnz = 0;
    std::vector<double> w, xi, eta, mu, M, M2, beta, betaVal, pnOps;
    std::vector<int32_t> betaColInd, betaRowPtr;
};

void build_tables(int N, int Q, HostTables &t);  // tables.cpp
void build_pn_operators(HostTables &t);          // tables.cpp (needs w, xi, eta, mu, M)

// Constants the closure needs, computed on the host once (libm), so that device and CPU see
// the same bits.
struct ClosureConsts {
    double m0;         // M[0][0] = 1/(2 sqrt(pi)) up to rounding (fobj.cpp:123 relies on it)
    double isoAlpha0;  // 2 sqrt(pi) log(0.5/sqrt(pi)), optBatch.cpp:117
    double minf;       // sqrt((2p+1)/(2 pi)), p = floor(sqrt(nm))+1, optBatch.cpp:136-137
};

// Launch shapes of the closure kernels that a context may choose among (c3d_tune).
struct Tune {
    int mixedWaves = 3;  // C3D_LAUNCH_MIXED_WAVES, kept for the ABI: no sweep reads it
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// The warm-start network of a context (warmstart.hip)
struct Mlp {
    bool have = false;
    float normHost[4 * C3D_MLP_IN] = {};  // x mean, x std, alpha mean, alpha std
    DevBuf params;                        // FP32 master parameters
    DevBuf opt;                           // gradient, Adam moments
    DevBuf data;                          // training features/targets and minibatch
    DevBuf work;                          // activations
    DevBuf w16, bias, normDev;            // inference images
};

}  // namespace c3d

struct c3d_ctx_s {
    int device = 0;
    int numCU = 0;
    c3d::HostTables t;
    c3d::ClosureConsts cc{};
    int nqPad = 0;  // nq rounded up to kChunk
    // device tables
    double *d_w = nullptr, *d_xi = nullptr, *d_eta = nullptr, *d_mu = nullptr;
    double *d_M = nullptr;     // [nqPad][nm], rows >= nq zero
    double *d_MT = nullptr;    // [nm][nqPad], the transpose of d_M
    double *d_wpad = nullptr;  // [nqPad], entries >= nq zero
    uint16_t *d_P16 = nullptr;  // the pair image (nm <= kPairMaxNm), nqPad / kChunk chunks
    double *d_R = nullptr;      // the row image (nm <= kPairMaxNm), nqPad / kChunk chunks
    double *d_M2 = nullptr;    // [nqPad][nm2]
    double *d_beta = nullptr, *d_betaVal = nullptr, *d_pnOps = nullptr;
    int32_t *d_betaColInd = nullptr, *d_betaRowPtr = nullptr;
    // control words
    int *d_status = nullptr;                  // OR of per-cell error flags since c3d_sync
    // scratch owned by the context, grown on demand
    c3d::DevBuf ansatz, uOld, flux;
    c3d::DevBuf closureState;  // per-cell Newton state, Hessians, worklists
    c3d::DevBuf statsPart;     // c3d_stats_summary's per-block partial sums
    c3d::Mlp mlp;
    bool warm = false;         // closure starts from the network's proposal where it is better
    c3d::DevBuf warmBuf;       // features and proposals of the closure's cells
    int maxPasses = 100000;     // sweep passes per closure call before cells are reported
    int precision = C3D_PREC_FP64;
    double mixedCondMax = 1e8;  // C3D_PREC_MIXED hand-over to the FP64 path
    int mixedMaxIter = 40;
    double mixedSwitch = 5e-4;
    double mixedFinishCond = 1e5;  // hand-over to FP64 Hessians at ||g|| <= mixedSwitch above this
    // profiling of the last closure call
    bool profiling = false;
    std::vector<hipEvent_t> evPool;
    std::vector<std::pair<int, int>> evSpans;  // (0: sweep+step pass, 4 events | 1: tail, 2)
    int evUsed = 0;
    int lastPasses = 0;
    int lastTailLaunches = 0;
    hipStream_t lastStream = nullptr;
    unsigned long long *d_sweepCount = nullptr;  // [0] bulk sweep kernel, [1] tail kernel
    // fused tail: taken once at most tailCells cells remain listed, tailPasses per launch
    int tailCells = 8192;
    int tailPasses = 32;
    int tailWideCells = 512;  // four waves per cell (speculative line-search trials) below this
    // > 0: the asynchronous schedule with this many device-gated bulk passes (no host sync)
    int asyncPasses = 0;
    c3d::Tune tune;
    int gauntDirect[2] = {-1, -1};  // Gaunt form (dense, CSR) equals the direct form: 1, 0, -1 unknown
    std::mutex mtx;
};

namespace c3d {

hipError_t grow(DevBuf &b, size_t bytes);

// closure_fp64.hip
hipError_t launch_closure_fp64(c3d_ctx_s *ctx, int64_t ncells, double *u, double *alpha,
                               const c3d_closure_opts &o, c3d_cell_stats *stats,
                               hipStream_t s, const DevBuf *ext = nullptr);
// the device address of c3d_closure_noalloc's scratch array, looked up once per device
hipError_t closure_scratch_setup(int device);
hipError_t launch_dense_solve(int nm, int64_t n, const double *H, const double *g, double *d,
                              double *rcond, int *info, double *L, int lapackOrder, hipStream_t s);
hipError_t launch_hessian_fp64(c3d_ctx_s *ctx, int64_t ncells, const double *alpha,
                               double *hs, hipStream_t s);

// warmstart.hip
hipError_t mlp_pack(c3d_ctx_s *ctx, hipStream_t s);
hipError_t mlp_infer(c3d_ctx_s *ctx, int64_t n, const float *X, double *out, hipStream_t s);
hipError_t mlp_features(c3d_ctx_s *ctx, int64_t n, const double *u, const double *alpha,
                        const float *normDev, float *X, float *T, hipStream_t s);
hipError_t mlp_forward(c3d_ctx_s *ctx, int64_t n, const float *p, const float *X, float *Y,
                       hipStream_t s);
hipError_t mlp_loss_grad(c3d_ctx_s *ctx, int64_t n, const float *p, const float *X,
                         const float *T, float *grad, float *loss, hipStream_t s);
int mlp_train(c3d_ctx_s *ctx, int64_t n, const double *u, const double *alpha, int mseSteps,
              int dualSteps, int batch, float lr, uint64_t seed, float *lossFirst,
              float *lossLast, hipStream_t s);
hipError_t mlp_dual_loss_grad(c3d_ctx_s *ctx, int64_t n, const float *p, const float *X,
                              const double *u, const float *normDev, float *grad, float *loss,
                              hipStream_t s);
int mlp_set(c3d_ctx_s *ctx, const float *params, const float *norm);
int mlp_get(c3d_ctx_s *ctx, float *params, float *norm);

// grid.hip
hipError_t launch_ghost_fill(int nx, int ny, int nz, int nm, double *u, hipStream_t s);
hipError_t launch_ansatz(c3d_ctx_s *ctx, int nx, int ny, int nz, const double *alpha,
                         double *A, int isPn, hipStream_t s);
hipError_t launch_stencil(c3d_ctx_s *ctx, int nx, int ny, int nz, const double *A,
                          double *flux, double *u, const double *sigmaT,
                          const double *sigmaS, double dt, double dx, double dy, double dz,
                          double theta, int fuseEuler, int zeroGhosts, hipStream_t s);
hipError_t launch_flux_moment(c3d_ctx_s *ctx, int nx, int ny, int nz, const double *u,
                              double *flux, double dx, double dy, double dz, hipStream_t s);
hipError_t launch_euler(int nx, int ny, int nz, int nm, double *u, const double *flux,
                        const double *sigmaT, const double *sigmaS, double dt, hipStream_t s);
hipError_t launch_rk2(int nx, int ny, int nz, int nm, double *u, double *uOld, int kind,
                      hipStream_t s);
hipError_t launch_init_grid(int nx, int ny, int nz, int nm, const double *a, const double *d,
                            int ic, double sigma, double floorv, double noise, uint64_t seed,
                            double cross, double *u, double *alpha, double *sigmaT,
                            double *sigmaS, hipStream_t s);
hipError_t launch_make_batch(c3d_ctx_s *ctx, int64_t n, uint64_t seed, double smin,
                             double smax, double *u, double *alpha0, double *alphaStar,
                             hipStream_t s);
hipError_t launch_stats_partials(int64_t n, const c3d_cell_stats *st, long long *part,
                                 int blocks, hipStream_t s);

}  // namespace c3d
