// capi.hip -- the extern "C" surface of libc3d_hip.so (declared in include/c3d_hip.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "c3d_internal.h"

namespace c3d {

hipError_t grow(DevBuf &b, size_t bytes)
{
    if (b.bytes >= bytes)
        return hipSuccess;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        if (e != hipSuccess)
            return e;
        b.p = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e == hipSuccess)
        b.bytes = bytes;
    return e;
}

namespace {

template <class T>
hipError_t upload(T **dst, const T *src, size_t n)
{
    hipError_t e = hipMalloc((void **)dst, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess || n == 0)
        return e;
    return hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
}

// The mixed closure's pair image of the tables (layout and scale in c3d_internal.h); ordinates
// >= nq and pairs >= nm (nm + 1) / 2 are zero.
void pair_image(const HostTables &t, int nqPad, std::vector<uint16_t> &img)
{
    const int nm = t.nm, ch = pair_chunk_halves(nm);
    img.assign((size_t)nqPad / kChunk * ch, 0);
    for (int q = 0; q < t.nq; ++q) {
        const double *m = &t.M[(size_t)q * nm];
        uint16_t *dst = img.data() + (size_t)(q / kChunk) * ch;
        for (int i = 0; i < nm; ++i)
            for (int j = 0; j <= i; ++j) {
                const int p = i * (i + 1) / 2 + j;
                const float v = (float)(kPairScale * (m[i] * m[j]));
                const _Float16 hi = (_Float16)v;
                const _Float16 lo = (_Float16)(v - (float)hi);
                dst[pair_offset(nm, q % kChunk, p, 0)] = __builtin_bit_cast(uint16_t, hi);
                dst[pair_offset(nm, q % kChunk, p, 1)] = __builtin_bit_cast(uint16_t, lo);
            }
    }
}

int finish_ctx(c3d_ctx_s *ctx)
{
    HostTables &t = ctx->t;
    if (hipGetDevice(&ctx->device) != hipSuccess)
        return C3D_ERR_HIP;
    if (hipDeviceGetAttribute(&ctx->numCU, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
        hipSuccess)
        return C3D_ERR_HIP;
    if (closure_scratch_setup(ctx->device) != hipSuccess)
        return C3D_ERR_HIP;
    ctx->cc.m0 = t.M[0];
    ctx->cc.isoAlpha0 = 2.0 * sqrt(M_PI) * log(0.5 / sqrt(M_PI));
    {
        int pnOrder = (int)floor(sqrt((double)t.nm)) + 1;
        ctx->cc.minf = sqrt((2 * pnOrder + 1) / (2 * M_PI));
    }
    ctx->nqPad = (t.nq + kChunk - 1) / kChunk * kChunk;
    std::vector<double> Mp((size_t)ctx->nqPad * t.nm, 0.0), wp(ctx->nqPad, 0.0);
    std::memcpy(Mp.data(), t.M.data(), t.M.size() * sizeof(double));
    std::memcpy(wp.data(), t.w.data(), t.w.size() * sizeof(double));
    hipError_t e = hipSuccess;
#define UP(dst, vec)                                                                           \
    if (e == hipSuccess)                                                                       \
    e = upload(&ctx->dst, vec.data(), vec.size())
    UP(d_w, t.w);
    UP(d_xi, t.xi);
    UP(d_eta, t.eta);
    UP(d_mu, t.mu);
    UP(d_M, Mp);
    {
        std::vector<double> MTp((size_t)ctx->nqPad * t.nm, 0.0);
        for (int q = 0; q < t.nq; ++q)
            for (int i = 0; i < t.nm; ++i)
                MTp[(size_t)i * ctx->nqPad + q] = t.M[(size_t)q * t.nm + i];
        UP(d_MT, MTp);
    }
    UP(d_wpad, wp);
    if (t.nm <= kPairMaxNm) {
        std::vector<uint16_t> img;
        pair_image(t, ctx->nqPad, img);
        UP(d_P16, img);
        const int nms = t.nm | 1, rc = row_chunk_doubles(t.nm);
        std::vector<double> rows((size_t)ctx->nqPad / kChunk * rc, 0.0);
        for (int q = 0; q < t.nq; ++q) {
            double *dst = rows.data() + (size_t)(q / kChunk) * rc;
            for (int i = 0; i < t.nm; ++i)
                dst[(q % kChunk) * nms + i] = t.M[(size_t)q * t.nm + i];
            dst[kChunk * nms + q % kChunk] = t.w[q];
        }
        UP(d_R, rows);
    }
    if (!t.M2.empty()) {
        std::vector<double> M2p((size_t)ctx->nqPad * t.nm2, 0.0);
        std::memcpy(M2p.data(), t.M2.data(), t.M2.size() * sizeof(double));
        UP(d_M2, M2p);
    }
    UP(d_beta, t.beta);
    UP(d_betaVal, t.betaVal);
    UP(d_betaColInd, t.betaColInd);
    UP(d_betaRowPtr, t.betaRowPtr);
    UP(d_pnOps, t.pnOps);
#undef UP
    if (e == hipSuccess)
        e = hipMalloc((void **)&ctx->d_sweepCount, 3 * sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipMalloc((void **)&ctx->d_status, sizeof(int));
    if (e == hipSuccess)
        e = hipMemset(ctx->d_status, 0, sizeof(int));
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    return e == hipSuccess ? C3D_OK : C3D_ERR_HIP;
}

bool valid_order(int N) { return N >= 1 && N <= 4; }

// The Gaunt form of the Hessian (fobj.cpp:59-95) is H_ij = sum_k beta_ijk G_k with
// G_k = sum_q w_q e_q m2_qk.  Where beta reproduces the basis products at every ordinate,
// sum_k beta_ijk m2_qk = m_qi m_qj, that H is sum_q w_q e_q m_qi m_qj, the direct form's,
// up to the identity's residual.  Products of real spherical harmonics of degree <= N are
// combinations of those of degree <= 2N, so the identity holds once the quadrature resolves
// degree 2N, and the device then evaluates the direct form.  Checked on the host once per
// context and form (dense beta, or its CSR without the entries |beta| <= 1e-13); the caller
// holds ctx->mtx.
bool gaunt_matches_direct(c3d_ctx_s *ctx, bool sparse)
{
    int &flag = ctx->gauntDirect[sparse ? 1 : 0];
    if (flag >= 0)
        return flag == 1;
    const HostTables &t = ctx->t;
    const int nm = t.nm, nm2 = t.nm2, nq = t.nq;
    bool ok = nm2 > 0 && t.M2.size() == (size_t)nq * nm2 &&
              (sparse ? t.betaRowPtr.size() == (size_t)nm * nm + 1
                      : t.beta.size() == (size_t)nm * nm * nm2);
    double worst = 0.0, scale = 0.0;
    for (int q = 0; ok && q < nq; ++q) {
        const double *m = &t.M[(size_t)q * nm], *m2 = &t.M2[(size_t)q * nm2];
        for (int i = 0; i < nm; ++i)
            for (int j = 0; j < nm; ++j) {
                const int ij = i * nm + j;
                double s = 0.0;
                if (sparse) {
                    for (int e = t.betaRowPtr[ij]; e < t.betaRowPtr[ij + 1]; ++e) {
                        const int k = t.betaColInd[e];
                        if (k < 0 || k >= nm2) {
                            ok = false;
                            break;
                        }
                        s += t.betaVal[e] * m2[k];
                    }
                } else {
                    for (int k = 0; k < nm2; ++k)
                        s += t.beta[(size_t)ij * nm2 + k] * m2[k];
                }
                const double prod = m[i] * m[j];
                worst = std::fmax(worst, std::fabs(s - prod));
                scale = std::fmax(scale, std::fabs(prod));
            }
    }
    ok = ok && std::isfinite(worst) && worst <= 1e-10 * scale;
    flag = ok ? 1 : 0;
    return ok;
}

// A closure or Hessian request in the Gaunt form: served by the direct form where the two
// agree (above); not in the reference-order mode, whose contract is the reference's bits.
int gaunt_status(c3d_ctx_s *ctx, int useGaunt, int useGauntSparse)
{
    if (!useGaunt)
        return C3D_OK;
    if (ctx->precision == C3D_PREC_FP64_REFORDER)
        return C3D_ERR_UNSUPPORTED;
    return gaunt_matches_direct(ctx, useGauntSparse != 0) ? C3D_OK : C3D_ERR_UNSUPPORTED;
}

int hip_status(hipError_t e) { return e == hipSuccess ? C3D_OK : C3D_ERR_HIP; }

}  // namespace
}  // namespace c3d

using namespace c3d;

extern "C" {

int c3d_create(c3d_ctx *out, int moment_order, int quad_order)
{
    if (!out || !valid_order(moment_order) || quad_order < 1 || quad_order > 512)
        return C3D_ERR_ARG;
    *out = nullptr;
    c3d_ctx_s *ctx = new c3d_ctx_s();
    build_tables(moment_order, quad_order, ctx->t);
    int rc = finish_ctx(ctx);
    if (rc != C3D_OK) {
        c3d_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return C3D_OK;
}

int c3d_create_from_tables(c3d_ctx *out, int moment_order, int quad_order, const double *w,
                           const double *xi, const double *eta, const double *mu,
                           const double *M, const double *M2, const double *beta, int nnz,
                           const double *betaVal, const int32_t *betaColInd,
                           const int32_t *betaRowPtr)
{
    if (!out || !valid_order(moment_order) || quad_order < 1 || !w || !xi || !eta || !mu || !M)
        return C3D_ERR_ARG;
    *out = nullptr;
    c3d_ctx_s *ctx = new c3d_ctx_s();
    HostTables &t = ctx->t;
    t.N = moment_order;
    t.Q = quad_order;
    t.nm = (moment_order + 1) * (moment_order + 1);
    t.nm2 = (2 * moment_order + 1) * (2 * moment_order + 1);
    t.nq = 2 * quad_order * quad_order;
    t.w.assign(w, w + t.nq);
    t.xi.assign(xi, xi + t.nq);
    t.eta.assign(eta, eta + t.nq);
    t.mu.assign(mu, mu + t.nq);
    t.M.assign(M, M + (size_t)t.nq * t.nm);
    if (M2)
        t.M2.assign(M2, M2 + (size_t)t.nq * t.nm2);
    if (beta)
        t.beta.assign(beta, beta + (size_t)t.nm * t.nm * t.nm2);
    if (betaVal && betaColInd && betaRowPtr && nnz > 0) {
        t.nnz = nnz;
        t.betaVal.assign(betaVal, betaVal + nnz);
        t.betaColInd.assign(betaColInd, betaColInd + nnz);
        t.betaRowPtr.assign(betaRowPtr, betaRowPtr + t.nm * t.nm + 1);
    }
    build_pn_operators(t);
    int rc = finish_ctx(ctx);
    if (rc != C3D_OK) {
        c3d_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return C3D_OK;
}

int c3d_destroy(c3d_ctx ctx)
{
    if (!ctx)
        return C3D_OK;
    void *ptrs[] = {ctx->d_w, ctx->d_xi, ctx->d_eta, ctx->d_mu, ctx->d_M, ctx->d_MT,
                    ctx->d_wpad, ctx->d_P16, ctx->d_R, ctx->d_M2, ctx->d_beta, ctx->d_betaVal, ctx->d_pnOps, ctx->d_betaColInd,
                    ctx->d_betaRowPtr, ctx->d_sweepCount, ctx->d_status, ctx->ansatz.p,
                    ctx->uOld.p, ctx->flux.p, ctx->closureState.p};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    for (hipEvent_t e : ctx->evPool)
        (void)hipEventDestroy(e);
    delete ctx;
    return C3D_OK;
}

int c3d_tables_host(int moment_order, int quad_order, double *w, double *xi, double *eta,
                    double *mu, double *M, double *M2, double *beta, double *pnOps, int *nnz,
                    int nnzCap, double *betaVal, int32_t *betaColInd, int32_t *betaRowPtr)
{
    if (!valid_order(moment_order) || quad_order < 1 || quad_order > 512)
        return C3D_ERR_ARG;
    HostTables t;
    build_tables(moment_order, quad_order, t);
    auto cp = [](auto *dst, const auto &v) {
        if (dst && !v.empty())
            std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(w, t.w);
    cp(xi, t.xi);
    cp(eta, t.eta);
    cp(mu, t.mu);
    cp(M, t.M);
    cp(M2, t.M2);
    cp(beta, t.beta);
    cp(pnOps, t.pnOps);
    if (nnz)
        *nnz = t.nnz;
    if (betaRowPtr)
        cp(betaRowPtr, t.betaRowPtr);
    if (betaVal && betaColInd && nnzCap >= t.nnz) {
        cp(betaVal, t.betaVal);
        cp(betaColInd, t.betaColInd);
    }
    return C3D_OK;
}

int c3d_sizes(c3d_ctx ctx, int *nm, int *nm2, int *nq, int *nnz)
{
    if (!ctx)
        return C3D_ERR_ARG;
    if (nm)
        *nm = ctx->t.nm;
    if (nm2)
        *nm2 = ctx->t.nm2;
    if (nq)
        *nq = ctx->t.nq;
    if (nnz)
        *nnz = ctx->t.nnz;
    return C3D_OK;
}

int c3d_get_tables(c3d_ctx ctx, double *w, double *xi, double *eta, double *mu, double *M,
                   double *M2, double *beta, double *betaVal, int32_t *betaColInd,
                   int32_t *betaRowPtr, double *pnOps)
{
    if (!ctx)
        return C3D_ERR_ARG;
    const HostTables &t = ctx->t;
    auto cp = [](auto *dst, const auto &v) {
        if (dst && !v.empty())
            std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(w, t.w);
    cp(xi, t.xi);
    cp(eta, t.eta);
    cp(mu, t.mu);
    cp(M, t.M);
    cp(M2, t.M2);
    cp(beta, t.beta);
    cp(betaVal, t.betaVal);
    cp(betaColInd, t.betaColInd);
    cp(betaRowPtr, t.betaRowPtr);
    cp(pnOps, t.pnOps);
    return C3D_OK;
}

int c3d_closure(c3d_ctx ctx, int64_t ncells, double *u, double *alpha,
                const c3d_closure_opts *opts, c3d_cell_stats *stats, void *stream)
{
    if (!ctx || !opts || ncells < 0 || (ncells > 0 && (!u || !alpha)))
        return C3D_ERR_ARG;
    if (opts->maxIter < 1)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    const int g = gaunt_status(ctx, opts->useGaunt, opts->useGauntSparse);
    if (g != C3D_OK)
        return g;
    return hip_status(launch_closure_fp64(ctx, ncells, u, alpha, *opts, stats,
                                          (hipStream_t)stream));
}

int c3d_set_profiling(c3d_ctx ctx, int enable)
{
    if (!ctx)
        return C3D_ERR_ARG;
    ctx->profiling = enable != 0;
    return C3D_OK;
}

int c3d_get_closure_profile(c3d_ctx ctx, c3d_closure_profile *out)
{
    if (!ctx || !out)
        return C3D_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->passes = ctx->lastPasses;
    out->tail_launches = ctx->lastTailLaunches;
    if (!ctx->profiling)
        return C3D_OK;
    if (hipStreamSynchronize(ctx->lastStream) != hipSuccess)
        return C3D_ERR_HIP;
    unsigned long long n[3] = {0, 0, 0};
    if (hipMemcpy(n, ctx->d_sweepCount, sizeof(n), hipMemcpyDeviceToHost) != hipSuccess)
        return C3D_ERR_HIP;
    out->cell_sweeps = (int64_t)n[0];
    out->tail_cell_sweeps = (int64_t)n[1];
    out->mixed_cell_sweeps = (int64_t)n[2];
    for (const auto &sp : ctx->evSpans) {
        const int k = sp.second;
        float a = 0.f, b = 0.f, c = 0.f;
        if (sp.first == 0) {
            if (hipEventElapsedTime(&a, ctx->evPool[k], ctx->evPool[k + 1]) != hipSuccess ||
                hipEventElapsedTime(&b, ctx->evPool[k + 1], ctx->evPool[k + 2]) != hipSuccess ||
                hipEventElapsedTime(&c, ctx->evPool[k + 2], ctx->evPool[k + 3]) != hipSuccess)
                return C3D_ERR_HIP;
            out->mixed_sweep_ms += a;
            out->sweep_ms += b;
            out->step_ms += c;
        } else {
            if (hipEventElapsedTime(&a, ctx->evPool[k], ctx->evPool[k + 1]) != hipSuccess)
                return C3D_ERR_HIP;
            (sp.first == 1 ? out->tail_ms : out->warm_ms) += a;
        }
    }
    return C3D_OK;
}

int c3d_set_tail(c3d_ctx ctx, int max_cells, int passes_per_launch)
{
    if (!ctx || max_cells < 0 || passes_per_launch < 1 || passes_per_launch > 4096)
        return C3D_ERR_ARG;
    ctx->tailCells = max_cells;
    ctx->tailPasses = passes_per_launch;
    return C3D_OK;
}

int c3d_set_tail_wide(c3d_ctx ctx, int max_cells)
{
    if (!ctx || max_cells < 0)
        return C3D_ERR_ARG;
    ctx->tailWideCells = max_cells;
    return C3D_OK;
}

int c3d_set_launch(c3d_ctx ctx, int param, int value)
{
    if (!ctx)
        return C3D_ERR_ARG;
    switch (param) {
    case C3D_LAUNCH_MIXED_WAVES:
        if (value < 2 || value > 3)
            return C3D_ERR_ARG;
        ctx->tune.mixedWaves = value;
        return C3D_OK;
    case C3D_LAUNCH_MIXED_FINISH_COND:
        if (value < 0 || value > 300)
            return C3D_ERR_ARG;
        ctx->mixedFinishCond = std::pow(10.0, value / 10.0);
        return C3D_OK;
    case C3D_LAUNCH_ASYNC:
        if (value < 0 || value > 10000)
            return C3D_ERR_ARG;
        ctx->asyncPasses = value;
        return C3D_OK;
    default:
        return C3D_ERR_ARG;
    }
}

int c3d_set_precision(c3d_ctx ctx, int mode)
{
    if (!ctx || (mode != C3D_PREC_FP64 && mode != C3D_PREC_FP64_REFORDER &&
                 mode != C3D_PREC_MIXED))
        return C3D_ERR_ARG;
    ctx->precision = mode;
    return C3D_OK;
}

int c3d_set_mixed(c3d_ctx ctx, double cond_max, int max_iter, double switch_tol)
{
    if (!ctx || !(cond_max >= 1.0) || max_iter < 1 || !(switch_tol > 0.0))
        return C3D_ERR_ARG;
    ctx->mixedCondMax = cond_max;
    ctx->mixedMaxIter = max_iter;
    ctx->mixedSwitch = switch_tol;
    return C3D_OK;
}

int c3d_mlp_train(c3d_ctx ctx, int64_t n, const double *u, const double *alpha, int mse_steps,
                  int dual_steps, int batch, float lr, uint64_t seed, float *loss_first,
                  float *loss_last, void *stream)
{
    if (!ctx || n < 1 || n > INT32_MAX || !u || !alpha || mse_steps < 0 || dual_steps < 0 ||
        mse_steps + dual_steps < 1 || batch < 1 || batch > (1 << 22) || !(lr > 0.0f))
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    return mlp_train(ctx, n, u, alpha, mse_steps, dual_steps, batch, lr, seed, loss_first,
                     loss_last, (hipStream_t)stream);
}

int c3d_mlp_get(c3d_ctx ctx, float *params, c3d_mlp_norm *norm)
{
    if (!ctx)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    return mlp_get(ctx, params, norm ? &norm->x_mean[0] : nullptr);
}

int c3d_mlp_set(c3d_ctx ctx, const float *params, const c3d_mlp_norm *norm)
{
    if (!ctx || !params || !norm)
        return C3D_ERR_ARG;
    for (int k = 0; k < C3D_MLP_IN; ++k)
        if (!(norm->x_std[k] > 0.0f) || !(norm->alpha_std[k] > 0.0f))
            return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    return mlp_set(ctx, params, &norm->x_mean[0]);
}

int c3d_set_warmstart(c3d_ctx ctx, int enable)
{
    if (!ctx)
        return C3D_ERR_ARG;
    if (enable && !ctx->mlp.have)
        return C3D_ERR_ARG;
    ctx->warm = enable != 0;
    return C3D_OK;
}

int c3d_mlp_features(c3d_ctx ctx, int64_t n, const double *u, const double *alpha,
                     const c3d_mlp_norm *norm, float *X, float *T, void *stream)
{
    if (!ctx || n < 0 || !u || !norm || !X || (T && !alpha))
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    hipStream_t s = (hipStream_t)stream;
    c3d::DevBuf tmp;  // the caller's normalisation (the network's own copy stays intact)
    hipError_t e = grow(tmp, sizeof(c3d_mlp_norm));
    if (e == hipSuccess)
        e = hipMemcpyAsync(tmp.p, norm, sizeof(c3d_mlp_norm), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = mlp_features(ctx, n, u, T ? alpha : nullptr, (const float *)tmp.p, X, T, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (tmp.p)
        (void)hipFree(tmp.p);
    return hip_status(e);
}

int c3d_mlp_forward(c3d_ctx ctx, int64_t n, const float *params, const float *X, float *Y,
                    void *stream)
{
    if (!ctx || n < 0 || !params || !X || !Y)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    return hip_status(mlp_forward(ctx, n, params, X, Y, (hipStream_t)stream));
}

int c3d_mlp_infer(c3d_ctx ctx, int64_t n, const float *X, double *Yd, void *stream)
{
    if (!ctx || n < 0 || !X || !Yd)
        return C3D_ERR_ARG;
    if (!ctx->mlp.have)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    return hip_status(mlp_infer(ctx, n, X, Yd, (hipStream_t)stream));
}

int c3d_mlp_loss_grad(c3d_ctx ctx, int64_t n, const float *params, const float *X,
                      const float *T, float *grad, float *loss, void *stream)
{
    if (!ctx || n < 1 || !params || !X || !T || !grad || !loss)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    return hip_status(mlp_loss_grad(ctx, n, params, X, T, grad, loss, (hipStream_t)stream));
}

int c3d_mlp_dual_loss_grad(c3d_ctx ctx, int64_t n, const float *params, const float *X,
                           const double *u, const c3d_mlp_norm *norm, float *grad, float *loss,
                           void *stream)
{
    if (!ctx || n < 1 || n > INT32_MAX || !params || !X || !u || !norm || !grad || !loss)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    hipStream_t s = (hipStream_t)stream;
    c3d::DevBuf tmp;
    hipError_t e = grow(tmp, sizeof(c3d_mlp_norm));
    if (e == hipSuccess)
        e = hipMemcpyAsync(tmp.p, norm, sizeof(c3d_mlp_norm), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = mlp_dual_loss_grad(ctx, n, params, X, u, (const float *)tmp.p, grad, loss, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (tmp.p)
        (void)hipFree(tmp.p);
    return hip_status(e);
}

int c3d_init_grid(c3d_ctx ctx, int nx, int ny, int nz, const double *a, const double *d, int ic,
                  double sigma, double floor_value, double noise, uint64_t seed,
                  double cross_section, double *u, double *alpha, double *sigmaT,
                  double *sigmaS, void *stream)
{
    if (!ctx || nx < 1 || ny < 1 || nz < 1 || !a || !d || ic < 0 || ic > 2 || !u || !sigmaT ||
        !sigmaS || (ic != 2 && !(sigma > 0.0)))
        return C3D_ERR_ARG;
    return hip_status(launch_init_grid(nx, ny, nz, ctx->t.nm, a, d, ic, sigma, floor_value,
                                       noise, seed, cross_section, u, alpha, sigmaT, sigmaS,
                                       (hipStream_t)stream));
}

int c3d_set_pass_budget(c3d_ctx ctx, int max_passes)
{
    if (!ctx || max_passes < 1)
        return C3D_ERR_ARG;
    ctx->maxPasses = max_passes;
    return C3D_OK;
}

int c3d_hessian(c3d_ctx ctx, int64_t ncells, const double *alpha, double *hs, int useGaunt,
                int useGauntSparse, void *stream)
{
    if (!ctx || ncells < 0 || (ncells > 0 && (!alpha || !hs)))
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    const int g = gaunt_status(ctx, useGaunt, useGauntSparse);
    if (g != C3D_OK)
        return g;
    return hip_status(launch_hessian_fp64(ctx, ncells, alpha, hs, (hipStream_t)stream));
}

int c3d_dense_solve(int nm, int64_t n, const double *H, const double *g, double *d,
                    double *rcond, int32_t *info, double *L, int lapackOrder, void *stream)
{
    if (n < 0 || (n > 0 && (!H || !g || !d || !rcond || !info || !L)))
        return C3D_ERR_ARG;
    if (nm != 4 && nm != 9 && nm != 16 && nm != 25)
        return C3D_ERR_ARG;
    return hip_status(launch_dense_solve(nm, n, H, g, d, rcond, info, L, lapackOrder,
                                         (hipStream_t)stream));
}

int c3d_sync(c3d_ctx ctx, void *stream)
{
    if (!ctx)
        return C3D_ERR_ARG;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
        return C3D_ERR_HIP;
    int st = 0;
    if (hipMemcpy(&st, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return C3D_ERR_HIP;
    if (st != 0 && hipMemset(ctx->d_status, 0, sizeof(int)) != hipSuccess)
        return C3D_ERR_HIP;
    if (st & C3D_FLAG_U0_NONPOS)
        return C3D_ERR_U0;
    if (st & C3D_FLAG_LADDER)
        return C3D_ERR_LADDER;
    if (st & C3D_FLAG_BUDGET)
        return C3D_ERR_BUDGET;
    return C3D_OK;
}

int c3d_ghost_fill(c3d_ctx ctx, int nx, int ny, int nz, double *u, void *stream)
{
    if (!ctx || nx < 2 || ny < 2 || nz < 2 || !u)
        return C3D_ERR_ARG;
    return hip_status(launch_ghost_fill(nx, ny, nz, ctx->t.nm, u, (hipStream_t)stream));
}

int c3d_flux_kinetic(c3d_ctx ctx, int nx, int ny, int nz, const double *alpha, double *flux,
                     double dx, double dy, double dz, double theta, int isPn, void *stream)
{
    if (!ctx || nx < 1 || ny < 1 || nz < 1 || !alpha || !flux)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    const size_t ng = (size_t)(nx + 4) * (ny + 4) * (nz + 4);
    if (grow(ctx->ansatz, ng * ctx->t.nq * sizeof(double)) != hipSuccess)
        return C3D_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    double *A = (double *)ctx->ansatz.p;
    hipError_t e = launch_ansatz(ctx, nx, ny, nz, alpha, A, isPn, s);
    if (e == hipSuccess)
        e = launch_stencil(ctx, nx, ny, nz, A, flux, nullptr, nullptr, nullptr, 0.0, dx, dy, dz,
                           theta, 0, 0, s);
    return hip_status(e);
}

int c3d_flux_moment(c3d_ctx ctx, int nx, int ny, int nz, const double *u, double *flux,
                    double dx, double dy, double dz, void *stream)
{
    if (!ctx || nx < 1 || ny < 1 || nz < 1 || !u || !flux)
        return C3D_ERR_ARG;
    return hip_status(launch_flux_moment(ctx, nx, ny, nz, u, flux, dx, dy, dz,
                                         (hipStream_t)stream));
}

int c3d_euler(c3d_ctx ctx, int nx, int ny, int nz, double *u, const double *flux,
              const double *sigmaT, const double *sigmaS, double dt, void *stream)
{
    if (!ctx || nx < 1 || ny < 1 || nz < 1 || !u || !flux || !sigmaT || !sigmaS)
        return C3D_ERR_ARG;
    return hip_status(launch_euler(nx, ny, nz, ctx->t.nm, u, flux, sigmaT, sigmaS, dt,
                                   (hipStream_t)stream));
}

int c3d_rk2_combine(c3d_ctx ctx, int nx, int ny, int nz, double *u, double *uOld, int kind,
                    void *stream)
{
    if (!ctx || nx < 1 || ny < 1 || nz < 1 || !u || !uOld || (kind != 0 && kind != 1))
        return C3D_ERR_ARG;
    return hip_status(launch_rk2(nx, ny, nz, ctx->t.nm, u, uOld, kind, (hipStream_t)stream));
}

int c3d_update(c3d_ctx ctx, int nx, int ny, int nz, double *u, double *alpha,
               const double *sigmaT, const double *sigmaS, double dx, double dy, double dz,
               double theta, double dt, int solver, int pnUseKinetic,
               const c3d_closure_opts *opts, c3d_cell_stats *stats, void *stream)
{
    if (!ctx || nx < 2 || ny < 2 || nz < 2 || !u || !sigmaT || !sigmaS)
        return C3D_ERR_ARG;
    if (solver == 1 && (!alpha || !opts || opts->maxIter < 1))
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    if (solver == 1) {
        const int g = gaunt_status(ctx, opts->useGaunt, opts->useGauntSparse);
        if (g != C3D_OK)
            return g;
    }
    const int nm = ctx->t.nm;
    const size_t ng = (size_t)(nx + 4) * (ny + 4) * (nz + 4);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = grow(ctx->uOld, ng * nm * sizeof(double));
    const bool needAnsatz = solver == 1 || pnUseKinetic;
    if (e == hipSuccess && needAnsatz)
        e = grow(ctx->ansatz, ng * ctx->t.nq * sizeof(double));
    if (e == hipSuccess && !needAnsatz)
        e = grow(ctx->flux, ng * nm * sizeof(double));
    if (e != hipSuccess)
        return C3D_ERR_HIP;
    double *uOld = (double *)ctx->uOld.p, *A = (double *)ctx->ansatz.p,
           *flux = (double *)ctx->flux.p;
    e = launch_rk2(nx, ny, nz, nm, u, uOld, 0, s);
    for (int st = 0; st < 2 && e == hipSuccess; ++st) {
        e = launch_ghost_fill(nx, ny, nz, nm, u, s);
        if (e != hipSuccess)
            break;
        if (solver == 1) {
            c3d_closure_opts oc = *opts;
            if (oc.unbatched)
                oc.tolGamma = 1.1;  // solveOptimization passes 1.1 to opt(), update.cpp:57
            e = launch_closure_fp64(ctx, (int64_t)ng, u, alpha, oc,
                                    stats ? stats + st * ng : nullptr, s);
            if (e == hipSuccess)
                e = launch_ansatz(ctx, nx, ny, nz, alpha, A, 0, s);
            if (e == hipSuccess)
                e = launch_stencil(ctx, nx, ny, nz, A, nullptr, u, sigmaT, sigmaS, dt, dx, dy, dz,
                                   theta, 1, 0, s);
        } else if (!pnUseKinetic) {
            e = launch_flux_moment(ctx, nx, ny, nz, u, flux, dx, dy, dz, s);
            if (e == hipSuccess)
                e = launch_euler(nx, ny, nz, nm, u, flux, sigmaT, sigmaS, dt, s);
        } else {
            e = launch_ansatz(ctx, nx, ny, nz, u, A, 1, s);
            if (e == hipSuccess)
                e = launch_stencil(ctx, nx, ny, nz, A, nullptr, u, sigmaT, sigmaS, dt, dx, dy, dz,
                                   theta, 1, 0, s);
        }
    }
    if (e == hipSuccess)
        e = launch_rk2(nx, ny, nz, nm, u, uOld, 1, s);
    return hip_status(e);
}

int c3d_make_closure_batch(c3d_ctx ctx, int64_t ncells, uint64_t seed, double smin,
                           double smax, double *u, double *alpha0, double *alphaStar,
                           void *stream)
{
    if (!ctx || ncells < 0 || (ncells > 0 && (!u || !alpha0)))
        return C3D_ERR_ARG;
    return hip_status(launch_make_batch(ctx, ncells, seed, smin, smax, u, alpha0, alphaStar,
                                        (hipStream_t)stream));
}

int c3d_stats_summary(c3d_ctx ctx, int64_t ncells, const c3d_cell_stats *stats,
                      c3d_stats_sum *out, void *stream)
{
    if (!ctx || !out || ncells < 0 || (ncells > 0 && !stats))
        return C3D_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    if (ncells == 0)
        return C3D_OK;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    constexpr int kFields = 10;
    const int64_t want = (ncells + 255) / 256;
    const int blocks = (int)std::min<int64_t>(want, (int64_t)std::max(ctx->numCU, 1) * 4);
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = grow(ctx->statsPart, (size_t)blocks * kFields * sizeof(long long));
    std::vector<long long> part((size_t)blocks * kFields);
    if (e == hipSuccess)
        e = launch_stats_partials(ncells, stats, (long long *)ctx->statsPart.p, blocks, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(part.data(), ctx->statsPart.p, part.size() * sizeof(long long),
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return C3D_ERR_HIP;
    for (int b = 0; b < blocks; ++b) {
        const long long *v = &part[(size_t)b * kFields];
        out->cells += v[0];
        out->sumIter += v[1];
        out->sumIterGamma += v[2];
        out->maxIter = std::max<int64_t>(out->maxIter, v[3]);
        out->maxIterGamma = std::max<int64_t>(out->maxIterGamma, v[4]);
        for (int r = 0; r < 5; ++r)
            out->rung[r] += v[5 + r];
    }
    return C3D_OK;
}

/* ------------------------------------------------------------------------------
 * Drop-ins for src/cuda/cuda_headers.h.  Like the reference, these keep their device
 * state in file statics created on first use and block on every call.  Their staging
 * buffers (alphas/Hessians, the alpha and flux grids) are kept and grown on demand, so a
 * steady stream of calls of one size allocates nothing after the first.  They return void,
 * as the reference's do; failures are printed (as cuda_solveH.cu:13-20 prints CUDA errors)
 * and the last call's status is kept for c3d_seam_status().
 * ------------------------------------------------------------------------------ */
static std::mutex g_dropin_mtx;
static c3d_ctx g_h_ctx = nullptr;  // solveHInit_hip / solveH_hip
static c3d_ctx g_f_ctx = nullptr;  // solveFluxKinetic_hip
static int g_f_nm = 0, g_f_nq = 0;
static DevBuf g_h_alpha, g_h_hess;  // solveH_hip staging
static DevBuf g_f_alpha, g_f_flux;  // solveFluxKinetic_hip staging
static int g_seam_status = C3D_OK;

static int order_from_nm(int nm)
{
    int N = (int)std::lround(std::sqrt((double)nm)) - 1;
    return ((N + 1) * (N + 1) == nm) ? N : -1;
}
static int quad_from_nq(int nq)
{
    int Q = (int)std::lround(std::sqrt(nq / 2.0));
    return (2 * Q * Q == nq) ? Q : -1;
}

static void seam_fail(const char *who, int rc, hipError_t e)
{
    g_seam_status = rc;
    if (e != hipSuccess)
        std::fprintf(stderr, "%s: %s\n", who, hipGetErrorString(e));
    else
        std::fprintf(stderr, "%s: error %d\n", who, rc);
}

int c3d_seam_status(void)
{
    std::lock_guard<std::mutex> lk(g_dropin_mtx);
    return g_seam_status;
}

void solveHInit_hip(int nm, int nm2, int nq, int batchSize, double *w, double *p, double *p2,
                    double *beta, double *betaVal, int *betaColInd, int *betaRowPtr,
                    int useGaunt)
{
    (void)nm2;
    std::lock_guard<std::mutex> lk(g_dropin_mtx);
    g_seam_status = C3D_OK;
    int N = order_from_nm(nm), Q = quad_from_nq(nq);
    if (N < 1 || N > 4 || Q < 1 || !w || !p) {
        seam_fail("solveHInit_hip: unsupported nm/nq or missing tables", C3D_ERR_ARG,
                  hipSuccess);
        return;
    }
    if (g_h_ctx)
        c3d_destroy(g_h_ctx);
    g_h_ctx = nullptr;
    // xi/eta/mu are not needed for H; pass w as a placeholder of the right length
    int nnz = (useGaunt && betaRowPtr) ? betaRowPtr[nm * nm] : 0;
    int rc = c3d_create_from_tables(&g_h_ctx, N, Q, w, w, w, w, p, p2, useGaunt ? beta : nullptr,
                                    nnz, betaVal, betaColInd, betaRowPtr);
    if (rc != C3D_OK) {
        seam_fail("solveHInit_hip", rc, hipSuccess);
        return;
    }
    // staging for the batches the caller announces (optBatch.cpp:364-365 passes its size)
    hipError_t e = hipSuccess;
    if (batchSize > 0) {
        e = grow(g_h_alpha, (size_t)batchSize * nm * sizeof(double));
        if (e == hipSuccess)
            e = grow(g_h_hess, (size_t)batchSize * nm * nm * sizeof(double));
    }
    if (e != hipSuccess)
        seam_fail("solveHInit_hip", C3D_ERR_HIP, e);
}

void solveH_hip(int nm, int nm2, int nq, int batchSize, double *alphas, double *hs,
                int useGaunt, int useGauntSparse)
{
    (void)nm2;
    std::lock_guard<std::mutex> lk(g_dropin_mtx);
    g_seam_status = C3D_OK;
    if (!g_h_ctx || g_h_ctx->t.nm != nm || g_h_ctx->t.nq != nq || batchSize < 0 ||
        (batchSize > 0 && (!alphas || !hs))) {
        seam_fail("solveH_hip: not initialised for these sizes (solveHInit_hip) or bad args",
                  C3D_ERR_ARG, hipSuccess);
        return;
    }
    if (batchSize == 0)
        return;
    const size_t na = (size_t)batchSize * nm, nh = (size_t)batchSize * nm * nm;
    hipError_t e = grow(g_h_alpha, na * sizeof(double));
    if (e == hipSuccess)
        e = grow(g_h_hess, nh * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(g_h_alpha.p, alphas, na * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        seam_fail("solveH_hip", C3D_ERR_HIP, e);
        return;
    }
    int rc = c3d_hessian(g_h_ctx, batchSize, (const double *)g_h_alpha.p, (double *)g_h_hess.p,
                         useGaunt, useGauntSparse, nullptr);
    if (rc != C3D_OK) {
        seam_fail("solveH_hip", rc, hipSuccess);
        return;
    }
    e = hipMemcpy(hs, g_h_hess.p, nh * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        seam_fail("solveH_hip", C3D_ERR_HIP, e);
}

void solveFluxKinetic_hip(int nx, int ny, int nz, int nm, int nq, double dx, double dy,
                          double dz, double *alpha, double *flux, double *xi, double *eta,
                          double *mu, double *w, double *p, int isPn)
{
    std::lock_guard<std::mutex> lk(g_dropin_mtx);
    g_seam_status = C3D_OK;
    int N = order_from_nm(nm), Q = quad_from_nq(nq);
    if (N < 1 || N > 4 || Q < 1 || nx < 5 || ny < 5 || nz < 5 || !alpha || !flux || !xi ||
        !eta || !mu || !w || !p) {
        seam_fail("solveFluxKinetic_hip: unsupported sizes or missing arrays", C3D_ERR_ARG,
                  hipSuccess);
        return;
    }
    if (!g_f_ctx || g_f_nm != nm || g_f_nq != nq) {
        if (g_f_ctx)
            c3d_destroy(g_f_ctx);
        g_f_ctx = nullptr;
        int rc = c3d_create_from_tables(&g_f_ctx, N, Q, w, xi, eta, mu, p, nullptr, nullptr, 0,
                                        nullptr, nullptr, nullptr);
        if (rc != C3D_OK) {
            seam_fail("solveFluxKinetic_hip: context creation", rc, hipSuccess);
            return;
        }
        g_f_nm = nm;
        g_f_nq = nq;
    }
    const size_t n = (size_t)nx * ny * nz * nm;
    hipError_t e = grow(g_f_alpha, n * sizeof(double));
    if (e == hipSuccess)
        e = grow(g_f_flux, n * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(g_f_alpha.p, alpha, n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lk2(g_f_ctx->mtx);
        const int ix = nx - 4, iy = ny - 4, iz = nz - 4;
        const size_t ng = (size_t)nx * ny * nz;
        e = grow(g_f_ctx->ansatz, ng * nq * sizeof(double));
        double *A = (double *)g_f_ctx->ansatz.p;
        if (e == hipSuccess)
            e = launch_ansatz(g_f_ctx, ix, iy, iz, (const double *)g_f_alpha.p, A, isPn, nullptr);
        // theta = 2, as src/cuda/cuda_flux_kinetic.cu:82 hard-codes it
        if (e == hipSuccess)
            e = launch_stencil(g_f_ctx, ix, iy, iz, A, (double *)g_f_flux.p, nullptr, nullptr,
                               nullptr, 0.0, dx, dy, dz, 2.0, 0, 1, nullptr);
    }
    if (e == hipSuccess)
        e = hipMemcpy(flux, g_f_flux.p, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        seam_fail("solveFluxKinetic_hip", C3D_ERR_HIP, e);
}

}  // extern "C"
