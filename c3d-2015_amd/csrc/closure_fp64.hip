// closure_fp64.hip -- the M_N dual Newton closure in FP64 on gfx950, as a sequence of
// bounded passes over a compacted worklist of unconverged cells.
//
// Semantics: Optimization::opt with solveTheRest (src/opt/optBatch.cpp:143-282, 312-344,
// 350-472) or opt()/optScaled (src/opt/opt.cpp:139-355); the Hessian of solveHMn
// (src/opt/fobj.cpp:34-102); fobjMnUseH (fobj.cpp:171-185); the Armijo line search
// (src/opt/linesearch.cpp:29-81); LAPACK dpotrf/dpocon/dpotrs and BLAS dnrm2 restated
// lane-parallel below.
//
// One pass = two kernels over the current worklist W (cell ids):
//   k_sweep  evaluates, for every listed cell at its current alpha, e_q = exp(alpha.m_q),
//            sum_q w_q e_q and H = M^T diag(w e) M.  A wavefront takes 4 cells; the block's
//            16 cells share every M chunk staged in LDS, each lane computes one ordinate's
//            exponent for the 4 cells (fused multiply-adds in moment order), and the four
//            Hessians accumulate on the matrix cores (V_MFMA_F64_16X16X4_F64, 4 ordinates
//            per instruction, A = (w e m_q)^T, B = m_q; nm = 25 uses the three 16x16 tiles
//            of the lower block triangle).  Operand lanes: A[i][k], B[k][j] sit in lane
//            16k + i (16k + j); accumulator register r of lane l holds D[l/16 + 4r][l%16].
//   k_step   advances each listed cell by exactly one step of its state machine with that
//            sweep's result: either the Armijo test of a line-search trial, or one Newton
//            iteration (Cholesky in dpotf2 order, the dlacn2/dtrsv condition estimate, the
//            restart/ladder decisions, the solve of utils.cpp:71-86, dnrm2, modcfl, line-
//            search setup).  Row i of H lives in lane i; every scalar is reduced in the
//            CPU's sequential order without FMA contraction.  Unfinished cells append
//            themselves to the next worklist.
// Every loop in both kernels is bounded by sizes fixed at launch; the host caps the number
// of passes, so the closure always drains.
#include <algorithm>
#include <cfloat>

#include "c3d_detmath.h"
#include "c3d_internal.h"

#pragma clang fp contract(off)

namespace c3d {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kCellsPerWave = 4;
constexpr int kNumReg = 5;
__constant__ double kReg[kNumReg] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};  // opt.h:37-38

enum : int { kPhaseH = 0, kPhaseLS = 1 };

// per-cell state, [ncells] of each; vectors are [ncells][kVec][nm]
enum : int { kVa = 0, kVa0, kVaOld, kVd, kVg, kVu, kVuOrig, kVaIn, kVec };
enum : int { kSu0 = 0, kSf, kStls, kStprev, kSmaxF, kSgd, kSerr, kSgamma, kSr, kSrtol, kSshift,
             kScal };
enum : int { kIiter = 0, kItot, kIgam, kIreg, kIflags, kIphase, kImaxIt, kInt };

struct State {
    double *vec;   // [ncells][kVec][nm]
    double *scal;  // [ncells][kScal]
    int *ints;     // [ncells][kInt]
    double *H;     // [ncells][nm][nm]
    double *fdir;  // [ncells]  sum_q w_q e_q
};

struct KParams {
    int64_t ncells;
    int nq, nqPad, maxIter, unbatched;
    double tol, tolGamma, condHMax;
    double m0, isoAlpha0, minf;
    // mixed mode: a cell leaves the FP32-Hessian iteration for the FP64 path, restarting
    // from its input on condest > mixedCondMax, a failed factorisation or line search, or
    // mixedMaxIter iterations; continuing from its iterate once ||g|| <= mixedSwitch
    int mixed, mixedMaxIter;
    double mixedCondMax, mixedSwitch, mixedFinishCond;
};

// The worklists of one pass generation: cells in the mixed mode's FP32-Hessian phase and
// cells on the FP64 path (every cell outside the mixed mode).  Slot s of a pass is
// Wm[s] for s < *nm, else Wf[s - *nm].
struct Lists {
    int *Wm, *Wf;
    int *nm, *nf;
};

__device__ __forceinline__ int list_cell(const Lists &L, int slot, int nWm)
{
    return slot < nWm ? L.Wm[slot] : L.Wf[slot - nWm];
}

__device__ __forceinline__ void list_append(const Lists &L, bool mixedPhase, int c)
{
    if (mixedPhase)
        L.Wm[atomicAdd(L.nm, 1)] = c;
    else
        L.Wf[atomicAdd(L.nf, 1)] = c;
}

// Appends of a block to the next pass's lists, gathered in LDS and flushed with one global
// atomic per list and batch (a global atomic per cell on one counter serialises at the L2).
// push: one lane per cell, at most perRound pushes between two flush() calls; flush:
// called by the whole block at a block-uniform point.
constexpr int kAggCap = 256;
struct BlockAppender {
    int *buf;   // LDS [2][kAggCap]
    int *cnt;   // LDS [2]
    int *base;  // LDS [2]
    __device__ __forceinline__ void init()
    {
        if (threadIdx.x < 2)
            cnt[threadIdx.x] = 0;
        __syncthreads();
    }
    __device__ __forceinline__ void push(bool mixedPhase, int c)
    {
        const int l = mixedPhase ? 0 : 1;
        buf[l * kAggCap + atomicAdd(&cnt[l], 1)] = c;
    }
    // after a barrier: flush if forced or if another round of pushes might overflow
    __device__ __forceinline__ void flush(const Lists &L, bool force, int perRound)
    {
        __syncthreads();
        const int c0 = cnt[0], c1 = cnt[1];
        if (!force && c0 <= kAggCap - perRound && c1 <= kAggCap - perRound)
            return;
        if (threadIdx.x == 0) {
            base[0] = c0 ? atomicAdd(L.nm, c0) : 0;
            base[1] = c1 ? atomicAdd(L.nf, c1) : 0;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < c0; t += blockDim.x)
            L.Wm[base[0] + t] = buf[t];
        for (int t = threadIdx.x; t < c1; t += blockDim.x)
            L.Wf[base[1] + t] = buf[kAggCap + t];
        __syncthreads();
        if (threadIdx.x < 2)
            cnt[threadIdx.x] = 0;
        __syncthreads();
    }
};

#define C3D_BLOCK_APPENDER(name)                                                          \
    __shared__ int name##_buf[2 * kAggCap];                                                \
    __shared__ int name##_cnt[2];                                                          \
    __shared__ int name##_base[2];                                                         \
    BlockAppender name{name##_buf, name##_cnt, name##_base};                               \
    name.init()

__device__ __forceinline__ double rdl(double v, int src)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1)
        v = v + __shfl_xor(v, o);
    return v;
}

// Sum over the lane's 16-lane row (butterfly: every lane of the row gets the same bits).
__device__ __forceinline__ double row16_sum(double v)
{
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1)
        v = v + __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ float row16_max(float v)
{
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// A cell's lanes: the whole wave (G = 64, one cell per wave; broadcasts by readlane) or one
// 16-lane row (G = 16, four cells per wave; broadcasts by DPP row_newbcast).  Lane indices
// passed to the algebra below are indices within the group.  The arithmetic a cell sees is
// the same in both layouts.
template <int G>
struct Grp;

template <>
struct Grp<64> {
    static __device__ __forceinline__ double bc(double v, int j) { return rdl(v, j); }
    static __device__ __forceinline__ double bc_dyn(double v, int j) { return rdl(v, j); }
    static __device__ __forceinline__ bool all(bool p) { return __all(p); }
};

template <int J>
__device__ __forceinline__ double row_bcast(double v)
{
    // every lane of an active row receives lane J's value, so no old value is needed
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + J, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + J, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

template <>
struct Grp<16> {
    // j is a constant after unrolling; the switch folds away
    static __device__ __forceinline__ double bc(double v, int j)
    {
        switch (j) {
        case 0: return row_bcast<0>(v);
        case 1: return row_bcast<1>(v);
        case 2: return row_bcast<2>(v);
        case 3: return row_bcast<3>(v);
        case 4: return row_bcast<4>(v);
        case 5: return row_bcast<5>(v);
        case 6: return row_bcast<6>(v);
        case 7: return row_bcast<7>(v);
        case 8: return row_bcast<8>(v);
        case 9: return row_bcast<9>(v);
        case 10: return row_bcast<10>(v);
        case 11: return row_bcast<11>(v);
        case 12: return row_bcast<12>(v);
        case 13: return row_bcast<13>(v);
        case 14: return row_bcast<14>(v);
        default: return row_bcast<15>(v);
        }
    }
    // a lane index only known at run time (uniform within the row)
    static __device__ __forceinline__ double bc_dyn(double v, int j)
    {
        return __shfl(v, (int)(__lane_id() & 48u) + j);
    }
    static __device__ __forceinline__ bool all(bool p)
    {
        const unsigned long long b = __ballot(p);
        return ((b >> (__lane_id() & 48u)) & 0xFFFFull) == 0xFFFFull;
    }
};

// Pairwise (tree) reductions of NM values: depth ceil(log2 NM) instead of NM.
template <int NM>
__device__ __forceinline__ double tree_sum(double (&t)[NM])
{
#pragma unroll
    for (int w = 1; w < NM; w *= 2)
#pragma unroll
        for (int i = 0; i + w < NM; i += 2 * w)
            t[i] = t[i] + t[i + w];
    return t[0];
}
// the same over the first n entries (n fixed after unrolling)
template <int NM>
__device__ __forceinline__ double tree_sum_n(double (&t)[NM], int n)
{
#pragma unroll
    for (int w = 1; w < NM; w *= 2)
#pragma unroll
        for (int i = 0; i + w < NM; i += 2 * w)
            if (i + w < n)
                t[i] = t[i] + t[i + w];
    return n > 0 ? t[0] : 0.0;
}
template <int NM>
__device__ __forceinline__ double tree_max(double (&t)[NM])
{
#pragma unroll
    for (int w = 1; w < NM; w *= 2)
#pragma unroll
        for (int i = 0; i + w < NM; i += 2 * w)
            t[i] = t[i + w] > t[i] ? t[i + w] : t[i];
    return t[0];
}

// BLAS level-1 reductions over a lane vector (lane i < NM holds element i).  EXACT keeps the
// reference BLAS's sequential order; otherwise the NM broadcasts are independent and the
// reduction is a tree (other rounding; idamax keeps the first index of the maximum).
template <int NM, int G = 64, bool EXACT = true>
struct LaneVec {
    static __device__ __forceinline__ double dasum(double x)
    {
        if constexpr (!EXACT) {
            double t[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i)
                t[i] = fabs(Grp<G>::bc(x, i));
            return tree_sum(t);
        }
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i)
            s = s + fabs(Grp<G>::bc(x, i));
        return s;
    }
    static __device__ __forceinline__ int idamax(double x)
    {
        if constexpr (!EXACT) {
            double t[NM];
            int k[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                t[i] = fabs(Grp<G>::bc(x, i));
                k[i] = i;
            }
#pragma unroll
            for (int w = 1; w < NM; w *= 2)
#pragma unroll
                for (int i = 0; i + w < NM; i += 2 * w)
                    if (t[i + w] > t[i]) {
                        t[i] = t[i + w];
                        k[i] = k[i + w];
                    }
            return k[0];
        }
        int im = 0;
        double m = fabs(Grp<G>::bc(x, 0));
#pragma unroll
        for (int i = 1; i < NM; ++i) {
            double v = fabs(Grp<G>::bc(x, i));
            if (v > m) {
                m = v;
                im = i;
            }
        }
        return im;
    }
    static __device__ __forceinline__ double dnrm2(double x)  // reference BLAS scaling
    {
        if constexpr (!EXACT) {  // max-scaled sum of squares, one reciprocal
            double v[NM], t[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                v[i] = fabs(Grp<G>::bc(x, i));
                t[i] = v[i];
            }
            const double scale = tree_max(t);
            if (!(scale > 0.0) || !isfinite(scale))
                return scale;  // 0, inf or NaN
            const double rs = 1.0 / scale;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double r = v[i] * rs;
                t[i] = r * r;
            }
            return scale * sqrt(tree_sum(t));
        }
        double scale = 0.0, ssq = 1.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            const double xi = Grp<G>::bc(x, i);
            const double a = fabs(xi);
            const bool up = scale < a;
            const double r = (up ? scale : a) / (up ? a : scale);  // one division either way
            if (xi != 0.0) {
                if (up) {
                    ssq = 1.0 + ssq * (r * r);
                    scale = a;
                } else {
                    ssq = ssq + r * r;
                }
            }
        }
        return scale * sqrt(ssq);
    }
};

// Wave-resident factor: lane i < NM holds row i (L after Cholesky; L(i,i) in row[i]) and
// the wave's LDS copy of the factor (sT[i ld + j] = L(i,j), j <= i) serves the transposed
// solves.  EXACT divides by the diagonal as LAPACK/BLAS do, one row at a time.  Otherwise
// the triangular solves run by 4 x 4 blocks: the inverse T of each diagonal block is formed
// once per factor (lane i keeps row i%4 and column i%4 of its block's T), and a block step
// is one small product with T plus the update of the later rows, so a solve's dependent
// chain is NM/4 block steps instead of NM scalar ones (the same solution, other rounding).
template <int NM, bool EXACT = true, int G = 64>
struct Dense {
    static constexpr int NB = (NM + 3) / 4;  // 4 x 4 blocks of the blocked solves
    double row[NM];
    const double *sT = nullptr;  // LDS copy of L, row-major with leading dimension ld
    int ld = 0;
    double dinv;
    double tr[4], tc[4];  // row and column lane%4 of the lane's diagonal-block inverse

    // L(j, lane) for j > lane (the column of L below the diagonal, read from LDS)
    __device__ __forceinline__ double col(int j, int lane) const
    {
        return (j > lane && lane < NM) ? sT[j * ld + lane] : 0.0;
    }

    __device__ __forceinline__ double diag(int j) const { return Grp<G>::bc(row[j], j); }

    // dpotf2 'L': ajj = a(j,j) - ddot; column by dgemv then dscal(1/ajj)
    __device__ __forceinline__ int cholesky(int lane)
    {
        if constexpr (!EXACT)
            return cholesky_short(lane);
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < j; ++k)
                s = s + row[k] * row[k];
            double ajj = Grp<G>::bc(row[j] - s, j);
            if (!(ajj > 0.0))
                return j + 1;
            ajj = sqrt(ajj);
            double y = row[j];
#pragma unroll
            for (int k = 0; k < j; ++k) {
                double temp = -Grp<G>::bc(row[k], j);
                y = y + temp * row[k];
            }
            double rr = 1.0 / ajj;
            if (lane == j)
                row[j] = ajj;
            else if (lane > j)
                row[j] = rr * y;
        }
        return 0;
    }

    // The same factor with a shorter dependent chain per column: the dot products as trees,
    // and sqrt(ajj), 1/sqrt(ajj) from the hardware reciprocal square root refined by two
    // Newton steps (rsq' = rsq (3 - a rsq^2) / 2) instead of a correctly rounded sqrt and a
    // division (the factor differs from dpotf2's in the last bits).
    __device__ __forceinline__ int cholesky_short(int lane)
    {
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            double t[NM];
#pragma unroll
            for (int k = 0; k < j; ++k)
                t[k] = row[k] * row[k];
            const double s = tree_sum_n(t, j);
            const double ajj = Grp<G>::bc(row[j] - s, j);
            if (!(ajj > 0.0))
                return j + 1;
#pragma unroll
            for (int k = 0; k < j; ++k)
                t[k] = Grp<G>::bc(row[k], j) * row[k];
            const double y = row[j] - tree_sum_n(t, j);
            double q;
            if (ajj > 1e-280 && ajj < 1e280) {  // far from the exponent range's ends
                q = __builtin_amdgcn_rsq(ajj);
                q = q * (1.5 - (0.5 * ajj) * (q * q));
                q = q * (1.5 - (0.5 * ajj) * (q * q));
            } else {
                q = 1.0 / sqrt(ajj);
            }
            if (lane == j)
                row[j] = ajj * q;
            else if (lane > j)
                row[j] = q * y;
        }
        return 0;
    }

    __device__ __forceinline__ void build_columns(double *sTw, int ldw, int lane)
    {
        sT = sTw;
        ld = ldw;
        if constexpr (!EXACT) {
            double djj = 1.0;
#pragma unroll
            for (int j = 0; j < NM; ++j)
                djj = (lane == j) ? row[j] : djj;
            dinv = 1.0 / djj;
        }
        wave_lds_sync();
        if (lane < NM) {
#pragma unroll
            for (int j = 0; j < NM; ++j)
                if (j <= lane)
                    sTw[lane * ldw + j] = row[j];
        }
        wave_lds_sync();
        if constexpr (!EXACT)
            block_inverse(lane);
    }

    // T = inverse of the lane's 4 x 4 diagonal block of L (identity rows beyond NM), by
    // forward substitution of L_bb T = I; the diagonal reciprocals are dinv's.
    __device__ __forceinline__ void block_inverse(int lane)
    {
        const int b0 = (lane >> 2) * 4, r = lane & 3;
        double L[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j)
                L[i][j] = (b0 + i < NM) ? sT[(b0 + i) * ld + b0 + j] : (i == j ? 1.0 : 0.0);
        double T[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
            T[i][i] = 1.0 / L[i][i];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int i = c + 1; i < 4; ++i) {
                double s = 0.0;
#pragma unroll
                for (int k = c; k < i; ++k)
                    s = s + L[i][k] * T[k][c];
                T[i][c] = -s * T[i][i];
            }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double vr = 0.0, vc = 0.0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (i == r) {
                    vr = (c <= i) ? T[i][c] : 0.0;
                    vc = (c >= i) ? T[c][i] : 0.0;
                }
            }
            tr[c] = vr;
            tc[c] = vc;
        }
    }

    // x <- L^-1 x by blocks (non-EXACT)
    __device__ __forceinline__ void trsv_ln_blocked(double &x, int lane)
    {
        const int myb = lane >> 2, r = lane & 3;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            double xb[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                xb[c] = (4 * b + c < NM) ? Grp<G>::bc(x, 4 * b + c) : 0.0;
            if (myb == b && lane < NM) {
                double y = 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (c <= r)
                        y = y + tr[c] * xb[c];
                x = y;
            }
            if (b + 1 < NB) {
                double yb[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    yb[c] = (4 * b + c < NM) ? Grp<G>::bc(x, 4 * b + c) : 0.0;
                if (lane >= 4 * b + 4 && lane < NM) {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (4 * b + c < NM)
                            x = x - row[4 * b + c] * yb[c];
                }
            }
        }
    }

    // x <- L^-T x by blocks (non-EXACT)
    __device__ __forceinline__ void trsv_lt_blocked(double &x, int lane)
    {
        const int myb = lane >> 2, r = lane & 3;
#pragma unroll
        for (int b = NB - 1; b >= 0; --b) {
            double xb[4];
#pragma unroll
            for (int c = 0; c < 4; ++c)
                xb[c] = (4 * b + c < NM) ? Grp<G>::bc(x, 4 * b + c) : 0.0;
            if (myb == b && lane < NM) {
                double y = 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (c >= r)
                        y = y + tc[c] * xb[c];
                x = y;
            }
            if (b > 0) {
                double yb[4];
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    yb[c] = (4 * b + c < NM) ? Grp<G>::bc(x, 4 * b + c) : 0.0;
                if (lane < 4 * b) {
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        if (4 * b + c < NM)
                            x = x - col(4 * b + c, lane) * yb[c];
                }
            }
        }
    }

    __device__ __forceinline__ void trsv_ln(double &x, int lane)  // dtrsv('L','N','N')
    {
        if constexpr (!EXACT) {
            trsv_ln_blocked(x, lane);
            return;
        }
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            double xj = Grp<G>::bc(x, j);
            if (xj != 0.0) {
                if constexpr (EXACT)
                    xj = xj / diag(j);
                else
                    xj = xj * Grp<G>::bc(dinv, j);
                if (lane == j)
                    x = xj;
                else if (lane > j)
                    x = x - xj * row[j];
            }
        }
    }
    __device__ __forceinline__ void trsv_lt(double &x, int lane)  // dtrsv('L','T','N')
    {
        if constexpr (!EXACT) {
            trsv_lt_blocked(x, lane);
            return;
        }
#pragma unroll
        for (int j = NM - 1; j >= 0; --j) {
            if (lane == j)
                x = EXACT ? x / row[j] : x * dinv;
            double xj = Grp<G>::bc(x, j);
            if (lane < j)
                x = x - col(j, lane) * xj;
        }
    }
    __device__ __forceinline__ void potrs_custom(double &d, int lane)  // utils.cpp:71-86
    {
        if constexpr (!EXACT) {
            trsv_ln_blocked(d, lane);
            trsv_lt_blocked(d, lane);
            return;
        }
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            if (lane == j)
                d = EXACT ? d / row[j] : d * dinv;
            double dj = Grp<G>::bc(d, j);
            if (lane > j)
                d = d - row[j] * dj;
        }
#pragma unroll
        for (int j = NM - 1; j >= 0; --j) {
            if (lane == j)
                d = EXACT ? d / row[j] : d * dinv;
            double dj = Grp<G>::bc(d, j);
            if (lane < j)
                d = d - col(j, lane) * dj;
        }
    }
    // LAPACK dpotrs('L'): dtrsm(L) then dtrsm(L^T), reference BLAS order (opt.cpp:285).  Only
    // the unbatched semantics use it, so it is kept compact (rolled loops, factor from LDS).
    __device__ __forceinline__ void potrs_lapack(double &d, int lane)
    {
#pragma unroll 1
        for (int k = 0; k < NM; ++k) {
            double dk = Grp<G>::bc_dyn(d, k);
            if (dk != 0.0) {
                dk = EXACT ? dk / sT[k * ld + k] : dk * Grp<G>::bc_dyn(dinv, k);
                if (lane == k)
                    d = dk;
                else if (lane > k && lane < NM)
                    d = d - dk * sT[lane * ld + k];
            }
        }
#pragma unroll 1
        for (int i = NM - 1; i >= 0; --i) {
            double temp = Grp<G>::bc_dyn(d, i);
#pragma unroll 1
            for (int k = i + 1; k < NM; ++k)
                temp = temp - sT[k * ld + i] * Grp<G>::bc_dyn(d, k);
            temp = EXACT ? temp / sT[i * ld + i] : temp * Grp<G>::bc_dyn(dinv, i);
            if (lane == i)
                d = temp;
        }
    }

    // (max_i L_ii / min_i L_ii)^2 over the factor's diagonal
    __device__ __forceinline__ double diag_ratio_sq(int lane) const
    {
        double djj = 1.0;
#pragma unroll
        for (int j = 0; j < NM; ++j)
            djj = (lane == j) ? row[j] : djj;
        double dmax = Grp<G>::bc(djj, 0), dmin = dmax;
#pragma unroll
        for (int i = 1; i < NM; ++i) {
            const double v = Grp<G>::bc(djj, i);
            dmax = fmax(dmax, v);
            dmin = fmin(dmin, v);
        }
        const double r = dmax / dmin;
        return r * r;
    }

    // dpocon('L'): dlacn2 (Higham 1988, the Hager/Higham 1-norm estimator) driving
    // dlatrs, which for these factors takes its dtrsv path.  At most 11 applications.
    __device__ __forceinline__ double pocon(double anorm, int lane)
    {
        if (anorm == 0.0)
            return 0.0;
        const bool own = lane < NM;
        double x = own ? 1.0 / (double)NM : 0.0;
        int isgn = 0;
        double est = 0.0;
        int isave1 = 1, isave2 = 0, isave3 = 0;
        for (int guard = 0; guard < 16; ++guard) {
            trsv_ln(x, lane);
            trsv_lt(x, lane);
            bool l120 = false;
            if (isave1 == 1) {
                est = LaneVec<NM, G, EXACT>::dasum(x);
                if (own) {
                    x = (x >= 0.0) ? 1.0 : -1.0;
                    isgn = (x >= 0.0) ? 1 : -1;
                }
                isave1 = 2;
                continue;
            } else if (isave1 == 2) {
                isave2 = LaneVec<NM, G, EXACT>::idamax(x);
                isave3 = 2;
                x = (lane == isave2) ? 1.0 : 0.0;
                isave1 = 3;
                continue;
            } else if (isave1 == 3) {
                double estold = est;
                est = LaneVec<NM, G, EXACT>::dasum(x);
                bool same = !own || (((x >= 0.0) ? 1 : -1) == isgn);
                if (Grp<G>::all(same) || est <= estold) {
                    l120 = true;
                } else {
                    if (own) {
                        x = (x >= 0.0) ? 1.0 : -1.0;
                        isgn = (x >= 0.0) ? 1 : -1;
                    }
                    isave1 = 4;
                    continue;
                }
            } else if (isave1 == 4) {
                int jlast = isave2;
                isave2 = LaneVec<NM, G, EXACT>::idamax(x);
                if (Grp<G>::bc_dyn(x, jlast) != fabs(Grp<G>::bc_dyn(x, isave2)) && isave3 < 5) {
                    ++isave3;
                    x = (lane == isave2) ? 1.0 : 0.0;
                    isave1 = 3;
                    continue;
                }
                l120 = true;
            } else {
                double temp = 2.0 * (LaneVec<NM, G, EXACT>::dasum(x) / (double)(3 * NM));
                if (temp > est)
                    est = temp;
                break;
            }
            if (l120) {
                x = own ? (((lane & 1) ? -1.0 : 1.0) * (1.0 + (double)lane / (double)(NM - 1)))
                        : 0.0;
                isave1 = 5;
            }
        }
        return est != 0.0 ? (1.0 / est) / anorm : 0.0;
    }
};

// ---------------------------------------------------------------- init pass
// preProcess (optBatch.cpp:312-331 / opt.cpp:336-346) and the first worklist.
template <int NM>
__device__ __forceinline__ void init_cell(const KParams &P, const double *__restrict__ u_g,
                                          const double *__restrict__ alpha_g, const State &S,
                                          int64_t c, c3d_cell_stats *stats, int *status,
                                          BlockAppender &app, int lane)
{
    const bool own = lane < NM;
    double uu = own ? u_g[c * NM + lane] : 0.0;
    double a = own ? alpha_g[c * NM + lane] : 0.0;
    const double u0 = rdl(uu, 0);
    const double alphaScale = det_log(u0) * 2.0 * sqrt(M_PI);
    if (lane == 0)
        a = a - alphaScale;
    uu = uu / u0;
    if (u0 <= 0) {
        if (lane == 0) {
            atomicOr(status, C3D_FLAG_U0_NONPOS);
            if (stats) {
                c3d_cell_stats s0{0.0, 0.0, 0.0, 0, 0, 0, C3D_FLAG_U0_NONPOS};
                stats[c] = s0;
            }
        }
        return;
    }
    double *v = S.vec + c * kVec * NM;
    if (own) {
        v[kVa * NM + lane] = a;
        v[kVa0 * NM + lane] = a;
        v[kVaOld * NM + lane] = a;
        v[kVd * NM + lane] = 0.0;
        v[kVg * NM + lane] = 0.0;
        v[kVu * NM + lane] = uu;
        v[kVuOrig * NM + lane] = uu;
        v[kVaIn * NM + lane] = a;
    }
    const double rtol = P.unbatched ? LaneVec<NM>::dnrm2(own ? uu : 0.0) * P.tol : P.tol;
    if (lane == 0) {
        double *s = S.scal + c * kScal;
        s[kSu0] = u0;
        s[kSf] = 0.0;
        s[kStls] = 1.0;
        s[kStprev] = 1.0;
        s[kSmaxF] = 0.0;
        s[kSgd] = 0.0;
        s[kSerr] = 0.0;
        s[kSgamma] = 0.0;
        s[kSr] = 0.0;
        s[kSrtol] = rtol;
        s[kSshift] = 0.0;
        int *in = S.ints + c * kInt;
        in[kIiter] = 0;
        in[kItot] = 0;
        in[kIgam] = 0;
        in[kIreg] = 0;
        in[kIflags] = 0;
        in[kIphase] = kPhaseH;
        in[kImaxIt] = P.maxIter;
        app.push(P.mixed != 0, (int)c);
    }
}

template <int NM>
__global__ __launch_bounds__(kBlock) void k_init(KParams P, const double *__restrict__ u_g,
                                                  const double *__restrict__ alpha_g, State S,
                                                  Lists L, c3d_cell_stats *stats, int *status)
{
    C3D_BLOCK_APPENDER(app);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t nBlocksNeeded = (P.ncells + kWavesPerBlock - 1) / kWavesPerBlock;
    for (int64_t blk = blockIdx.x; blk < nBlocksNeeded; blk += gridDim.x) {
        const int64_t c = blk * kWavesPerBlock + wv;
        if (c < P.ncells)
            init_cell<NM>(P, u_g, alpha_g, S, c, stats, status, app, lane);
        app.flush(L, blk + gridDim.x >= nBlocksNeeded, kWavesPerBlock);
    }
}


// ---------------------------------------------------------------- sweep building blocks
// FP32 copy of an M chunk for the matrix-core operands, laid out so that lane (rq, gq) of
// the MFMA finds M[4k + gq][rq] for k = 0..15 in one 16-float row: row (col * 4 + gq), rows
// kRS floats apart (16-byte aligned), element k in 4-float group (k >> 2) ^ ((col >> 2) & 3)
// of the row, so that the 16 lanes of a group reading group kb of their rows cover all 64
// banks.  Rows of columns nm..16 NT-1 stay zero.
constexpr int kRS = 20;
template <int NM>
struct MixedChunk {
    static constexpr int NT = (NM + 15) / 16;
    static constexpr int kFloats = 16 * NT * 4 * kRS;
    static __device__ __forceinline__ int index(int rr, int cc)
    {
        const int k = rr >> 2;
        return (cc * 4 + (rr & 3)) * kRS + 4 * ((k >> 2) ^ ((cc >> 2) & 3)) + (k & 3);
    }
    // zero the padding columns of a buffer (once per kernel, any number of threads)
    static __device__ __forceinline__ void clear_pad(float *sMfT, int tid, int nthreads)
    {
        for (int e = tid; e < (16 * NT - NM) * 4 * kRS; e += nthreads)
            sMfT[NM * 4 * kRS + e] = 0.0f;
    }
};

// M and w chunks of kChunk ordinates staged through double-buffered LDS by a whole block:
// each thread holds its share of the next chunk in registers while the block consumes the
// current one, so one barrier per chunk publishes a chunk and the global loads overlap the
// matrix-core work.
template <int NM>
struct ChunkStager {
    static constexpr int NMS = NM | 1;
    static constexpr int kPer = (kChunk * NM + kBlock - 1) / kBlock;
    double pre[kPer];
    double prew;
    __device__ __forceinline__ void fetch(const double *__restrict__ Mg,
                                          const double *__restrict__ wg, int ch, int tid)
    {
        const double *src = Mg + (size_t)ch * kChunk * NM;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int e = tid + kBlock * k;
            pre[k] = (e < kChunk * NM) ? src[e] : 0.0;
        }
        prew = tid < kChunk ? wg[(size_t)ch * kChunk + tid] : 0.0;
    }
    __device__ __forceinline__ void put(double *sM, double *sW, int tid) const
    {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int e = tid + kBlock * k;
            if (e < kChunk * NM) {
                const int rr = e / NM, cc = e - rr * NM;
                sM[rr * NMS + cc] = pre[k];
            }
        }
        if (tid < kChunk)
            sW[tid] = prew;
    }
};

// The chunk ordinate whose exponent and weight lane l computes in the FP64 sweep.
__device__ __forceinline__ int mfma_ordinate(int lane) { return lane; }

// Matrix-core accumulation of one LDS chunk of kChunk ordinates for CW cells: lane l owns
// ordinate mfma_ordinate(l) of the chunk for the exponent (wl is that ordinate's weight),
// then the MFMA loop consumes 4 ordinates per instruction.  The operation sequence per cell
// does not depend on CW or on the other cells of the wave, so a cell's H is the same bits
// whichever kernel evaluates it.
template <int NM, int CW>
__device__ __forceinline__ void chunk_mfma(const double *sMc, double wl, const double *sAc,
                                           d4 (&acc)[CW][(NM + 15) / 16 * ((NM + 15) / 16 + 1) / 2],
                                           double (&fpart)[CW], int lane)
{
    constexpr int NT = (NM + 15) / 16;
    constexpr int NMS = NM | 1;
    const int gq = lane >> 4, rq = lane & 15;
    const double *rowp = &sMc[mfma_ordinate(lane) * NMS];
    double we[CW];
#pragma unroll
    for (int c = 0; c < CW; ++c) {
        double x = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i)
            x = fma(sAc[c * NM + i], rowp[i], x);
        we[c] = wl * exp(x);
        fpart[c] = fpart[c] + we[c];
    }
#pragma unroll
    for (int k = 0; k < kChunk / 4; ++k) {
        const int ql = 4 * k + gq;
        const double *mr = &sMc[ql * NMS];
        const double b0 = (rq < NM) ? mr[rq] : 0.0;
        const double b1 = (NT == 2 && 16 + rq < NM) ? mr[16 + rq] : 0.0;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const double wq = __shfl(we[c], ql);
            acc[c][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(wq * b0, b0, acc[c][0], 0, 0, 0);
            if constexpr (NT == 2) {
                acc[c][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(wq * b1, b0, acc[c][1], 0, 0, 0);
                acc[c][2] = __builtin_amdgcn_mfma_f64_16x16x4f64(wq * b1, b1, acc[c][2], 0, 0, 0);
            }
        }
    }
}

// chunk_mfma for four cells and nm <= 9 on V_MFMA_F64_4X4X4_4B_F64 (as fast per flop as the
// 16x16 instruction): the lower-triangle 4x4 blocks (R, C) of H, one instruction per block
// with the four cells in its four block slots, in place of a 16x16 tile per cell of which
// nm^2 entries are used.  Lane l: cell (l >> 2) & 3, row or column l & 3 of the block, K index
// l >> 4 (ordinate 4 k + (l >> 4) of K step k, as in chunk_mfma).  Each element's chain of
// fused multiply-adds over the ordinates, and its operands w_q m_qi (rounded) and m_qj, are
// chunk_mfma's, so H has the same bits; the exponents and f are chunk_mfma's code.
template <int NM>
struct Blk4 {
    static constexpr int NR = (NM + 3) / 4;
    static constexpr int NBLK = NR * (NR + 1) / 2;
};
template <int NM>
__device__ __forceinline__ void chunk_mfma4(const double *sMc, double wl, const double *sAc,
                                            double (&acc)[Blk4<NM>::NBLK], double (&fpart)[4],
                                            int lane)
{
    static_assert(NM <= 9, "the 16x16 tile serves larger orders");
    constexpr int NR = Blk4<NM>::NR;
    constexpr int NMS = NM | 1;
    const double *rowp = &sMc[mfma_ordinate(lane) * NMS];
    double we[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double x = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i)
            x = fma(sAc[c * NM + i], rowp[i], x);
        we[c] = wl * exp(x);
        fpart[c] = fpart[c] + we[c];
    }
    const int ii = lane & 3, kq = lane >> 4, cb = (lane >> 2) & 3;
#pragma unroll
    for (int k = 0; k < kChunk / 4; ++k) {
        const int ql = 4 * k + kq;
        const double *mr = &sMc[ql * NMS];
        double mv[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r)
            mv[r] = (4 * r + ii < NM) ? mr[4 * r + ii] : 0.0;
        double w4[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            w4[c] = __shfl(we[c], ql);
        const double wq = cb == 0 ? w4[0] : (cb == 1 ? w4[1] : (cb == 2 ? w4[2] : w4[3]));
        int p = 0;
#pragma unroll
        for (int R = 0; R < NR; ++R) {
            const double a = wq * mv[R];
#pragma unroll
            for (int C = 0; C <= R; ++C) {
                acc[p] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, mv[C], acc[p], 0, 0, 0);
                ++p;
            }
        }
    }
}

// H of the lane's cell from chunk_mfma4's blocks: element (4 R + (l >> 4), 4 C + (l & 3)),
// lower triangle, mirrored (store_H's writes).
template <int NM>
__device__ __forceinline__ void store_H4(const double (&acc)[Blk4<NM>::NBLK], double *H, int lane)
{
    const int i0 = lane >> 4, j0 = lane & 3;
    int p = 0;
#pragma unroll
    for (int R = 0; R < Blk4<NM>::NR; ++R)
#pragma unroll
        for (int C = 0; C <= R; ++C) {
            const int i = 4 * R + i0, j = 4 * C + j0;
            if (i < NM && j < NM && j <= i) {
                H[i * NM + j] = acc[p];
                H[j * NM + i] = acc[p];
            }
            ++p;
        }
}

// H from one cell's accumulators: lower triangle, mirrored (fobj.cpp:96-101).
template <int NM>
__device__ __forceinline__ void store_H(const d4 *acc, double *H, int lane)
{
    constexpr int NT = (NM + 15) / 16;
    constexpr int NB = NT * (NT + 1) / 2;
    const int gq = lane >> 4, rq = lane & 15;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int R = (b == 0) ? 0 : 1, Cb = (b == 2) ? 1 : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = 16 * R + gq + 4 * k, j = 16 * Cb + rq;
            if (i < NM && j < NM && j <= i) {
                H[i * NM + j] = acc[b][k];
                H[j * NM + i] = acc[b][k];
            }
        }
    }
}

// One cell's FP64 sweep for nm > 16 on a block of four waves, with chunk_mfma's bits: wave
// b < 3 runs the matrix-core chain of H's accumulator block b only (rows 16 R.., columns
// 16 Cb.., as store_H numbers them), every working wave evaluates the chunk's exponents and
// weights as chunk_mfma does, and wave 0 sums f.  The chunks are the row image's (rows with
// leading dimension nm | 1, then the chunk's weights), staged by the whole block into two LDS
// buffers with one barrier per chunk.  Every thread of the block calls it; it returns f in
// wave 0 (0 elsewhere) and leaves H in sH, published by its last barrier.
template <int NM>
__device__ double sweep_fp64_blocks(int nqPad, const double *__restrict__ Rimg,
                                    const double *sAc, double *sR0, double *sR1, double *sH,
                                    int lane, int wv)
{
    static_assert(NM > 16, "one accumulator block per wave");
    constexpr int NMS = NM | 1;
    constexpr int KD = row_chunk_doubles(NM);
    constexpr int kPer = (KD + kBlock - 1) / kBlock;
    const int tid = wv * kWave + lane;
    const int nch = nqPad / kChunk;
    const int gq = lane >> 4, rq = lane & 15;
    const bool lowRows = wv == 0, highCols = wv == 2;  // block (R, Cb) = (0,0), (1,0), (1,1)
    double pre[kPer];
    auto fetch = [&](int ch) {
        const double *src = Rimg + (size_t)ch * KD;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int e = tid + kBlock * k;
            pre[k] = e < KD ? src[e] : 0.0;
        }
    };
    auto put = [&](double *dst) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int e = tid + kBlock * k;
            if (e < KD)
                dst[e] = pre[k];
        }
    };
    fetch(0);
    put(sR0);
    __syncthreads();
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    double fpart = 0.0;
    for (int ch = 0; ch < nch; ++ch) {
        const double *sMc = (ch & 1) ? sR1 : sR0;
        if (ch + 1 < nch)
            fetch(ch + 1);
        if (wv < 3) {
            const int qo = mfma_ordinate(lane);
            const double wl = sMc[kChunk * NMS + qo];
            const double *rowp = &sMc[qo * NMS];
            double x = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i)
                x = fma(sAc[i], rowp[i], x);
            const double we = wl * exp(x);
            fpart = fpart + we;
#pragma unroll
            for (int k = 0; k < kChunk / 4; ++k) {
                const int ql = 4 * k + gq;
                const double *mr = &sMc[ql * NMS];
                const double b0 = mr[rq];
                const double b1 = 16 + rq < NM ? mr[16 + rq] : 0.0;
                const double wq = __shfl(we, ql);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(wq * (lowRows ? b0 : b1),
                                                           highCols ? b1 : b0, acc, 0, 0, 0);
            }
        }
        if (ch + 1 < nch)
            put((ch & 1) ? sR0 : sR1);
        __syncthreads();
    }
    if (wv < 3) {
        const int R = lowRows ? 0 : 1, Cb = highCols ? 1 : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = 16 * R + gq + 4 * k, j = 16 * Cb + rq;
            if (i < NM && j < NM && j <= i) {
                sH[i * NM + j] = acc[k];
                sH[j * NM + i] = acc[k];
            }
        }
    }
    __syncthreads();
    return wv == 0 ? wave_sum(fpart) : 0.0;
}

// Entry-per-lane assignment of the lower triangle for the reference-order sweep.
template <int NM>
struct RefEntries {
    static constexpr int NE = NM * (NM + 1) / 2;
    static constexpr int NS = (NE + kWave - 1) / kWave;
    int ei[NS], ej[NS];
    __device__ __forceinline__ explicit RefEntries(int lane)
    {
#pragma unroll
        for (int s2 = 0; s2 < NS; ++s2) {
            int e = lane + kWave * s2, i = 0;
            while (i < NM && (i + 1) * (i + 2) / 2 <= e)
                ++i;
            ei[s2] = e < NE ? i : 0;
            ej[s2] = e < NE ? e - i * (i + 1) / 2 : 0;
        }
    }
};

// One chunk in the reference's order for one cell (one wave): the exponent of the lane's
// ordinate as in fobj.cpp:44-47 with the deterministic exp, then each lane adds
// ((w_q m_qi) m_qj) e_q for its entries in ordinate order (fobj.cpp:62-69) and
// sum_q w_q e_q in ordinate order (fobj.cpp:141-164).
template <int NM>
__device__ __forceinline__ void chunk_ref(const double *sMc, const double *sWc, double *sEc,
                                          const double *sAc, const RefEntries<NM> &E,
                                          double (&h)[RefEntries<NM>::NS], double &f, int nqc,
                                          int lane)
{
    constexpr int NMS = NM | 1;
    double x = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
        x = x + sAc[i] * sMc[lane * NMS + i];
    sEc[lane] = det_exp(x);
    wave_lds_sync();
    for (int q = 0; q < nqc; ++q) {  // padded ordinates skipped
        const double wq = sWc[q], tq = sEc[q];
#pragma unroll
        for (int s2 = 0; s2 < RefEntries<NM>::NS; ++s2)
            if (lane + kWave * s2 < RefEntries<NM>::NE)
                h[s2] += wq * sMc[q * NMS + E.ei[s2]] * sMc[q * NMS + E.ej[s2]] * tq;
        f = f + wq * tq;
    }
    wave_lds_sync();
}

template <int NM>
__device__ __forceinline__ void store_H_ref(const RefEntries<NM> &E,
                                            const double (&h)[RefEntries<NM>::NS], double *H,
                                            int lane)
{
#pragma unroll
    for (int s2 = 0; s2 < RefEntries<NM>::NS; ++s2)
        if (lane + kWave * s2 < RefEntries<NM>::NE) {
            H[E.ei[s2] * NM + E.ej[s2]] = h[s2];
            H[E.ej[s2] * NM + E.ei[s2]] = h[s2];
        }
}

// ---------------------------------------------------------------- sweep pass
// A wavefront takes 4 cells; the block's 16 cells share every M chunk staged in LDS.
template <int NM>
__global__ __launch_bounds__(kBlock) void k_sweep(KParams P, const double *__restrict__ Mg,
                                                   const double *__restrict__ wg, State S,
                                                   const int *__restrict__ W,
                                                   const int *__restrict__ nWp)
{
    constexpr int NT = (NM + 15) / 16;
    constexpr int NB = NT * (NT + 1) / 2;
    constexpr int NMS = NM | 1;
    constexpr int CW = kCellsPerWave;
    __shared__ double sM[2][kChunk * NMS];
    __shared__ double sW[2][kChunk];
    __shared__ double sA[kWavesPerBlock][CW * NM];

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nch = P.nqPad / kChunk;
    const int nW = uniform_int(*nWp);
    const int groups = (nW + CW - 1) / CW;
    const int nBlocksNeeded = (groups + kWavesPerBlock - 1) / kWavesPerBlock;
    for (int blk = blockIdx.x; blk < nBlocksNeeded; blk += gridDim.x) {
        const int grp = blk * kWavesPerBlock + wv;
        int cells[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int slot = grp * CW + c;
            cells[c] = (grp < groups && slot < nW) ? W[slot] : -1;
        }
#pragma unroll
        for (int c = 0; c < CW; ++c)
            if (lane < NM)
                sA[wv][c * NM + lane] =
                    cells[c] >= 0 ? S.vec[(size_t)cells[c] * kVec * NM + kVa * NM + lane] : 0.0;
        constexpr bool B4 = NM <= 9;  // 4x4 blocks, the four cells in the block slots
        static_assert(!B4 || CW == 4, "one cell per block slot");
        d4 acc[B4 ? 1 : CW][B4 ? 1 : NB];
        double acc4[Blk4<B4 ? NM : 4>::NBLK];
#pragma unroll
        for (int c = 0; c < (B4 ? 1 : CW); ++c)
#pragma unroll
            for (int b = 0; b < (B4 ? 1 : NB); ++b)
                acc[c][b] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int p = 0; p < Blk4<B4 ? NM : 4>::NBLK; ++p)
            acc4[p] = 0.0;
        double fpart[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c)
            fpart[c] = 0.0;
        const bool any = cells[0] >= 0;
        ChunkStager<NM> st;
        st.fetch(Mg, wg, 0, tid);
        __syncthreads();  // the previous cells' last chunk is consumed
        for (int ch = 0; ch < nch; ++ch) {
            const int b = ch & 1;
            st.put(sM[b], sW[b], tid);
            __syncthreads();
            if (ch + 1 < nch)
                st.fetch(Mg, wg, ch + 1, tid);
            if (any) {
                if constexpr (B4)
                    chunk_mfma4<NM>(sM[b], sW[b][mfma_ordinate(lane)], sA[wv], acc4, fpart, lane);
                else
                    chunk_mfma<NM, CW>(sM[b], sW[b][mfma_ordinate(lane)], sA[wv], acc, fpart,
                                       lane);
            }
        }
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const double fs = wave_sum(fpart[c]);
            if (cells[c] < 0)
                continue;
            if (lane == 0)
                S.fdir[cells[c]] = fs;
            if constexpr (!B4)
                store_H<NM>(acc[c], S.H + (size_t)cells[c] * NM * NM, lane);
        }
        if constexpr (B4) {
            const int cb = (lane >> 2) & 3;
            if (cells[cb] >= 0)
                store_H4<NM>(acc4, S.H + (size_t)cells[cb] * NM * NM, lane);
        }
    }
}

// ---------------------------------------- sweep pass, reference operation order
// One wavefront per cell; bit-identical to a CPU evaluation of solveHMn / fobjMn with the
// same exp.
template <int NM>
__global__ __launch_bounds__(kBlock) void k_sweep_ref(KParams P, const double *__restrict__ Mg,
                                                       const double *__restrict__ wg, State S,
                                                       const int *__restrict__ W,
                                                       const int *__restrict__ nWp)
{
    constexpr int NMS = NM | 1;
    __shared__ double sM[kChunk * NMS];
    __shared__ double sW[kChunk];
    __shared__ double sE[kWavesPerBlock][kChunk];
    __shared__ double sA[kWavesPerBlock][NM];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nW = uniform_int(*nWp);
    const int nBlocksNeeded = (nW + kWavesPerBlock - 1) / kWavesPerBlock;
    const RefEntries<NM> E(lane);
    for (int blk = blockIdx.x; blk < nBlocksNeeded; blk += gridDim.x) {
        const int slot = blk * kWavesPerBlock + wv;
        const int cell = slot < nW ? uniform_int(W[slot]) : -1;
        if (lane < NM)
            sA[wv][lane] = cell >= 0 ? S.vec[(size_t)cell * kVec * NM + kVa * NM + lane] : 0.0;
        double h[RefEntries<NM>::NS];
#pragma unroll
        for (int s2 = 0; s2 < RefEntries<NM>::NS; ++s2)
            h[s2] = 0.0;
        double f = 0.0;
        for (int ch = 0; ch < P.nqPad / kChunk; ++ch) {
            __syncthreads();
            for (int e = tid; e < kChunk * NM; e += kBlock) {
                const int rr = e / NM, cc = e - rr * NM;
                sM[rr * NMS + cc] = Mg[(size_t)ch * kChunk * NM + e];
            }
            if (tid < kChunk)
                sW[tid] = wg[(size_t)ch * kChunk + tid];
            __syncthreads();
            if (cell >= 0)
                chunk_ref<NM>(sM, sW, sE[wv], sA[wv], E, h, f, min(kChunk, P.nq - ch * kChunk),
                              lane);
        }
        if (cell < 0)
            continue;
        store_H_ref<NM>(E, h, S.H + (size_t)cell * NM * NM, lane);
        if (lane == 0)
            S.fdir[cell] = f;
    }
}

typedef float f4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------- mixed-precision sweep
// Phase one of the mixed mode, on the FP16 and FP64 matrix cores (the column sweep below):
//   * FP64: the exponents alpha.m_q and the gradient moments sum_q w_q e_q m_q on the FP64
//     matrix cores, the shifted weights w_q e^{alpha.m_q - shift} (the shift is log sum_q w e
//     at the cell's last Newton iterate) and their sum f (the Armijo test keeps FP64
//     resolution) on the vector units;
//   * the Hessian on V_MFMA_F32_16X16X32_F16 (32 ordinates per instruction, FP32
//     accumulation) as three products of FP16 halves of v = 2^12 w e and of the basis pair
//     products, v_hi P_hi + v_hi P_lo + v_lo P_hi (each split by rounding to nearest, then the
//     rounded remainder); about 21 significant bits; the fixed scale 2^12 and the range check
//     below keep the operands in FP16's range.
// Operand layouts of the 16x16x32 FP16 instruction: A[i][k] in lane i + 16 (k / 8), element
// k % 8; B[k][j] in lane j + 16 (k / 8), element k % 8; D[i][j] in lane j + 16 (i / 4), FP32
// register i % 4.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));


// e^x for the weights of the mixed sweeps (every mixed path computes them with it, so that they
// agree bit for bit): x = k ln2 / 64 + r by Cody-Waite reduction with fused multiply-adds
// (|r| <= ln2 / 128), e^x = 2^(k >> 6) 2^((k & 63) / 64) e^r with the middle factor from a
// 64-entry table (correctly rounded, 512 bytes that stay in the vector caches) and e^r by its
// degree-5 Taylor polynomial (truncation below 4e-17): relative error within 2 ulp for
// |x| < 700, NaN propagates.
__constant__ double kExp2Tab[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0,
    0x1.0874518759bc8p+0, 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0,
    0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, 0x1.172b83c7d517bp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0,
    0x1.2d285a6e4030bp+0, 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0,
    0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, 0x1.3dea64c123422p+0,
    0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0,
    0x1.56f4736b527dap+0, 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0,
    0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, 0x1.6a09e667f3bcdp+0,
    0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0,
    0x1.868d99b4492edp+0, 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0,
    0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, 0x1.9c49182a3f090p+0,
    0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0,
    0x1.bcc1e904bc1d2p+0, 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0,
    0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, 0x1.d5818dcfba487p+0,
    0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0,
    0x1.fa7c1819e90d8p+0,
};
// tab: the table's 64 entries, kExp2Tab or a copy of it in LDS.  Where LDS DMA copies or
// prefetched loads are in flight, the copy in LDS keeps the lookup from waiting for them (the
// vector-memory counter drains in order).
__device__ __forceinline__ double mexp(double x, const double *tab)
{
    const double kf = __builtin_rint(x * 0x1.71547652b82fep+6);  // 64 / ln2
    double r = fma(kf, -0x1.62e42fefa39efp-7, x);                // ln2 / 64, high part
    r = fma(kf, -0x1.abc9e3b39803fp-62, r);                      // and the rest
    const int k = (int)kf;
    double p = fma(r, 1.0 / 120.0, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(tab[k & 63] * p, k >> 6);
}
__device__ __forceinline__ double mexp(double x) { return mexp(x, kExp2Tab); }

// The FP16 operands carry the weights scaled by 2^12.  At the shift's own iterate the weights
// sum to one, so the largest lies in [1/nq, 1] and every operand in FP16's range with full
// precision for the terms that matter; each lane keeps a running maximum of its weights, and a
// sweep whose largest scaled weight leaves [2^-4, 2^14] (a trial point far from the shift's
// iterate) marks its H invalid (see h16_range_ok); a cell accepting such a point re-centres
// the shift there and sweeps again before its Newton iteration.
constexpr float kWScale = 4096.0f;
__device__ __forceinline__ bool h16_range_ok(float wmax)
{
    const float m = wmax * kWScale;
    return m >= 0.0625f && m <= 16384.0f;
}

// Exponent operands of V_MFMA_F64_4X4X4_4B_F64 with the cells as rows: the A operand of K
// step s is alpha[lane & 3][4 s + (lane >> 4)] (zero beyond NM and for rows >= the cells
// present), kept in registers for a whole sweep.
template <int NM>
struct ExpA {
    static constexpr int NKS = (NM + 3) / 4;
    double a[NKS];
    // alpha of row r at sA + r * NM for rows r < rows
    __device__ __forceinline__ void load(const double *sA, int rows, int lane)
    {
        const int r = lane & 3, k = lane >> 4;
#pragma unroll
        for (int s = 0; s < NKS; ++s) {
            const int m = 4 * s + k;
            a[s] = (r < rows && m < NM) ? sA[r * NM + m] : 0.0;
        }
    }
};

// The exponents x[i][q] = sum_m alpha[i][m] M[q][m] of ordinates q = 16 g + (lane & 15),
// g = 0..3, for row i = lane >> 4: one chain of K steps on the FP64 matrix cores per group,
// the four chains interleaved (K step outer); b(g, s) returns M[16 g + (lane & 15)][4 s +
// (lane >> 4)], zero for moments >= NM.
template <int NM, typename BF>
__device__ __forceinline__ void exp_arg4(const ExpA<NM> &A, BF b, double (&x)[4])
{
#pragma unroll
    for (int g = 0; g < 4; ++g)
        x[g] = 0.0;
#pragma unroll
    for (int s = 0; s < ExpA<NM>::NKS; ++s)
#pragma unroll
        for (int g = 0; g < 4; ++g)
            x[g] = __builtin_amdgcn_mfma_f64_4x4x4f64(A.a[s], b(g, s), x[g], 0, 0, 0);
}

// Row i of the packed lower triangle starts at i (i + 1) / 2.
__host__ __device__ constexpr int tri(int i) { return i * (i + 1) / 2; }

// ---------------------------------------------------------------- column sweep (nm <= 25)
// The mixed sweep for nm <= kPairMaxNm with the cells as the columns of its matrix products,
// 16 cells a wave, cell c in lanes l with l & 15 = c:
//   exponents  X = M alpha^T on V_MFMA_F64_16X16X4_F64 per block of 16 ordinates (A = the M
//              rows of the block, B = the cells' alpha, kept in registers);
//   gradient   G = W M on the same instruction, A = the weights exactly as the exponent products
//              leave them in the lanes (accumulator register r of lane l holds row (l >> 4) + 4 r,
//              column l & 15), so no transposition;
//   Hessian    the pair products H_p = sum_q v_q P_qp over the pairs p of the packed lower
//              triangle on V_MFMA_F32_16X16X32_F16, A = the FP16 halves of the scaled weights,
//              again from the lanes' own registers (the pair image's K order follows them,
//              c3d_internal.h), B = the pair image (16 pairs a tile), three products per 32
//              ordinates, hi hi + hi lo + lo hi.
// Each output element of a matrix product depends on its own row and column only, so a cell's
// results have the same bits in whichever column of whichever wave they are computed.
template <int NM>
struct PairHess {
    static constexpr int NP = pair_count(NM);
    static constexpr int PT = pair_tiles(NM);
    static constexpr int kHalves = pair_chunk_halves(NM);  // one chunk's image
    static constexpr int kUnits = kHalves / 512;            // its 1 KB units
};

// H = scale * acc with the cell's e^shift: the weights carry 2^12, the image kPairScale
__device__ __forceinline__ double pair_scale(double es) { return es / (kWScale * kPairScale); }

// The exponent products' B operand for K step s: alpha[4 s + (lane >> 4)] of the lane's cell
// (zero beyond NM and where alpha is null).
template <int NM>
struct ExpB {
    static constexpr int NKS = (NM + 3) / 4;
    double a[NKS];
    __device__ __forceinline__ void load(const double *alpha, int lane)
    {
#pragma unroll
        for (int s = 0; s < NKS; ++s) {
            const int m = 4 * s + (lane >> 4);
            a[s] = (alpha != nullptr && m < NM) ? alpha[m] : 0.0;
        }
    }
};

// Exponents of the four 16-ordinate blocks of a chunk: x[b][r] = alpha . m_q of the lane's
// cell at ordinate q = 16 b + (lane >> 4) + 4 r; am(b, s) returns M[16 b + (lane & 15)][4 s +
// (lane >> 4)] (zero for moments >= NM).  The four K-chains are interleaved (K step outer).
template <int NM, typename AF>
__device__ __forceinline__ void col_exponents(const ExpB<NM> &aT, AF am, d4 (&x)[4])
{
#pragma unroll
    for (int b = 0; b < 4; ++b)
        x[b] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < ExpB<NM>::NKS; ++s)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            x[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(am(b, s), aT.a[s], x[b], 0, 0, 0);
}

// The weight of one ordinate from its exponent: w_q e^{x - shift} (FP64), the running sum fo
// it belongs to, the running maximum, and the FP16 halves of the weight scaled by 2^12 (the
// FP32 difference v - hi is exact).
__device__ __forceinline__ double col_weight(double x, double wq, double shiftOwn, double &fo,
                                             float &wmaxOwn, _Float16 &hi, _Float16 &lo)
{
    const double we = wq * mexp(x - shiftOwn);
    fo = fo + we;
    const float wf = (float)we;
    wmaxOwn = fmaxf(wmaxOwn, wf);
    const float v = wf * kWScale;
    hi = (_Float16)v;
    lo = (_Float16)(v - (float)hi);
    return we;
}

// The gradient products of block b: gacc[t][b] += we[r] M[q][16 t + (lane & 15)] for
// q = 16 b + (lane >> 4) + 4 r, r ascending (accumulator b, one per block).
template <int NM>
__device__ __forceinline__ void col_grad_block(const double *sMc, const d4 &we, int b,
                                               d4 (&gacc)[(NM + 15) / 16][4], int lane)
{
    constexpr int NMS = NM | 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int q = 16 * b + (lane >> 4) + 4 * r;
#pragma unroll
        for (int t = 0; t < (NM + 15) / 16; ++t) {
            const int col = 16 * t + (lane & 15);
            const double bm = col < NM ? sMc[q * NMS + col] : 0.0;
            gacc[t][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(we[r], bm, gacc[t][b], 0, 0, 0);
        }
    }
}

// The pair products of a chunk into acc[i] for the tiles t = t0 + ts i < PT: ah/al the A
// operands of the two 32-ordinate blocks (the halves of blocks 2 kb and 2 kb + 1), sP the
// chunk's pair image.
template <int NM, int NA>
__device__ __forceinline__ void pair_hess(const _Float16 *sP, const h8 (&ah)[2], const h8 (&al)[2],
                                          f4 (&acc)[NA], int t0, int ts, int lane)
{
    constexpr int PT = PairHess<NM>::PT;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int t = t0 + ts * i;
            if (t < PT) {
                const _Float16 *bp = &sP[(kb * PT + t) * 1024 + lane * 8];
                const h8 bh = *reinterpret_cast<const h8 *>(bp);
                const h8 bl = *reinterpret_cast<const h8 *>(bp + 512);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[kb], bh, acc[i], 0, 0, 0);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[kb], bl, acc[i], 0, 0, 0);
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[kb], bh, acc[i], 0, 0, 0);
            }
        }
    }
}

// One chunk of the column sweep for the 16 cells of a wave (cell lane & 15): exponents,
// weights (running sums fo[r] of the ordinates o = (lane >> 4) + 4 r of every block, over
// blocks and chunks in order; maximum wmaxOwn), gradient products (accumulator per block) and
// pair products; sMc the chunk's FP64 rows (leading dimension NM | 1), sWc its weights, sPc its
// pair image, complete once image_ready() returns (called before the pair products).
template <int NM, typename RF>
__device__ __forceinline__ void chunk_cols(const double *sMc, const double *sWc,
                                           const _Float16 *sPc, const ExpB<NM> &aT,
                                           double shiftOwn, float &wmaxOwn, double (&fo)[4],
                                           d4 (&gacc)[(NM + 15) / 16][4],
                                           f4 (&acc)[PairHess<NM>::PT], int lane, RF image_ready)
{
    constexpr int NMS = NM | 1;
    const int i16 = lane & 15, g = lane >> 4;
    d4 x[4];
    col_exponents<NM>(aT, [&](int b, int s) {
        const int m = 4 * s + g;
        return (NM % 4 == 0 || m < NM) ? sMc[(16 * b + i16) * NMS + m] : 0.0;
    }, x);
    h8 ah[2], al[2];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            _Float16 hi, lo;
            x[b][r] = col_weight(x[b][r], sWc[16 * b + g + 4 * r], shiftOwn, fo[r], wmaxOwn, hi,
                                 lo);
            ah[b >> 1][4 * (b & 1) + r] = hi;
            al[b >> 1][4 * (b & 1) + r] = lo;
        }
#pragma unroll
    for (int b = 0; b < 4; ++b)
        col_grad_block<NM>(sMc, x[b], b, gacc, lane);
    image_ready();
    pair_hess<NM, PairHess<NM>::PT>(sPc, ah, al, acc, 0, 1, lane);
}

// f of the lane's cell from the running sums of its 16 ordinate slots o = (lane >> 4) + 4 r:
// the butterfly of row16_sum over o (partners o ^ 8, o ^ 4, o ^ 2, o ^ 1: registers r ^ 2,
// r ^ 1, then lanes ^ 32, ^ 16), so that a lone cell whose slot o sits in lane o of a 16-lane
// row, summed by row16_sum, gets the same bits; every lane of the column gets them.
__device__ __forceinline__ double col_f(const double (&fo)[4])
{
    double s1[4], s2[4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        s1[r] = fo[r] + fo[r ^ 2];
#pragma unroll
    for (int r = 0; r < 4; ++r)
        s2[r] = s1[r] + s1[r ^ 1];
    double f = s2[0];
    f = f + __shfl_xor(f, 32);
    return f + __shfl_xor(f, 16);
}
__device__ __forceinline__ float col_max(float v)
{
    v = fmaxf(v, __shfl_xor(v, 16));
    return fmaxf(v, __shfl_xor(v, 32));
}
// the gradient moments from the per-block accumulators (register r: cell (l >> 4) + 4 r)
__device__ __forceinline__ double col_g(const d4 (&ga)[4], int r)
{
    return (ga[0][r] + ga[1][r]) + (ga[2][r] + ga[3][r]);
}

// The LDS DMA (global_load_lds_dwordx4): lane i of the issuing wave copies the 16 bytes at
// src + 16 i to dst + 16 i.  dma_units copies `units` 1 KB units, unit j by the wave j % nw of
// the caller's nw waves (wv its index), one instruction per unit; the copy completes
// asynchronously, and each wave waits for its own (dma_wait: vmcnt 0; dma_wait_n<N>: all but
// its last N vector-memory instructions) before a barrier publishes it.  lds_reads_done waits
// for the wave's outstanding LDS reads, before a copy overwrites what they read.
__device__ __forceinline__ void dma_units(const void *src, void *dst, int units, int wv, int nw,
                                          int lane)
{
    // dst points into LDS: its low 32 bits are the LDS address (no aperture test needed)
    const uint32_t lds = (uint32_t)(uintptr_t)dst;
    for (int j = wv; j < units; j += nw)
        __builtin_amdgcn_global_load_lds(
            (const void *)((const char *)src + j * 1024 + lane * 16),
            (__attribute__((address_space(3))) void *)(uintptr_t)(lds + j * 1024), 16, 0, 0);
}
__device__ __forceinline__ void dma_wait() { __builtin_amdgcn_s_waitcnt(0x0F70); }
template <int N>
__device__ __forceinline__ void dma_wait_n()
{
    static_assert(N >= 0 && N < 64, "vmcnt holds 6 bits");
    __builtin_amdgcn_s_waitcnt(0x0F70 | (N & 15) | ((N >> 4) << 14));
}
__device__ __forceinline__ void lds_reads_done() { __builtin_amdgcn_s_waitcnt(0xC07F); }

// The row image's geometry (c3d_internal.h): a chunk's rows at 0, leading dimension NM | 1, its
// weights after them
template <int NM>
struct RowImg {
    static constexpr int NMS = NM | 1;
    static constexpr int kDoubles = row_chunk_doubles(NM);
    static constexpr int kUnits = kDoubles * 8 / 1024;
    static constexpr int kW = kChunk * NMS;  // offset of the weights
};

// One cell's mixed sweep for nm <= kPairMaxNm on NWV = 1 or 4 waves, with the bulk column
// sweep's bits for that cell.  FP64 matrix products round as sequential fused multiply-adds
// over K in order (both V_MFMA_F64_16X16X4_F64 and V_MFMA_F64_4X4X4_4B_F64), so the cell's
// exponents come from the cheaper four-row products (the cell in row 0: lanes 0..15 hold its
// ordinates 16 b + lane) and its gradient from vector FMA chains in the bulk's order (lane
// 16 b + j: moments 16 t + j of block b's accumulators, K order r then k); the Hessian's pair
// products
// run on the FP16 matrix cores with the cell in row 0 (A rows 1..15 zero).  With four waves,
// wave w computes the exponents of block w and the pair products of tiles w, w + 4, ..;
// wave 1 the gradient.  f sums the weights per ordinate slot o = lane (blocks and chunks in
// order) and reduces by row16_sum, the bulk's order.  Staging by LDS DMA: one wave
// single-buffers (the next chunk's rows copied after the gradient has read the rows, its pair
// image after the pair products; the waits count the wave's copies), four waves double-buffer
// (sR1, sP1).  sWE: 64 doubles; sRed: 4 floats.  Results: H (packed lower triangle, first entry
// NaN where the weights left the FP16 range, if mark) in sH, the gradient moments in sG; f
// returned (wave 0).
template <int NM, int NWV>
__device__ __forceinline__ double sweep_lone(int nqPad, const double *__restrict__ Rimg,
                                             const _Float16 *__restrict__ Pimg, const double *sA,
                                             double shift, double *sR0, double *sR1,
                                             _Float16 *sP0, _Float16 *sP1, double *sWE,
                                             float *sRed, double *sH, double *sG, bool mark,
                                             int lane, int wv, const double *xtab)
{
    using PH = PairHess<NM>;
    using RI = RowImg<NM>;
    constexpr int NMS = NM | 1;
    constexpr int NA = (PH::PT + NWV - 1) / NWV;
    static_assert(NWV == 1 || NWV == 4, "one wave or four");
    const int o = lane & 15;
    ExpA<NM> ax;
    ax.load(sA, 1, lane);
    const double shiftOwn = (lane >> 4) == 0 ? shift : 0.0;
    constexpr int NT = (NM + 15) / 16;
    float wmaxOwn = 0.0f;
    double fo = 0.0, gl[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
        gl[t] = 0.0;
    f4 acc[NA];
#pragma unroll
    for (int i = 0; i < NA; ++i)
        acc[i] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    auto sync = []() __attribute__((always_inline)) {
        if constexpr (NWV == 1)
            wave_lds_sync();
        else
            __syncthreads();
    };
    // one chunk from its rows sM (weights after them), its pair image sP (complete once
    // image_ready returns)
    auto chunk = [&](const double *sM, const _Float16 *sP,
                     auto image_ready) __attribute__((always_inline)) {
        const double *sW = sM + RI::kW;
#pragma unroll
        for (int bb = 0; bb < 4 / NWV; ++bb) {  // the blocks of this wave, one chain each
            const int b = NWV == 1 ? bb : wv;
            double x = 0.0;
#pragma unroll
            for (int s = 0; s < ExpA<NM>::NKS; ++s) {
                const int m = 4 * s + (lane >> 4);
                const double bm = (NM % 4 == 0 || m < NM) ? sM[(16 * b + o) * NMS + m] : 0.0;
                x = __builtin_amdgcn_mfma_f64_4x4x4f64(ax.a[s], bm, x, 0, 0, 0);
            }
            const double we = sW[16 * b + o] * mexp(x - shiftOwn, xtab);
            wmaxOwn = fmaxf(wmaxOwn, (float)we);
            if (lane < 16)
                sWE[16 * b + o] = we;
        }
        sync();
        if (wv == 0 && lane < 16) {
#pragma unroll
            for (int b = 0; b < 4; ++b)
                fo = fo + sWE[16 * b + o];
        }
        if (wv == (NWV == 1 ? 0 : 1)) {
            const int bg = lane >> 4;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int q = 16 * bg + k + 4 * r;
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        if (16 * t + o < NM)
                            gl[t] = fma(sWE[q], sM[q * NMS + 16 * t + o], gl[t]);
                }
        }
        h8 ah[2], al[2];
        const bool arow = o == 0;
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int q = 32 * kb + 16 * (e >> 2) + (lane >> 4) + 4 * (e & 3);
                const float v = arow ? (float)sWE[q] * kWScale : 0.0f;
                const _Float16 hi = (_Float16)v;
                ah[kb][e] = hi;
                al[kb][e] = (_Float16)(v - (float)hi);
            }
        image_ready();
        pair_hess<NM, NA>(sP, ah, al, acc, wv, NWV, lane);
    };
    const int nch = nqPad / kChunk;
    if constexpr (NWV == 1) {
        wave_lds_sync();
        lds_reads_done();  // earlier readers of sR0 and sP0 are done
        dma_units(Rimg, sR0, RI::kUnits, 0, 1, lane);
        dma_units(Pimg, sP0, PH::kUnits, 0, 1, lane);
        for (int ch = 0; ch < nch; ++ch) {
            const bool more = ch + 1 < nch;
            // the chunk's rows, issued before its pair image (of which up to 63 copies, the
            // counter's limit, may still be in flight)
            dma_wait_n<(PH::kUnits < 63 ? PH::kUnits : 63)>();
            wave_lds_sync();
            chunk(sR0, sP0, [&] {
                lds_reads_done();  // the rows are read
                if (more) {
                    dma_units(Rimg + (size_t)(ch + 1) * RI::kDoubles, sR0, RI::kUnits, 0, 1,
                              lane);
                    dma_wait_n<RI::kUnits>();  // the pair image (the next rows may be in flight)
                } else {
                    dma_wait();
                }
                wave_lds_sync();
            });
            if (more) {
                lds_reads_done();  // the pair products have read sP0
                dma_units(Pimg + (size_t)(ch + 1) * PH::kHalves, sP0, PH::kUnits, 0, 1, lane);
            }
        }
    } else if constexpr (NM <= 16) {
        __syncthreads();  // earlier readers of the buffers are done
        dma_units(Rimg, sR0, RI::kUnits, wv, 4, lane);
        dma_units(Pimg, sP0, PH::kUnits, wv, 4, lane);
        for (int ch = 0; ch < nch; ++ch) {
            dma_wait();
            __syncthreads();  // the chunk's images; the previous chunk's readers are done
            if (ch + 1 < nch) {
                dma_units(Rimg + (size_t)(ch + 1) * RI::kDoubles, (ch & 1) ? sR0 : sR1,
                          RI::kUnits, wv, 4, lane);
                dma_units(Pimg + (size_t)(ch + 1) * PH::kHalves, (ch & 1) ? sP0 : sP1,
                          PH::kUnits, wv, 4, lane);
            }
            chunk((ch & 1) ? sR1 : sR0, (ch & 1) ? sP1 : sP0, [] {});
        }
    } else {  // nm = 25: one pair-image buffer, copied at the chunk's start (sP1 unused)
        __syncthreads();  // earlier readers of the buffers are done
        dma_units(Rimg, sR0, RI::kUnits, wv, 4, lane);
        for (int ch = 0; ch < nch; ++ch) {
            dma_wait();
            __syncthreads();  // the chunk's rows; the previous chunk's readers are done
            dma_units(Pimg + (size_t)ch * PH::kHalves, sP0, PH::kUnits, wv, 4, lane);
            if (ch + 1 < nch)
                dma_units(Rimg + (size_t)(ch + 1) * RI::kDoubles, (ch & 1) ? sR0 : sR1,
                          RI::kUnits, wv, 4, lane);
            chunk((ch & 1) ? sR1 : sR0, sP0, [] {
                dma_wait();
                __syncthreads();
            });
        }
    }
    const float wRow = row16_max(wmaxOwn);
    float wm = __shfl(wRow, 0);
    if constexpr (NWV == 4) {
        if (lane == 0)
            sRed[wv] = wm;
        __syncthreads();
        wm = fmaxf(fmaxf(sRed[0], sRed[1]), fmaxf(sRed[2], sRed[3]));
    }
    const double es = exp(shift);
    const double sc = pair_scale(es);
    const bool ok = !mark || h16_range_ok(wm);
    if (lane < 16) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int t = wv + NWV * i, p = 16 * t + lane;
            if (t < PH::PT && p < PH::NP)
                sH[p] = (p == 0 && !ok) ? __builtin_nan("") : (double)acc[i][0] * sc;
        }
    }
    if (wv == (NWV == 1 ? 0 : 1)) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const double g0 = __shfl(gl[t], o), g1 = __shfl(gl[t], 16 + o);
            const double g2 = __shfl(gl[t], 32 + o), g3 = __shfl(gl[t], 48 + o);
            if (lane < 16 && 16 * t + lane < NM)
                sG[16 * t + lane] = ((g0 + g1) + (g2 + g3)) * es;
        }
    }
    return wv == 0 ? __shfl(row16_sum(fo), 0) * es : 0.0;
}

constexpr int kCCells = 16;  // cells per wave of the column sweep
// Launch shape of the column sweep: 8 waves per block (two per SIMD) with double-buffered
// pair images up to nm = 16; at nm = 25 (84 KB of pair image per chunk, 84 accumulator
// registers) 4 waves per block, one per SIMD, and a single pair-image buffer
template <int NM>
struct ColShape {
    static constexpr int kWaves = NM <= 16 ? 8 : 4;
    static constexpr int kBlock = kWaves * kWave;
    static constexpr int kPairBufs = NM <= 16 ? 2 : 1;
};

// Sweep pass of the mixed phase (the column sweep): 16 cells per wave, the block's waves
// (ColShape) sharing each chunk; the chunk's row image and pair image staged by LDS DMA, the
// rows double-buffered, the pair image double-buffered (LDS 90 KB at nm = 16) or, at nm = 25,
// copied at the chunk's start and waited for before its pair products (a second barrier;
// LDS 110 KB).  Writes H (packed lower
// triangle, first entry NaN where the weights left the FP16 range), sum_q w e (fdir) and the
// FP64 gradient moments (g slot).
template <int NM>
__global__ __launch_bounds__(ColShape<NM>::kBlock) void k_sweep_cols(KParams P,
                                                                     const double *__restrict__ Rimg,
                                                        const _Float16 *__restrict__ Pimg,
                                                        State S, const int *__restrict__ W,
                                                        const int *__restrict__ nWp)
{
    using PH = PairHess<NM>;
    using RI = RowImg<NM>;
    constexpr int NT = (NM + 15) / 16;
    constexpr int PT = PH::PT;
    constexpr int kCWaves = ColShape<NM>::kWaves;
    constexpr int PB = ColShape<NM>::kPairBufs;
    __shared__ __align__(16) _Float16 sP[PB][PH::kHalves];
    __shared__ __align__(16) double sR[2][RI::kDoubles];
    const int lane = threadIdx.x & 63, wv = uniform_int(threadIdx.x >> 6);
    const int i16 = lane & 15, g = lane >> 4;
    const int nch = P.nqPad / kChunk;
    const int nW = uniform_int(*nWp);
    const int groups = (nW + kCCells - 1) / kCCells;
    const int nBlocksNeeded = (groups + kCWaves - 1) / kCWaves;
    for (int blk = blockIdx.x; blk < nBlocksNeeded; blk += gridDim.x) {
        const int grp = blk * kCWaves + wv;
        const bool any = grp < groups;
        const int slot = grp * kCCells + i16;
        const int cell = (any && slot < nW) ? W[slot] : -1;  // the lane's cell
        const double shift = cell >= 0 ? S.scal[(size_t)cell * kScal + kSshift] : 0.0;
        ExpB<NM> aT;
        aT.load(cell >= 0 ? S.vec + (size_t)cell * kVec * NM + kVa * NM : nullptr, lane);
        float wmaxOwn = 0.0f;
        double fb[4] = {0.0, 0.0, 0.0, 0.0};
        d4 gacc[NT][4];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                gacc[t][b] = d4{0.0, 0.0, 0.0, 0.0};
        f4 acc[PT];
#pragma unroll
        for (int t = 0; t < PT; ++t)
            acc[t] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        __syncthreads();  // the previous cells' last chunk is consumed
        dma_units(Rimg, sR[0], RI::kUnits, wv, kCWaves, lane);
        if constexpr (PB == 2)
            dma_units(Pimg, sP[0], PH::kUnits, (wv + RI::kUnits) % kCWaves, kCWaves, lane);
        for (int ch = 0; ch < nch; ++ch) {
            const int b = ch & 1;
            dma_wait();
            __syncthreads();  // the chunk's rows; every wave is done with the previous chunk
            if constexpr (PB == 1)
                dma_units(Pimg + (size_t)ch * PH::kHalves, sP[0], PH::kUnits,
                          (wv + RI::kUnits) % kCWaves, kCWaves, lane);
            if (ch + 1 < nch) {
                dma_units(Rimg + (size_t)(ch + 1) * RI::kDoubles, sR[b ^ 1], RI::kUnits, wv,
                          kCWaves, lane);
                if constexpr (PB == 2)
                    dma_units(Pimg + (size_t)(ch + 1) * PH::kHalves, sP[b ^ 1], PH::kUnits,
                              (wv + RI::kUnits) % kCWaves, kCWaves, lane);
            }
            if constexpr (PB == 2) {
                if (any)
                    chunk_cols<NM>(sR[b], sR[b] + RI::kW, sP[b], aT, shift, wmaxOwn, fb, gacc,
                                   acc, lane, [] {});
            } else {  // every wave takes the barrier before the pair products
                chunk_cols<NM>(sR[b], sR[b] + RI::kW, sP[0], aT, shift, wmaxOwn, fb, gacc, acc,
                               lane, [] {
                                   dma_wait();
                                   __syncthreads();
                               });
            }
        }
        if (!any)
            continue;
        const double f = col_f(fb);
        const float wm = col_max(wmaxOwn);
        const double es = exp(shift);
        if (g == 0 && cell >= 0)
            S.fdir[cell] = f * es;
        // H: accumulator k of lane l is the cell of column 4 (l >> 4) + k, pair 16 t + (l & 15)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int src = 4 * g + k;
            const int ck = __shfl(cell, src);
            const double sc = pair_scale(__shfl(es, src));
            const bool ok = h16_range_ok(__shfl(wm, src));
            if (ck >= 0) {
                double *H = S.H + (size_t)ck * NM * NM;
#pragma unroll
                for (int t = 0; t < PT; ++t) {
                    const int p = 16 * t + i16;
                    if (p < PH::NP)
                        H[p] = (p == 0 && !ok) ? __builtin_nan("") : (double)acc[t][k] * sc;
                }
            }
        }
        // gradient: register r of lane l is the cell of column (l >> 4) + 4 r, moment
        // 16 t + (l & 15)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int src = g + 4 * r;
            const int cr = __shfl(cell, src);
            const double esr = __shfl(es, src);
            if (cr >= 0) {
                double *gm = S.vec + (size_t)cr * kVec * NM + kVg * NM;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int col = 16 * t + i16;
                    if (col < NM)
                        gm[col] = col_g(gacc[t], r) * esr;
                }
            }
        }
    }
}

// ---------------------------------------------------------------- one cell's state machine
// Register image of a cell's Newton state over a group of G lanes; vectors are distributed
// (group lane i < NM holds component i), scalars are uniform over the group.  The ladder's
// restart point a0, the input alpha and the unscaled moments stay in memory (vb): they are read
// and written only on restarts and rungs, as are the scaled moments uu (read every step).
template <int NM, int G = 64>
struct Cell {
    double a, aOld, d, g, uu;
    double *vb;  // the cell's vector block in the state
    double u0, f, tls, tprev, maxF, gd, err, gamma, r, rtol, shift;
    int iter, tot, igam, reg, flags, phase, maxIt;

    // mixed mode and not yet handed to the FP64 path
    __device__ __forceinline__ bool mixed(const KParams &P) const
    {
        return P.mixed && !(flags & (C3D_FLAG_FP64_FALLBACK | C3D_FLAG_FP64_FINISH));
    }

    __device__ __forceinline__ void load(const State &S, int c, int lane)
    {
        const bool own = lane < NM;
        const double *v = S.vec + (size_t)c * kVec * NM;
        vb = S.vec + (size_t)c * kVec * NM;
        a = own ? v[kVa * NM + lane] : 0.0;
        aOld = own ? v[kVaOld * NM + lane] : 0.0;
        d = own ? v[kVd * NM + lane] : 0.0;
        g = 0.0;  // computed from each sweep before it is used (the g slot carries the sweep's
                  // gradient moments in the mixed phase)
        uu = own ? v[kVu * NM + lane] : 0.0;
        const double *sc = S.scal + (size_t)c * kScal;
        u0 = sc[kSu0];
        f = sc[kSf];
        tls = sc[kStls];
        tprev = sc[kStprev];
        maxF = sc[kSmaxF];
        gd = sc[kSgd];
        err = sc[kSerr];
        gamma = sc[kSgamma];
        r = sc[kSr];
        rtol = sc[kSrtol];
        shift = sc[kSshift];
        const int *in = S.ints + (size_t)c * kInt;
        iter = in[kIiter];
        tot = in[kItot];
        igam = in[kIgam];
        reg = in[kIreg];
        flags = in[kIflags];
        phase = in[kIphase];
        maxIt = in[kImaxIt];
    }

    __device__ __forceinline__ void store(const State &S, int c, int lane) const
    {
        double *v = S.vec + (size_t)c * kVec * NM;
        if (lane < NM) {
            v[kVa * NM + lane] = a;
            v[kVaOld * NM + lane] = aOld;
            v[kVd * NM + lane] = d;
        }
        if (lane == 0) {
            double *sc = S.scal + (size_t)c * kScal;
            sc[kSf] = f;
            sc[kStls] = tls;
            sc[kStprev] = tprev;
            sc[kSmaxF] = maxF;
            sc[kSgd] = gd;
            sc[kSerr] = err;
            sc[kSgamma] = gamma;
            sc[kSr] = r;
            sc[kSrtol] = rtol;
            sc[kSshift] = shift;
            int *in = S.ints + (size_t)c * kInt;
            in[kIiter] = iter;
            in[kItot] = tot;
            in[kIgam] = igam;
            in[kIreg] = reg;
            in[kIflags] = flags;
            in[kIphase] = phase;
            in[kImaxIt] = maxIt;
        }
    }

    // Mixed mode hands the cell to the FP64 path: the state k_init set up, from the input
    // (rtol as for the batched semantics).
    __device__ __forceinline__ void restart_fp64(const KParams &P, int lane)
    {
        const bool own = lane < NM;
        const double in = own ? vb[kVaIn * NM + lane] : 0.0;
        a = in;
        if (own)
            vb[kVa0 * NM + lane] = in;
        aOld = in;
        d = 0.0;
        g = 0.0;
        uu = own ? vb[kVuOrig * NM + lane] : 0.0;
        if (own)
            vb[kVu * NM + lane] = uu;
        f = 0.0;
        tls = 1.0;
        tprev = 1.0;
        maxF = 0.0;
        gd = 0.0;
        err = 0.0;
        gamma = 0.0;
        r = 0.0;
        rtol = P.tol;  // the caller rescales it for the unbatched semantics
        shift = 0.0;
        iter = 0;
        tot = 0;
        igam = 0;
        reg = 0;
        flags = C3D_FLAG_FP64_FALLBACK;
        phase = kPhaseH;
        maxIt = P.maxIter;
    }

    // postProcess (optBatch.cpp:333-344 / opt.cpp:347-353) and the per-cell statistics
    __device__ __forceinline__ void finish(int done, int c, double *u_g, double *alpha_g,
                                           c3d_cell_stats *stats, int *status, int lane)
    {
        if (done == 2)
            flags |= C3D_FLAG_LADDER;
        const double alphaScale = det_log(u0) * 2.0 * sqrt(M_PI);
        if (lane < NM) {
            u_g[(size_t)c * NM + lane] = uu * u0;
            alpha_g[(size_t)c * NM + lane] = (lane == 0) ? a + alphaScale : a;
        }
        if (lane == 0) {
            if (done == 2)
                atomicOr(status, C3D_FLAG_LADDER);
            if (stats) {
                c3d_cell_stats s0{r, err, gamma, tot, igam, reg, flags};
                stats[c] = s0;
            }
        }
    }

    // Armijo test of a line-search trial at step t and point at (this lane's component),
    // f(alpha) = fdir - alpha.u (linesearch.cpp:60-66); alpha.u in order (EXACT) or as a tree
    template <bool EXACT>
    __device__ __forceinline__ bool accepts_at(double fdir, double at, double t) const
    {
        double fNew = fdir;
        const double au = at * uu;  // each lane its own product
        if constexpr (EXACT) {
#pragma unroll
            for (int i = 0; i < NM; ++i)
                fNew = fNew - Grp<G>::bc(au, i);
        } else {
            double v[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i)
                v[i] = Grp<G>::bc(au, i);
            fNew = fNew - tree_sum(v);
        }
        return fNew <= f + 1e-3 * t * gd;
    }
    // ... of the current trial
    template <bool EXACT>
    __device__ __forceinline__ bool ls_accepts(double fdir) const
    {
        return accepts_at<EXACT>(fdir, a, tls);
    }

    // The next trial of the backtracking search at step tls, or its end (linesearch.cpp:46-54):
    // alpha restored and tprev = 0, which the next iteration turns into a ladder step.
    __device__ __forceinline__ void ls_next(int lane)
    {
        if (tls * maxF > DBL_EPSILON / 2) {
            a = lane < NM ? aOld + tls * d : 0.0;
            phase = kPhaseLS;
        } else {
            a = aOld;
            tprev = 0.0;
            phase = kPhaseH;
        }
    }

    // A rejected line-search trial: halve the step (linesearch.cpp:75-78).
    __device__ __forceinline__ void ls_reject(int lane)
    {
        tls = 0.5 * tls;
        ls_next(lane);
    }

    // One step of the state machine with the sweep result at the current alpha (H row-major,
    // in LDS or global memory; fdir = sum_q w_q e_q): either the Armijo test of a line-search
    // trial, or one Newton iteration of solveTheRest / optScaled.  sT: NM x (NM+1) doubles of
    // LDS for the transposed factor.  Returns 0 to continue, 1 converged, 2 ladder exhausted.
    // H is not read when a line-search trial is rejected (it may then be null).
    template <bool EXACT>
    __device__ __forceinline__ int advance(const KParams &P, const double *H, double fdir,
                                           const double *gm, double *sT, int lane)
    {
        constexpr int LD = NM + 1;
        const bool own = lane < NM;
        const double tolGamma = P.tolGamma;  // opt.cpp:149 / optBatch.cpp:268
        if (phase == kPhaseLS) {
            if (ls_accepts<EXACT>(fdir)) {
                tprev = tls;  // accepted; H of this sweep belongs to the new alpha
            } else {
                ls_reject(lane);
                return 0;
            }
        }
        // the mixed sweep marks an H whose scaled weights left FP16's range (a point far from
        // the shift's iterate): re-centre the shift at this point and sweep it again (once: an
        // unchanged shift falls through, and the marked H fails the factorisation)
        if (mixed(P) && !(H[0] == H[0])) {
            const double lf = log(fdir);
            const double ns = (lf > -700.0 && lf < 700.0) ? lf : 0.0;
            if (ns != shift) {
                shift = ns;
                phase = kPhaseH;
                return 0;
            }
        }
        // ================================ one Newton iteration (solveTheRest)
        const bool mx = mixed(P);
        Dense<NM, EXACT, G> D;
        if (mx) {  // the mixed sweep's packed lower triangle (all the mixed phase reads)
#pragma unroll
            for (int j = 0; j < NM; ++j)
                D.row[j] = (own && j <= lane) ? H[tri(lane) + j] : 0.0;
        } else {
#pragma unroll
            for (int j = 0; j < NM; ++j)
                D.row[j] = own ? H[lane * NM + j] : 0.0;
        }
        int maxItNow = maxIt;
        if (!P.unbatched && reg == kNumReg - 1)
            maxItNow = 10 * maxIt;
        ++iter;
        ++tot;
        // fobjMnUseH; the mixed phase has the FP64 gradient moments of its sweep
        g = own ? (mx ? gm[lane] : D.row[0] / P.m0) - uu : 0.0;
        if (mx)  // f from the FP64 sum, as the line search evaluates it (fobj.cpp:135-164)
            f = fdir;
        else
            f = Grp<G>::bc(D.row[0], 0) / (P.m0 * P.m0);
        const double au = a * uu;
        if constexpr (EXACT) {
#pragma unroll
            for (int i = 0; i < NM; ++i)
                f = f - Grp<G>::bc(au, i);
        } else {
            double t[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i)
                t[i] = Grp<G>::bc(au, i);
            f = f - tree_sum(t);
        }
        double norm1 = 0.0;  // utils_norm1: the largest absolute row sum
        if (!mx) {  // for dpocon (the FP64 path)
            if constexpr (EXACT) {
                double rs = 0.0;
#pragma unroll
                for (int j = 0; j < NM; ++j)
                    rs = rs + fabs(D.row[j]);
#pragma unroll
                for (int i = 0; i < NM; ++i) {
                    double vv = Grp<G>::bc(rs, i);
                    if (vv > norm1)
                        norm1 = vv;
                }
            } else {
                double t[NM];
#pragma unroll
                for (int j = 0; j < NM; ++j)
                    t[j] = fabs(D.row[j]);
                const double rs = tree_sum(t);
#pragma unroll
                for (int i = 0; i < NM; ++i)
                    t[i] = Grp<G>::bc(rs, i);
                norm1 = fmax(tree_max(t), 0.0);
            }
        }
        const int cholErr = D.cholesky(lane);
        double condest = 0.0;
        if (cholErr == 0) {
            D.build_columns(sT, LD, lane);
            // the reference's dpocon decides the FP64 path; the mixed phase's hand-over test
            // only needs an indicator: (max L_ii / min L_ii)^2, a lower bound on cond_2(H)
            condest = mx ? D.diag_ratio_sq(lane) : 1.0 / D.pocon(norm1, lane);
        }
        bool restart = false;  // the iteration starts over; rtol follows the (new) u
        if (mx) {
            const bool stalled = tprev == 0.0 || iter >= min(maxItNow, P.mixedMaxIter);
            const bool illCond = cholErr != 0 || !(condest <= P.mixedCondMax);
            if ((stalled || illCond) && tot > 1 && err <= P.mixedSwitch) {
                // stalled, or its split-FP16 Hessian too ill-conditioned to factor, near its
                // optimum (err: the last iteration's ||g||, FP64): FP64 Hessians continue from
                // this iterate, which the next (FP64) sweep evaluates
                flags |= C3D_FLAG_FP64_FINISH;
                tprev = 1.0;
                phase = kPhaseH;
                --iter;
                --tot;
                return 0;
            }
            if (illCond || stalled) {
                restart_fp64(P, lane);
                restart = true;
            } else {
                const double lf = log(fdir);  // shift of the sweeps from this iterate on
                shift = (lf > -700.0 && lf < 700.0) ? lf : 0.0;
            }
        } else if (tot == 1 && (cholErr != 0 || condest > P.condHMax)) {
            // isotropic restart, optBatch.cpp:208-218
            const double ai = (lane == 0) ? P.isoAlpha0 : 0.0;
            if (own)
                vb[kVa0 * NM + lane] = ai;
            a = ai;
            iter = 0;
            igam = 0;
            tprev = 1.0;
            flags |= C3D_FLAG_ISO_RESTART;
            phase = kPhaseH;
            restart = true;
        } else if (iter >= maxItNow || cholErr != 0 || tprev == 0.0 || condest > P.condHMax) {
            if (reg == kNumReg - 1)  // ladder, optBatch.cpp:222-248
                return 2;
            ++reg;
            r = kReg[reg];
            uu = (lane == 0) ? 1.0 : (own ? vb[kVuOrig * NM + lane] * (1.0 - r) : 0.0);
            if (own)
                vb[kVu * NM + lane] = uu;
            a = own ? vb[kVa0 * NM + lane] : 0.0;
            igam = 0;
            iter = 0;
            tprev = 1.0;
            if (P.unbatched && reg == kNumReg - 1)
                maxIt = 10 * maxIt;
            phase = kPhaseH;
            restart = true;
        }
        if (restart) {
            if (P.unbatched)  // opt.cpp:195, 240, 270
                rtol = LaneVec<NM, G>::dnrm2(uu) * P.tol;
            return 0;
        }
        d = g;
        if (P.unbatched)
            D.potrs_lapack(d, lane);
        else
            D.potrs_custom(d, lane);
        d = own ? -d : 0.0;
        err = LaneVec<NM, G, EXACT>::dnrm2(g);
        const double da1 = LaneVec<NM, G, EXACT>::dasum(d);
        // the deterministic exp in every mode: the device library's exp, inlined here, pushed
        // the step kernel past 256 registers (24 spilled)
        gamma = det_exp(2.0 * da1 * P.minf);
        // the stopping test on the FP64 gradient, in both phases (optBatch.cpp:268)
        if (err <= rtol && gamma <= tolGamma)
            return 1;
        if (err <= rtol && gamma > tolGamma)
            ++igam;
        // an ill-conditioned cell near its optimum finishes with FP64 Hessians (the split
        // FP16 products carry about 22 bits, so Newton there contracts by ~cond * 1e-7)
        if (mx && err <= P.mixedSwitch && condest > P.mixedFinishCond)
            flags |= C3D_FLAG_FP64_FINISH;
        aOld = a;  // linesearch.cpp:29-54
        maxF = 0.0;
        const double ratio = fabs(d) / (fabs(aOld) + 1e-100);  // each lane its own component
        const double gdl = g * d;
        if constexpr (EXACT) {
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double temp = Grp<G>::bc(ratio, i);
                if (temp > maxF)
                    maxF = temp;
            }
            gd = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i)
                gd = gd + Grp<G>::bc(gdl, i);
        } else {
            double t[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i)
                t[i] = Grp<G>::bc(ratio, i);
            maxF = fmax(tree_max(t), 0.0);
#pragma unroll
            for (int i = 0; i < NM; ++i)
                t[i] = Grp<G>::bc(gdl, i);
            gd = tree_sum(t);
        }
        tls = 1.0;
        ls_next(lane);
        return 0;
    }
};

// ---------------------------------------------------------------- step pass
// One state-machine step per listed cell.  Orders up to M_3 (NM <= 16) take four cells per
// wave, one per 16-lane row, so that the serial Cholesky / solve / condition-estimate chains
// of four cells run side by side; larger orders take one cell per wave.  Two waves per SIMD
// (at most 256 registers a lane) measured faster than one wave with no spills.
template <int NM, bool EXACT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void k_step(KParams P, State S, Lists Lc, Lists Ln,
                                                  double *u_g, double *alpha_g,
                                                  c3d_cell_stats *stats, int *status,
                                                  unsigned long long *sweepCount)
{
    constexpr int G = NM <= 16 ? 16 : kWave;
    constexpr int kCellsPerBlock = kBlock / G;
    constexpr int LD = NM + 1;
    __shared__ double sT[kCellsPerBlock][NM * LD];
    const int lane = threadIdx.x & (G - 1), grp = threadIdx.x / G;
    const int nWm = uniform_int(*Lc.nm);
    const int nW = nWm + uniform_int(*Lc.nf);
    if (blockIdx.x == 0 && threadIdx.x == 0 && sweepCount) {  // [0] FP64 sweeps, [2] mixed
        atomicAdd(sweepCount, (unsigned long long)(nW - nWm));
        atomicAdd(sweepCount + 2, (unsigned long long)nWm);
    }
    C3D_BLOCK_APPENDER(app);
    const int nBlocksNeeded = (nW + kCellsPerBlock - 1) / kCellsPerBlock;
    for (int blk = blockIdx.x; blk < nBlocksNeeded; blk += gridDim.x) {
        const int slot = blk * kCellsPerBlock + grp;
        if (slot < nW) {
            int c = list_cell(Lc, slot, nWm);
            if constexpr (G == kWave)
                c = uniform_int(c);
            Cell<NM, G> C;
            C.load(S, c, lane);
            const int done =
                C.template advance<EXACT>(P, S.H + (size_t)c * NM * NM, S.fdir[c],
                                          S.vec + (size_t)c * kVec * NM + kVg * NM, sT[grp], lane);
            if (done) {
                C.finish(done, c, u_g, alpha_g, stats, status, lane);
            } else {
                C.store(S, c, lane);
                if (lane == 0)
                    app.push(C.mixed(P), c);
            }
        }
        app.flush(Ln, blk + gridDim.x >= nBlocksNeeded, kCellsPerBlock);
    }
}

// f alone for one cell at kw <= KW trial points (one wave; the points in LDS, KW x NM), for
// line-search trials in the tail kernel: sum_q w_q exp(alpha.m_q) with exactly the
// operations the sweep kernels use for that sum, so that a cell's result does not depend on
// which kernel evaluated it -- kFp64: chunk_mfma's exponent, per-lane partials over the
// chunks and final wave_sum; kFRef: chunk_ref's deterministic exp and ordinate-order sum
// (one point).  (The mixed phase's points: f_trials_mixed.)  M is read from the transposed
// table MT [NM][nqPad] (one coalesced load per component), two chunks in flight, each chunk's
// M shared by the kw points.
enum : int { kFp64 = 0, kFRef = 1 };
template <int NM, int MODE, int KW>
__device__ __forceinline__ void f_trials(const KParams &P, const double *__restrict__ MT,
                                         const double *__restrict__ wg, const double *sA,
                                         int kw, double *sW, double *sE, int lane,
                                         double (&out)[KW])
{
    static_assert(MODE != kFRef || KW == 1, "reference order: one point per wave");
    const int nch = P.nqPad / kChunk;
    const size_t ld = (size_t)P.nqPad;
    double mA[NM], mB[NM], wA = 0.0, wB = 0.0;
    // kFp64 evaluates ordinate mfma_ordinate(lane) of each chunk, as chunk_mfma does
    const int qo = MODE == kFp64 ? mfma_ordinate(lane) : lane;
    auto load = [&](double(&m)[NM], double &w, int ch) {
        const size_t q = (size_t)ch * kChunk + qo;
#pragma unroll
        for (int i = 0; i < NM; ++i)
            m[i] = MT[i * ld + q];
        w = wg[q];
    };
    double fpart[KW];
#pragma unroll
    for (int j = 0; j < KW; ++j)
        fpart[j] = 0.0;
    auto term = [&](const double(&m)[NM], double w, int ch) {
#pragma unroll
        for (int j = 0; j < KW; ++j) {
            if (j >= kw)
                continue;
            const double *a = sA + j * NM;
            double x = 0.0;
            {
                if constexpr (MODE == kFRef) {
#pragma unroll
                    for (int i = 0; i < NM; ++i)
                        x = x + a[i] * m[i];
                } else {
#pragma unroll
                    for (int i = 0; i < NM; ++i)
                        x = fma(a[i], m[i], x);
                }
                if constexpr (MODE == kFRef) {
                    wave_lds_sync();
                    sE[lane] = det_exp(x);
                    sW[lane] = w;
                    wave_lds_sync();
                    const int nqc = min(kChunk, P.nq - ch * kChunk);
                    for (int k = 0; k < nqc; ++k)
                        fpart[j] = fpart[j] + sW[k] * sE[k];
                } else {
                    fpart[j] = fpart[j] + w * exp(x);
                }
            }
        }
    };
    load(mA, wA, 0);
    if (nch > 1)
        load(mB, wB, 1);
    for (int ch = 0; ch < nch; ch += 2) {
        term(mA, wA, ch);
        if (ch + 2 < nch)
            load(mA, wA, ch + 2);
        if (ch + 1 < nch) {
            term(mB, wB, ch + 1);
            if (ch + 3 < nch)
                load(mB, wB, ch + 3);
        }
    }
#pragma unroll
    for (int j = 0; j < KW; ++j) {
        if (j >= kw)
            continue;
        if constexpr (MODE == kFRef)
            out[j] = fpart[j];
        else
            out[j] = wave_sum(fpart[j]);
    }
}

// f at up to four trial points in the mixed phase, as the column sweep evaluates it: the points are
// the rows of the exponent products (alpha of point j at sA + j * NM, j < kw), lane L takes
// point L >> 4 at ordinates 16 g + (L & 15) of each chunk and sums its weights in chunk and
// group order, then the 16-lane row sum and the e^shift rescale.  M is read from its
// transposed copy MT[m][q] (leading dimension nqPad), two chunks in flight (one at nm = 25).
template <int NM>
__device__ __forceinline__ void f_trials_mixed(const KParams &P, const double *__restrict__ MT,
                                               const double *__restrict__ wg, const double *sA,
                                               int kw, double shift, int lane, double (&out)[4],
                                               const double *xtab)
{
    constexpr int NKS = ExpA<NM>::NKS;
    const int nch = P.nqPad / kChunk;
    const size_t ld = (size_t)P.nqPad;
    const int r16 = lane & 15, kk = lane >> 4;
    ExpA<NM> ax;
    ax.load(sA, kw, lane);
    auto load = [&](double(&b)[4][NKS], double(&w)[4], int ch) {
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const size_t q = (size_t)ch * kChunk + 16 * g + r16;
#pragma unroll
            for (int t = 0; t < NKS; ++t) {
                const int m = 4 * t + kk;
                b[g][t] = (NM % 4 == 0 || m < NM) ? MT[m * ld + q] : 0.0;
            }
            w[g] = wg[q];
        }
    };
    double fOwn = 0.0;
    auto term = [&](const double(&b)[4][NKS], const double(&w)[4]) {
        double xg[4];
        exp_arg4<NM>(ax, [&](int g, int t) { return b[g][t]; }, xg);
#pragma unroll
        for (int g = 0; g < 4; ++g)
            fOwn = fOwn + w[g] * mexp(xg[g] - shift, xtab);
    };
    if constexpr (NKS <= 4) {
        // two chunks at a time: their eight exponent chains and exps are independent work for
        // one another's latencies; the lane's sum still takes chunk A's groups, then B's
        double bA[4][NKS], bB[4][NKS], wA[4], wB[4];
        load(bA, wA, 0);
        if (nch > 1)
            load(bB, wB, 1);
        for (int ch = 0; ch < nch; ch += 2) {
            if (ch + 1 < nch) {
                double xa[4], xb[4];
                exp_arg4<NM>(ax, [&](int g, int t) { return bA[g][t]; }, xa);
                exp_arg4<NM>(ax, [&](int g, int t) { return bB[g][t]; }, xb);
                double ea[4], eb[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    ea[g] = wA[g] * mexp(xa[g] - shift, xtab);
                    eb[g] = wB[g] * mexp(xb[g] - shift, xtab);
                }
                if (ch + 2 < nch)
                    load(bA, wA, ch + 2);
                if (ch + 3 < nch)
                    load(bB, wB, ch + 3);
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    fOwn = fOwn + ea[g];
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    fOwn = fOwn + eb[g];
            } else {
                term(bA, wA);
            }
        }
    } else {  // nm = 25: one chunk in flight (registers)
        double bA[4][NKS], wA[4];
        for (int ch = 0; ch < nch; ++ch) {
            load(bA, wA, ch);
            term(bA, wA);
        }
    }
    const double fRow = row16_sum(fOwn);
    const double es = exp(shift);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        out[j] = __shfl(fRow, 16 * j) * es;
}

// ---------------------------------------------------------------- fused tail pass
// For the last few cells: one block per cell runs up to `passes` rounds in one launch, with
// the cell's state in registers and its trial points, M chunks, H and the transposed factor
// in LDS.  The sweep performs the same operations in the same order as k_sweep /
// k_sweep_ref / k_sweep_cols do for that cell, so a cell's result does not depend on which
// kernel finished it.  A line-search trial needs f alone (H only once a trial is accepted).
// The TW waves of a block evaluate f at the next trial points of the backtracking search
// (tls, tls/2, tls/4, ...: one per wave, or KW per wave once the search has backtracked), so
// one round settles up to TW * KW trials of the sequential search.  Every wave keeps an
// identical copy of the state machine (same inputs, same operations), which keeps control
// flow uniform over the block; wave 0 sweeps H and writes the results.  Launch: TW * 64
// threads per block; the blocks stride over the listed cells (any grid).
template <int NM, bool REF, bool MIXED, int TW>
__global__ __launch_bounds__(TW * kWave) void k_tail(KParams P, const double *__restrict__ Mg,
                                                      const double *__restrict__ MT,
                                                      const double *__restrict__ wg,
                                                      const _Float16 *__restrict__ Pimg,
                                                      const double *__restrict__ Rimg, State S,
                                                      Lists Lc, Lists Ln, double *u_g,
                                                      double *alpha_g, c3d_cell_stats *stats,
                                                      int *status, int passes,
                                                      unsigned long long *sweepCount)
{
    constexpr int NT = (NM + 15) / 16;
    constexpr int NB = NT * (NT + 1) / 2;
    constexpr int NMS = NM | 1;
    // trial points per wave in a wide round: as many as keep a round's latency near one f
    // evaluation's (f costs nq (NM + exp) per point)
    constexpr int KW = REF ? 1 : (NM <= 9 ? 8 : (NM <= 16 ? 2 : 1));
    // in the mixed phase the points are the rows of the exponent products: four per wave
    constexpr int KWM = 4;
    constexpr int KWA = MIXED && KWM > KW ? KWM : KW;
    static_assert(!(REF && MIXED), "one arithmetic mode");
    // the column sweep's images (nm <= kPairMaxNm, mixed): double-buffered with four waves,
    // else single; the first rows buffer also serves as the other paths' chunk rows (sM)
    constexpr bool COLS = MIXED && NM <= kPairMaxNm;
    constexpr bool COLS4 = COLS && TW == 4;
    // FP64 Newton sweeps of nm > 16 with four waves: H's accumulator blocks over the waves,
    // the chunks staged through both row buffers
    constexpr bool SPLIT = !REF && TW == 4 && NB > 1;
    __shared__ __align__(16) double sR[COLS4 || SPLIT ? 2 : 1][COLS || SPLIT ? RowImg<NM>::kDoubles : 1];
    __shared__ __align__(16) _Float16 sP[COLS4 && NM <= 16 ? 2 : 1][COLS ? PairHess<NM>::kHalves : 8];
    __shared__ double sMo[COLS || SPLIT ? 1 : kChunk * NMS];
    double *const sM = COLS || SPLIT ? &sR[0][0] : sMo;
    __shared__ double sH[NM * NM];
    // the state machine runs on 16-lane rows for NM <= 16 (DPP broadcasts; the four rows of a
    // wave hold identical copies), else on the whole wave
    constexpr int G = NM <= 16 ? 16 : kWave;
    // the rows of a wave run identical copies of the state machine and write identical values:
    // one factor workspace per wave serves them all
    __shared__ double sT[TW][NM * (NM + 1)];
    __shared__ double sW[TW][kChunk];
    __shared__ double sE[TW][kChunk];
    __shared__ double sA[TW][KWA * NM];
    __shared__ int sOk[TW * KWA];       // Armijo verdict at each trial point
    __shared__ double sF[1];            // f of the sweep
    __shared__ double sWEl[COLS ? kChunk : 1];  // the lone-cell sweep's weights
    __shared__ double sG[NM];
    __shared__ float sRed[4];
    __shared__ double sX[MIXED ? 64 : 1];  // the mixed weights' exp table
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int nWm = uniform_int(*Lc.nm);
    const int nW = nWm + uniform_int(*Lc.nf);
    if ((int)blockIdx.x >= nW)
        return;
    if constexpr (MIXED) {
        for (int e = threadIdx.x; e < 64; e += TW * kWave)
            sX[e] = kExp2Tab[e];
        __syncthreads();
    }
    const int li = lane & (G - 1), row = lane / G;
    for (int slot = blockIdx.x; slot < nW; slot += gridDim.x) {
        const int c = uniform_int(list_cell(Lc, slot, nWm));
        Cell<NM, G> C;
        C.load(S, c, li);
        const int nch = P.nqPad / kChunk;
        int done = 0;
        int sweeps = 0;
        for (int p = 0; p < passes; ++p) {
            const bool mx = MIXED && C.mixed(P);
            if (C.phase == kPhaseLS) {
                // points k = wv * kw + j at tls / 2^k, formed as ls_next() forms them; the
                // mixed phase's points are the rows of one exponent product, so its four cost
                // what one does and are taken from the first trial on
                const int kw = mx ? KWM : (C.tls < 1.0 ? KW : 1);
                double tk = C.tls;
                for (int k = 0; k < wv * kw; ++k)
                    tk = 0.5 * tk;
                auto point = [&](int j, double t) {
                    return (wv == 0 && j == 0) ? C.a : (li < NM ? C.aOld + t * C.d : 0.0);
                };
                {
                    double aj[KWA], tj[KWA];
#pragma unroll
                    for (int j = 0; j < KWA; ++j) {
                        tj[j] = tk;
                        aj[j] = point(j, tk);
                        if (j < kw && lane < NM)
                            sA[wv][j * NM + lane] = aj[j];
                        tk = 0.5 * tk;
                    }
                    wave_lds_sync();
                    double ft[KWA];
                    if (mx) {  // the mixed sweeps' f: the same bits as the sweep's
                        double f4p[4];
                        f_trials_mixed<NM>(P, MT, wg, sA[wv], kw, C.shift, lane, f4p, sX);
#pragma unroll
                        for (int j = 0; j < KWA; ++j)
                            ft[j] = j < 4 ? f4p[j] : 0.0;
                    } else {
                        double fk[KW];
                        f_trials<NM, REF ? kFRef : kFp64, KW>(P, MT, wg, sA[wv], kw, sW[wv],
                                                              sE[wv], lane, fk);
#pragma unroll
                        for (int j = 0; j < KWA; ++j)
                            ft[j] = j < KW ? fk[j] : 0.0;
                    }
                    // each wave judges its own points; the search then walks the verdicts in
                    // order
#pragma unroll
                    for (int j = 0; j < KWA; ++j) {
                        if (j < kw) {
                            const bool ok = C.template accepts_at<REF>(ft[j], aj[j], tj[j]);
                            if (lane == 0)
                                sOk[wv * kw + j] = ok;
                        }
                    }
                }
                __syncthreads();
                bool accepted = false;
                for (int k = 0; k < TW * kw; ++k) {
                    if (sOk[k]) {
                        accepted = true;
                        break;
                    }
                    C.ls_reject(li);  // halve the step, or give up
                    if (C.phase != kPhaseLS)
                        break;
                }
                __syncthreads();  // sOk and sA are rewritten by the next round
                if (!accepted)
                    continue;
            }
            bool swept = false;
            if constexpr (MIXED && TW == 4) {
                static_assert(TW * kWave == kBlock, "the block sweep stages with kBlock threads");
                if (mx) {  // the four waves share the sweep (same bits as one wave's)
                    if (wv == 0 && lane < NM)
                        sA[0][lane] = C.a;
                    __syncthreads();
                    const double fdir = sweep_lone<NM, 4>(
                        P.nqPad, Rimg, Pimg, sA[0], C.shift, sR[0], sR[COLS4 ? 1 : 0], sP[0],
                        sP[COLS4 && NM <= 16 ? 1 : 0], sWEl, sRed, sH, sG, true, lane, wv, sX);
                    if (wv == 0) {
                        if (lane == 0)
                            sF[0] = fdir;
                        ++sweeps;
                    }
                    swept = true;
                }
            }
            if constexpr (SPLIT) {
                if (!mx && !swept) {  // the FP64 sweep, H's accumulator blocks over the waves
                    if (wv == 0 && lane < NM)
                        sA[0][lane] = C.a;
                    __syncthreads();
                    const double fdir =
                        sweep_fp64_blocks<NM>(P.nqPad, Rimg, sA[0], sR[0], sR[1], sH, lane, wv);
                    if (wv == 0) {
                        if (lane == 0)
                            sF[0] = fdir;
                        ++sweeps;
                    }
                    swept = true;
                }
            }
            if (wv == 0 && !swept) {
                double fdir;
                if (lane < NM)
                    sA[0][lane] = C.a;
                if (mx) {  // the mixed phase's lone-cell sweep
                    wave_lds_sync();  // sA[0]
                    fdir = sweep_lone<NM, 1>(P.nqPad, Rimg, Pimg, sA[0], C.shift, sR[0], sR[0],
                                             sP[0], sP[0], sWEl, sRed, sH, sG, true, lane, 0, sX);
                } else if constexpr (REF) {
                    const RefEntries<NM> E(lane);
                    double h[RefEntries<NM>::NS];
#pragma unroll
                    for (int s2 = 0; s2 < RefEntries<NM>::NS; ++s2)
                        h[s2] = 0.0;
                    double f = 0.0;
                    for (int ch = 0; ch < nch; ++ch) {
                        wave_lds_sync();
                        for (int e = lane; e < kChunk * NM; e += kWave) {
                            const int rr = e / NM, cc = e - rr * NM;
                            sM[rr * NMS + cc] = Mg[(size_t)ch * kChunk * NM + e];
                        }
                        sW[0][lane] = wg[(size_t)ch * kChunk + lane];
                        wave_lds_sync();
                        chunk_ref<NM>(sM, sW[0], sE[0], sA[0], E, h, f,
                                      min(kChunk, P.nq - ch * kChunk), lane);
                    }
                    wave_lds_sync();
                    store_H_ref<NM>(E, h, sH, lane);
                    fdir = f;
                } else {
                    d4 acc[1][NB];
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        acc[0][b] = d4{0.0, 0.0, 0.0, 0.0};
                    double fpart[1] = {0.0};
                    double pre[NM];
                    const int qo = mfma_ordinate(lane);  // the lane's weight
                    double prew = wg[qo];
#pragma unroll
                    for (int k = 0; k < NM; ++k)
                        pre[k] = Mg[lane + kWave * k];
                    for (int ch = 0; ch < nch; ++ch) {
                        wave_lds_sync();
#pragma unroll
                        for (int k = 0; k < NM; ++k) {
                            const int e = lane + kWave * k;
                            const int rr = e / NM, cc = e - rr * NM;
                            sM[rr * NMS + cc] = pre[k];
                        }
                        const double wl = prew;
                        wave_lds_sync();
                        if (ch + 1 < nch) {
                            const size_t base = (size_t)(ch + 1) * kChunk;
                            prew = wg[base + qo];
#pragma unroll
                            for (int k = 0; k < NM; ++k)
                                pre[k] = Mg[base * NM + lane + kWave * k];
                        }
                        chunk_mfma<NM, 1>(sM, wl, sA[0], acc, fpart, lane);
                    }
                    wave_lds_sync();
                    fdir = wave_sum(fpart[0]);
                    store_H<NM>(acc[0], sH, lane);
                }
                if (lane == 0)
                    sF[0] = fdir;
                ++sweeps;
            }
            __syncthreads();
            done = C.template advance<REF>(P, sH, sF[0], MIXED ? sG : nullptr, sT[wv], li);
            __syncthreads();  // sH and sF are rewritten by the next round
            if (done)
                break;
        }
        if (wv == 0 && row == 0) {
            if (lane == 0 && sweepCount)
                atomicAdd(sweepCount, (unsigned long long)sweeps);
            if (done) {
                C.finish(done, c, u_g, alpha_g, stats, status, li);
            } else {
                C.store(S, c, li);
                if (lane == 0)
                    list_append(Ln, C.mixed(P), c);
            }
        }
        __syncthreads();  // the block's LDS serves the next cell
    }
}

// ---------------------------------------------------------------- learned warm start
// Features of the listed cells for the network, row = worklist slot: x_k = (u_k/u0 - xm_k)/
// xs_k from the state's u0-scaled moments (k < NM; zero beyond and for slots >= *nWp).
template <int NM>
__global__ void k_warm_features(int64_t rows, State S, Lists L, const float *__restrict__ norm,
                                float *X)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * C3D_MLP_IN)
        return;
    const int64_t slot = e / C3D_MLP_IN;
    const int k = (int)(e % C3D_MLP_IN);
    float x = 0.0f;
    const int nWm = *L.nm;
    if (slot < nWm + *L.nf && k < NM) {
        const double uk = S.vec[(size_t)list_cell(L, (int)slot, nWm) * kVec * NM + kVu * NM + k];
        x = ((float)uk - norm[k]) / norm[C3D_MLP_IN + k];
    }
    X[e] = x;
}

// The gradient-decrease test: for each listed cell, ||grad h||_2 at the caller's alpha and at
// the network's proposal anet[slot] (sums in FP32: the test only ranks two candidates, and a
// proposal whose exponentials overflow loses it); the proposal becomes the starting point
// (alpha, alpha0 and the line-search anchor) only if its gradient is smaller.  Either way the
// mixed mode's first shift is log sum_q w e at the chosen point.  Eight cells per wave,
// sixteen candidate points: each lane forms the weights w_q e^{x} of its ordinate for the
// sixteen points (VALU), and the gradients sum_q w_q e_q m_q accumulate on the FP32 matrix
// cores, point v in row v of a 16x16 tile.
template <int NM>
__global__ __launch_bounds__(kBlock) void k_warm_select(KParams P, const double *__restrict__ Mg,
                                                         const double *__restrict__ wg, State S,
                                                         Lists L, const double *__restrict__ anet)
{
    constexpr int NT = (NM + 15) / 16;
    constexpr int NMS = NM | 1;
    constexpr int CW = 8;  // 16 candidate points fill the rows of one 16x16 tile
    constexpr int NV = 2 * CW;  // point v = 2 c + (0: the caller's alpha, 1: the proposal)
    constexpr int RS = kRS;
    constexpr int kPer = (kChunk * NM + kBlock - 1) / kBlock;
    static_assert(NV <= 16, "points fill the rows of one tile");
    __shared__ float sMr[2][kChunk * NMS];
    __shared__ __align__(16) float sMfT[2][MixedChunk<NM>::kFloats];
    __shared__ float sW[2][kChunk];
    __shared__ float sA[kWavesPerBlock][NV * NM];
    // weights by point, rows kChunk + 4 floats apart: the A operand's 16 rows fall in distinct
    // banks
    constexpr int kWES = kChunk + 4;
    __shared__ __align__(16) float sWE[kWavesPerBlock][NV * kWES];
    __shared__ float sN[kWavesPerBlock][NV];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int gq = lane >> 4, rq = lane & 15;
    const int nch = P.nqPad / kChunk;
    const int nWm = uniform_int(*L.nm);
    const int nW = nWm + uniform_int(*L.nf);
    const int groups = (nW + CW - 1) / CW;
    const int nBlocksNeeded = (groups + kWavesPerBlock - 1) / kWavesPerBlock;
    MixedChunk<NM>::clear_pad(sMfT[0], tid, kBlock);
    MixedChunk<NM>::clear_pad(sMfT[1], tid, kBlock);
    for (int blk = blockIdx.x; blk < nBlocksNeeded; blk += gridDim.x) {
        const int grp = blk * kWavesPerBlock + wv;
        int slots[CW], cells[CW];
        // iso bit c: the caller's alpha of cell c is isotropic (alpha_j = 0 for j >= 1).  Its
        // gradient there is -u_j for j >= 1 (the basis is orthogonal to the constant m_0 under
        // the quadrature) and e^{alpha_0 m_0} sum_q w_q m_0 - u_0 = e^{alpha_0 m_0} 4 pi m_0 - u_0
        // for j = 0, so that point's sums are not evaluated.
        unsigned iso = 0u;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            slots[c] = grp * CW + c;
            cells[c] = (grp < groups && slots[c] < nW) ? list_cell(L, slots[c], nWm) : -1;
            const bool ok = cells[c] >= 0;
            double ac = 0.0;
            if (lane < NM) {
                ac = ok ? S.vec[(size_t)cells[c] * kVec * NM + kVa * NM + lane] : 0.0;
                sA[wv][(2 * c) * NM + lane] = (float)ac;
                sA[wv][(2 * c + 1) * NM + lane] =
                    ok ? (float)anet[(size_t)slots[c] * NM + lane] : 0.0f;
            }
            if (__all(lane < 1 || ac == 0.0) && ok)
                iso |= 1u << c;
        }
        f4 acc[NT];
#pragma unroll
        for (int b = 0; b < NT; ++b)
            acc[b] = f4{0.0f, 0.0f, 0.0f, 0.0f};
        float fpart[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v)
            fpart[v] = 0.0f;
        const bool any = cells[0] >= 0;
        // chunk staging: the next chunk in registers, converted to FP32 when stored
        double pre[kPer], prew;
        auto fetch = [&](int ch) {
            const double *src = Mg + (size_t)ch * kChunk * NM;
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int e = tid + kBlock * k;
                pre[k] = (e < kChunk * NM) ? src[e] : 0.0;
            }
            prew = tid < kChunk ? wg[(size_t)ch * kChunk + tid] : 0.0;
        };
        fetch(0);
        __syncthreads();  // the previous cells' last chunk is consumed
        for (int ch = 0; ch < nch; ++ch) {
            const int b = ch & 1;
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int e = tid + kBlock * k;
                if (e < kChunk * NM) {
                    const int rr = e / NM, cc = e - rr * NM;
                    sMr[b][rr * NMS + cc] = (float)pre[k];
                    sMfT[b][MixedChunk<NM>::index(rr, cc)] = (float)pre[k];
                }
            }
            if (tid < kChunk)
                sW[b][tid] = (float)prew;
            __syncthreads();
            if (ch + 1 < nch)
                fetch(ch + 1);
            if (!any)
                continue;
            // weights of the lane's ordinate at the eight points, to LDS in MFMA order
            const float *rowp = &sMr[b][lane * NMS];
            const float wl = sW[b][lane];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                // four groups of four points, the scheduler kept from interleaving them: one
                // group's operands are live at a time (152 registers for nm = 16, three waves
                // per SIMD; 192 and two waves unsplit)
#pragma unroll
                for (int v = h * NV / 4; v < (h + 1) * NV / 4; ++v) {
                    if ((v & 1) == 0 && (iso >> (v >> 1)) & 1u) {  // wave-uniform
                        sWE[wv][v * kWES + (lane & 3) * 16 + (lane >> 2)] = 0.0f;
                        continue;
                    }
                    float x = 0.0f;
#pragma unroll
                    for (int i = 0; i < NM; ++i)
                        x = fmaf(sA[wv][v * NM + i], rowp[i], x);
                    const float we = wl * __expf(x);  // hardware exp2: ranking needs no more
                    fpart[v] += we;
                    sWE[wv][v * kWES + (lane & 3) * 16 + (lane >> 2)] = we;
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            wave_lds_sync();
            const float *m0 = &sMfT[b][(rq * 4 + gq) * RS];
            const float *a0 = &sWE[wv][(rq < NV ? rq : 0) * kWES + gq * 16];
            const int swz = (rq >> 2) & 3;  // the rows' group swizzle (MixedChunk)
#pragma unroll
            for (int kb = 0; kb < kChunk / 16; ++kb) {
                f4 av = *reinterpret_cast<const f4 *>(a0 + 4 * kb);
                if (rq >= NV)
                    av = f4{0.0f, 0.0f, 0.0f, 0.0f};
                const f4 b0v = *reinterpret_cast<const f4 *>(m0 + 4 * (kb ^ swz));
                f4 b1v = f4{0.0f, 0.0f, 0.0f, 0.0f};
                if constexpr (NT == 2)
                    b1v = *reinterpret_cast<const f4 *>(
                        &sMfT[b][((16 + rq) * 4 + gq) * RS + 4 * (kb ^ swz)]);
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[kk], b0v[kk], acc[0], 0, 0, 0);
                    if constexpr (NT == 2)
                        acc[1] =
                            __builtin_amdgcn_mfma_f32_16x16x4f32(av[kk], b1v[kk], acc[1], 0, 0, 0);
                }
            }
            wave_lds_sync();  // sWE is rewritten by the next chunk
        }
        if (!any)
            continue;
        // ||g_v - u||^2: lane (rq, gq) holds G[v = 4 gq + r][j = rq (+16)] in acc[.][r]
        float nv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int v = 4 * gq + r;
            const int c = (v < NV ? v : 0) >> 1;
            const bool isoRow = v < NV && (v & 1) == 0 && ((iso >> c) & 1u);
            float t = 0.0f;
#pragma unroll
            for (int bt = 0; bt < NT; ++bt) {
                const int j = 16 * bt + rq;
                if (j < NM && cells[c] >= 0) {
                    const float uj = (float)S.vec[(size_t)cells[c] * kVec * NM + kVu * NM + j];
                    // an isotropic row's sums are zero: its gradient is -u_j for j >= 1
                    const float g0 = __expf(sA[wv][(2 * c) * NM] * (float)P.m0) *
                                     (float)(4.0 * M_PI * P.m0);
                    const float d = (isoRow && j == 0) ? g0 - uj : acc[bt][r] - uj;
                    t = fmaf(d, d, t);
                }
            }
#pragma unroll
            for (int o = 8; o >= 1; o >>= 1)
                t += __shfl_xor(t, o);
            nv[r] = t;
        }
        if (rq == 0 && gq < NV / 4) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                sN[wv][4 * gq + r] = nv[r];
        }
        wave_lds_sync();
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            // false when the proposal overflowed (NaN or inf)
            const bool take = sN[wv][2 * c + 1] < sN[wv][2 * c];
            float fs = take ? fpart[2 * c + 1] : fpart[2 * c];
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1)
                fs += __shfl_xor(fs, o);
            if (cells[c] < 0)
                continue;
            double *vw = S.vec + (size_t)cells[c] * kVec * NM;
            if (take && lane < NM) {
                const double an = anet[(size_t)slots[c] * NM + lane];
                vw[kVa * NM + lane] = an;
                vw[kVa0 * NM + lane] = an;
                vw[kVaOld * NM + lane] = an;
            }
            if (lane == 0) {
                if (take)
                    S.ints[(size_t)cells[c] * kInt + kIflags] |= C3D_FLAG_WARM;
                if (P.mixed) {
                    const double lf = log((double)fs);
                    S.scal[(size_t)cells[c] * kScal + kSshift] =
                        (lf > -700.0 && lf < 700.0) ? lf : 0.0;
                }
            }
        }
        wave_lds_sync();  // sN and sA are rewritten for the next cells
    }
}


// Cells still listed after the pass budget: report them and leave their input untouched.
template <int NM>
__global__ void k_budget(Lists L, c3d_cell_stats *stats, int *status)
{
    const int nWm = *L.nm;
    const int nW = nWm + *L.nf;
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nW; s += gridDim.x * blockDim.x) {
        if (s == 0)
            atomicOr(status, C3D_FLAG_BUDGET);
        if (stats) {
            c3d_cell_stats s0{0.0, 0.0, 0.0, 0, 0, 0, C3D_FLAG_BUDGET};
            stats[list_cell(L, s, nWm)] = s0;
        }
    }
}

// H only (solveH_cuda / solveHMn semantics): one wave per cell.
template <int NM>
__global__ __launch_bounds__(kBlock) void k_hessian_fp64(int64_t ncells, int nqPad,
                                                        const double *__restrict__ Mg,
                                                        const double *__restrict__ wg,
                                                        const double *__restrict__ alpha_g,
                                                        double *hs)
{
    constexpr int NT = (NM + 15) / 16;
    constexpr int NB = NT * (NT + 1) / 2;
    constexpr int NMS = NM | 1;
    __shared__ double sM[kChunk * NMS];
    __shared__ double sW[kChunk];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t cell = (int64_t)blockIdx.x * kWavesPerBlock + wv;
    const bool have = cell < ncells;
    double as[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i)
        as[i] = have ? alpha_g[cell * NM + i] : 0.0;
    d4 acc[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b)
        acc[b] = d4{0.0, 0.0, 0.0, 0.0};
    const int gq = lane >> 4, rq = lane & 15;
    for (int c = 0; c < nqPad / kChunk; ++c) {
        __syncthreads();
        for (int e = tid; e < kChunk * NM; e += kBlock) {
            int rr = e / NM, cc = e - rr * NM;
            sM[rr * NMS + cc] = Mg[(size_t)c * kChunk * NM + e];
        }
        if (tid < kChunk)
            sW[tid] = wg[(size_t)c * kChunk + tid];
        __syncthreads();
        if (have) {
            double x = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i)
                x = x + as[i] * sM[lane * NMS + i];
            const double we = sW[lane] * exp(x);
#pragma unroll
            for (int k = 0; k < kChunk / 4; ++k) {
                const int ql = 4 * k + gq;
                const double wq = __shfl(we, ql);
                const double b0 = (rq < NM) ? sM[ql * NMS + rq] : 0.0;
                acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(wq * b0, b0, acc[0], 0, 0, 0);
                if constexpr (NT == 2) {
                    const double b1 = (16 + rq < NM) ? sM[ql * NMS + 16 + rq] : 0.0;
                    acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(wq * b1, b0, acc[1], 0, 0, 0);
                    acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(wq * b1, b1, acc[2], 0, 0, 0);
                }
            }
        }
    }
    if (!have)
        return;
    double *H = hs + cell * NM * NM;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int R = (b == 0) ? 0 : 1, Cb = (b == 2) ? 1 : 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = 16 * R + gq + 4 * k, j = 16 * Cb + rq;
            if (i < NM && j < NM && j <= i) {
                H[i * NM + j] = acc[b][k];
                H[j * NM + i] = acc[b][k];
            }
        }
    }
}

// H only, mixed mode, nm <= kPairMaxNm: the column sweep's Hessian under the cell's own shift
// (the largest exponent and quadrature weight, so that every weight is at most one and the
// largest at least min w / max w), one wave per cell in column 0; H row-major, both triangles.
template <int NM>
__global__ __launch_bounds__(kWave) void k_hessian_cols(int64_t ncells, int nqPad,
                                                        const double *__restrict__ Mg,
                                                        const double *__restrict__ wg,
                                                        const _Float16 *__restrict__ Pimg,
                                                        const double *__restrict__ Rimg,
                                                        const double *__restrict__ alpha_g,
                                                        double *hs)
{
    constexpr int NMS = NM | 1;
    __shared__ __align__(16) double sR[RowImg<NM>::kDoubles];
    double *const sM = sR;
    __shared__ double sW[kChunk];
    __shared__ __align__(16) _Float16 sP[PairHess<NM>::kHalves];
    __shared__ double sA[NM];
    __shared__ double sH[PairHess<NM>::NP];
    __shared__ double sG[NM];
    __shared__ double sWE[kChunk];
    __shared__ double sX[64];  // the exp table
    const int lane = threadIdx.x;
    const int64_t cell = blockIdx.x;
    if (cell >= ncells)
        return;
    sX[lane] = kExp2Tab[lane];
    if (lane < NM)
        sA[lane] = alpha_g[cell * NM + lane];
    double xmax = -DBL_MAX, wtop = 0.0;
    for (int c = 0; c < nqPad / kChunk; ++c) {
        wave_lds_sync();
        for (int e = lane; e < kChunk * NM; e += kWave) {
            const int rr = e / NM, cc = e - rr * NM;
            sM[rr * NMS + cc] = Mg[(size_t)c * kChunk * NM + e];
        }
        sW[lane] = wg[(size_t)c * kChunk + lane];
        wave_lds_sync();
        if (sW[lane] > 0.0) {
            double x = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i)
                x = fma(sA[i], sM[lane * NMS + i], x);
            xmax = fmax(xmax, x);
            wtop = fmax(wtop, sW[lane]);
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        xmax = fmax(xmax, __shfl_xor(xmax, o));
        wtop = fmax(wtop, __shfl_xor(wtop, o));
    }
    const double shift = xmax + log(wtop);
    wave_lds_sync();
    (void)sweep_lone<NM, 1>(nqPad, Rimg, Pimg, sA, shift, sR, sR, sP, sP, sWE, nullptr, sH, sG,
                            false, lane, 0, sX);
    wave_lds_sync();
    double *H = hs + cell * NM * NM;
    for (int p = lane; p < PairHess<NM>::NP; p += kWave) {
        int i = 0;
        while (tri(i + 1) <= p)
            ++i;
        const int j = p - tri(i);
        H[i * NM + j] = sH[p];
        H[j * NM + i] = sH[p];
    }
}

// The dense part of one Newton iteration for a batch of given Hessians (one wave each):
// info = dpotrf, rcond = dpocon(norm1(H)), d = -H^{-1} g (custom or LAPACK solve order).
template <int NM>
__global__ __launch_bounds__(kBlock) void k_dense_solve(int64_t n, const double *__restrict__ Hs,
                                                        const double *__restrict__ gs, double *ds,
                                                        double *rconds, int *infos, double *Ls,
                                                        int lapackOrder)
{
    constexpr int LD = NM + 1;
    __shared__ double sT[kWavesPerBlock][NM * LD];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t c = (int64_t)blockIdx.x * kWavesPerBlock + wv;
    if (c >= n)
        return;
    const bool own = lane < NM;
    Dense<NM> D;
#pragma unroll
    for (int j = 0; j < NM; ++j)
        D.row[j] = own ? Hs[c * NM * NM + lane * NM + j] : 0.0;
    double rs = 0.0;
#pragma unroll
    for (int j = 0; j < NM; ++j)
        rs = rs + fabs(D.row[j]);
    double norm1 = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        double vv = rdl(rs, i);
        if (vv > norm1)
            norm1 = vv;
    }
    const int info = D.cholesky(lane);
    double rcond = 0.0, d = own ? gs[c * NM + lane] : 0.0;
    if (info == 0) {
        D.build_columns(sT[wv], LD, lane);
        rcond = D.pocon(norm1, lane);
        if (lapackOrder)
            D.potrs_lapack(d, lane);
        else
            D.potrs_custom(d, lane);
        d = -d;
    }
    if (own) {
        ds[c * NM + lane] = d;
#pragma unroll
        for (int j = 0; j < NM; ++j)
            Ls[c * NM * NM + lane * NM + j] = (j <= lane) ? D.row[j] : 0.0;
    }
    if (lane == 0) {
        rconds[c] = rcond;
        infos[c] = info;
    }
}

__global__ void k_set_int(int *p, int v) { *p = v; }

// Pass control of the asynchronous schedule (one block): the pass's bulk kernels read their
// list lengths from eff.  While more than skipLE cells are listed, eff is the list's own
// lengths; at or below it the bulk kernels see empty lists and the list carries over to the
// next generation unchanged, for the tail kernels.
__global__ __launch_bounds__(256) void k_gate(Lists Lc, Lists Ln, int *eff, int skipLE)
{
    const int nm = *Lc.nm, nf = *Lc.nf;
    const bool skip = nm + nf <= skipLE;
    if (threadIdx.x == 0) {
        eff[0] = skip ? 0 : nm;
        eff[1] = skip ? 0 : nf;
    }
    if (!skip)
        return;
    for (int i = threadIdx.x; i < nm; i += blockDim.x)
        Ln.Wm[i] = Lc.Wm[i];
    for (int i = threadIdx.x; i < nf; i += blockDim.x)
        Ln.Wf[i] = Lc.Wf[i];
    if (threadIdx.x == 0) {
        *Ln.nm = nm;
        *Ln.nf = nf;
    }
}
__global__ void k_set_int2(int *p, int v)
{
    p[0] = v;
    p[1] = v;
}
__global__ void k_zero_u64(unsigned long long *p) { p[threadIdx.x] = 0ull; }

unsigned grid_cap(c3d_ctx_s *ctx, int perCU) { return (unsigned)(ctx->numCU * perCU); }

hipEvent_t *take_events(c3d_ctx_s *ctx, int n, int kind)
{
    if ((int)ctx->evPool.size() < ctx->evUsed + n) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t x;
            if (hipEventCreate(&x) != hipSuccess)
                return nullptr;
            ctx->evPool.push_back(x);
        }
    }
    ctx->evSpans.push_back({kind, ctx->evUsed});
    hipEvent_t *ev = &ctx->evPool[ctx->evUsed];
    ctx->evUsed += n;
    return ev;
}

// Device bytes of a closure over n cells of order NM: the per-cell Newton state, H, the
// worklists and their counters, then (warm start) the features and proposals at the next
// 256-byte boundary.
size_t closure_state_bytes(size_t n, int nm)
{
    return n * (kVec * nm + kScal + (size_t)nm * nm + 1) * sizeof(double) +
           n * (kInt + 4) * sizeof(int) + 8 * sizeof(int);
}
size_t warm_feature_bytes(size_t n) { return (n * C3D_MLP_IN * sizeof(float) + 255) / 256 * 256; }
size_t closure_total_bytes(size_t n, int nm, bool warm)
{
    size_t b = closure_state_bytes(n, nm);
    if (warm)
        b = (b + 255) / 256 * 256 + warm_feature_bytes(n) + n * nm * sizeof(double);
    return b;
}

// ext: a caller-provided device buffer of ext->bytes for all of the closure's scratch (the
// no-allocation entry point), or null for the context's own buffers, grown on demand.
// With ext the closure allocates nothing and records no profiling events.
template <int NM>
hipError_t run_closure(c3d_ctx_s *ctx, const KParams &P, double *u, double *alpha,
                       c3d_cell_stats *stats, int maxPasses, hipStream_t s, const DevBuf *ext)
{
    const size_t n = (size_t)P.ncells;
    const bool warm = ctx->warm && ctx->mlp.have;
    const size_t bytes = closure_state_bytes(n, NM);
    hipError_t e = hipSuccess;
    char *base;
    if (ext) {
        if (closure_total_bytes(n, NM, warm) > ext->bytes)
            return hipErrorInvalidValue;
        base = (char *)ext->p;
    } else {
        e = grow(ctx->closureState, bytes);
        if (e != hipSuccess)
            return e;
        base = (char *)ctx->closureState.p;
    }
    const bool prof = ctx->profiling && !ext;
    State S;
    S.vec = (double *)base;
    S.scal = S.vec + n * kVec * NM;
    S.H = S.scal + n * kScal;
    S.fdir = S.H + n * NM * NM;
    S.ints = (int *)(S.fdir + n);
    int *Wbase = S.ints + n * kInt;  // 4 lists of n: generation 0 (mixed, fp64), 1 (...)
    int *counts = Wbase + 4 * n;     // [2g] mixed, [2g+1] fp64 of generation g
    Lists Lg[2];
    for (int g = 0; g < 2; ++g) {
        Lg[g].Wm = Wbase + (2 * g) * n;
        Lg[g].Wf = Wbase + (2 * g + 1) * n;
        Lg[g].nm = counts + 2 * g;
        Lg[g].nf = counts + 2 * g + 1;
    }
    const bool ref = ctx->precision == C3D_PREC_FP64_REFORDER;
    const bool mixed = P.mixed != 0;
    hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, s, counts, 0);
    ctx->evUsed = 0;
    ctx->evSpans.clear();
    ctx->lastPasses = 0;
    ctx->lastTailLaunches = 0;
    ctx->lastStream = s;
    unsigned long long *cntBulk = prof ? ctx->d_sweepCount : nullptr;
    unsigned long long *cntTail = prof ? ctx->d_sweepCount + 1 : nullptr;
    if (prof)
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(3), 0, s, ctx->d_sweepCount);
    hipLaunchKernelGGL(k_init<NM>, dim3(grid_cap(ctx, 16)), dim3(kBlock), 0, s, P, u, alpha, S,
                       Lg[0], stats, ctx->d_status);
    if (warm) {
        const size_t xb = warm_feature_bytes(n);
        char *wbase;
        if (ext) {
            wbase = base + (bytes + 255) / 256 * 256;
        } else {
            e = grow(ctx->warmBuf, xb + n * NM * sizeof(double));
            if (e != hipSuccess)
                return e;
            wbase = (char *)ctx->warmBuf.p;
        }
        float *X = (float *)wbase;
        double *anet = (double *)(wbase + xb);
        hipEvent_t *ev = prof ? take_events(ctx, 2, 2) : nullptr;
        if (prof && !ev)
            return hipErrorOutOfMemory;
        if (ev)
            (void)hipEventRecord(ev[0], s);
        hipLaunchKernelGGL(k_warm_features<NM>, dim3((unsigned)((n * C3D_MLP_IN + 255) / 256)),
                           dim3(256), 0, s, (int64_t)n, S, Lg[0],
                           (const float *)ctx->mlp.normDev.p, X);
        e = mlp_infer(ctx, (int64_t)n, X, anet, s);
        if (e != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_warm_select<NM>, dim3(grid_cap(ctx, 8)), dim3(kBlock), 0, s, P,
                           ctx->d_M, ctx->d_wpad, S, Lg[0], anet);
        if (ev)
            (void)hipEventRecord(ev[1], s);
    }
    // one bulk pass over the listed cells of Lc (list lengths read through Lc.nm / Lc.nf)
    auto bulk_pass = [&](const Lists &Lc, const Lists &Ln) -> hipError_t {
        hipEvent_t *ev = prof ? take_events(ctx, 4, 0) : nullptr;
        if (prof && !ev)
            return hipErrorOutOfMemory;
        if (ev)
            (void)hipEventRecord(ev[0], s);
        if (mixed)
            hipLaunchKernelGGL(k_sweep_cols<NM>, dim3(grid_cap(ctx, 1)), dim3(ColShape<NM>::kBlock),
                               0, s, P, ctx->d_R, (const _Float16 *)ctx->d_P16, S, Lc.Wm, Lc.nm);
        if (ev)
            (void)hipEventRecord(ev[1], s);
        if (ref)
            hipLaunchKernelGGL(k_sweep_ref<NM>, dim3(grid_cap(ctx, 8)), dim3(kBlock), 0, s, P,
                               ctx->d_M, ctx->d_wpad, S, Lc.Wf, Lc.nf);
        else
            hipLaunchKernelGGL(k_sweep<NM>, dim3(grid_cap(ctx, 8)), dim3(kBlock), 0, s, P,
                               ctx->d_M, ctx->d_wpad, S, Lc.Wf, Lc.nf);
        if (ev)
            (void)hipEventRecord(ev[2], s);
        if (ref)
            hipLaunchKernelGGL((k_step<NM, true>), dim3(grid_cap(ctx, 16)), dim3(kBlock), 0, s,
                               P, S, Lc, Ln, u, alpha, stats, ctx->d_status, cntBulk);
        else
            hipLaunchKernelGGL((k_step<NM, false>), dim3(grid_cap(ctx, 16)), dim3(kBlock), 0,
                               s, P, S, Lc, Ln, u, alpha, stats, ctx->d_status, cntBulk);
        if (ev)
            (void)hipEventRecord(ev[3], s);
        return hipSuccess;
    };
    // one launch of the fused tail kernel: `grid` blocks stride over Lc's cells, K rounds each;
    // wide: four waves per cell
    auto tail_launch = [&](const Lists &Lc, const Lists &Ln, unsigned grid, int K,
                           bool wide) -> hipError_t {
        hipEvent_t *ev = prof ? take_events(ctx, 2, 1) : nullptr;
        if (prof && !ev)
            return hipErrorOutOfMemory;
        if (ev)
            (void)hipEventRecord(ev[0], s);
#define C3D_TAIL(R, X, TW)                                                                   \
    hipLaunchKernelGGL((k_tail<NM, R, X, TW>), dim3(grid), dim3(TW * kWave), 0, s, P, ctx->d_M,  \
                       ctx->d_MT, ctx->d_wpad, (const _Float16 *)ctx->d_P16, ctx->d_R, S, Lc, \
                       Ln, u, alpha, stats, ctx->d_status, K, cntTail)
        if (ref && wide)
            C3D_TAIL(true, false, 4);
        else if (ref)
            C3D_TAIL(true, false, 1);
        else if (mixed && (wide || NM > 16))  // nm = 25: the one-wave kernel's 110 KB of LDS
            C3D_TAIL(false, true, 4);         // leave one block per CU either way
        else if (mixed) {
            if constexpr (NM <= 16)
                C3D_TAIL(false, true, 1);
        }
        else if (wide)
            C3D_TAIL(false, false, 4);
        else
            C3D_TAIL(false, false, 1);
#undef C3D_TAIL
        if (ev)
            (void)hipEventRecord(ev[1], s);
        ++ctx->lastTailLaunches;
        return hipSuccess;
    };
    int cur = 0;
    if (ctx->asyncPasses > 0) {
        // Asynchronous schedule (no host synchronisation; capturable in a graph): a fixed number
        // of bulk passes, each gated on the device, which stop sweeping once at most tailCells
        // cells are listed; then one narrow tail launch of tailPasses rounds per cell and one
        // wide launch for the rest of the pass budget.  A cell's operations are the ones the
        // adaptive schedule performs, so the results are the same.
        int *eff = counts + 4;
        const Lists Le0{Lg[0].Wm, Lg[0].Wf, eff, eff + 1}, Le1{Lg[1].Wm, Lg[1].Wf, eff, eff + 1};
        const int B = std::min(ctx->asyncPasses, maxPasses);
        for (int p = 0; p < B; ++p) {
            const Lists &Lc = Lg[cur], &Ln = Lg[1 - cur];
            hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, s, Ln.nm, 0);
            hipLaunchKernelGGL(k_gate, dim3(1), dim3(256), 0, s, Lc, Ln, eff, ctx->tailCells);
            e = bulk_pass(cur == 0 ? Le0 : Le1, Ln);
            if (e != hipSuccess)
                return e;
            cur = 1 - cur;
        }
        int rest = maxPasses - B;
        const int K1 = std::min(ctx->tailPasses, rest);
        if (K1 > 0) {
            hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, s, Lg[1 - cur].nm, 0);
            e = tail_launch(Lg[cur], Lg[1 - cur], (unsigned)std::max(ctx->tailCells, 1024), K1,
                            false);
            if (e != hipSuccess)
                return e;
            cur = 1 - cur;
            rest -= K1;
        }
        if (rest > 0) {
            hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, s, Lg[1 - cur].nm, 0);
            e = tail_launch(Lg[cur], Lg[1 - cur], (unsigned)std::max(ctx->tailWideCells, 256),
                            rest, true);
            if (e != hipSuccess)
                return e;
            cur = 1 - cur;
        }
        ctx->lastPasses = maxPasses;
        hipLaunchKernelGGL(k_budget<NM>, dim3(64), dim3(256), 0, s, Lg[cur], stats,
                           ctx->d_status);
        return hipGetLastError();
    }
    int64_t left = (int64_t)n;  // upper bound on the listed cells until the first read-back
    int pass = 0;
    while (pass < maxPasses) {
        const Lists &Lc = Lg[cur], &Ln = Lg[1 - cur];
        hipLaunchKernelGGL(k_set_int2, dim3(1), dim3(1), 0, s, Ln.nm, 0);
        const bool tail = ctx->tailCells > 0 && left <= ctx->tailCells;
        if (tail) {
            const int K = std::min(ctx->tailPasses, maxPasses - pass);
            // few cells: four waves per cell settle up to four line-search trials a round
            e = tail_launch(Lc, Ln, (unsigned)left, K, left <= ctx->tailWideCells);
            if (e != hipSuccess)
                return e;
            pass += K;
        } else {
            e = bulk_pass(Lc, Ln);
            if (e != hipSuccess)
                return e;
            ++pass;
        }
        ctx->lastPasses = pass;
        cur = 1 - cur;
        if (tail || (pass & 7) == 0) {
            int lft[2] = {0, 0};
            e = hipMemcpyAsync(lft, counts + 2 * cur, sizeof(lft), hipMemcpyDeviceToHost, s);
            if (e == hipSuccess)
                e = hipStreamSynchronize(s);
            if (e != hipSuccess)
                return e;
            left = (int64_t)lft[0] + lft[1];
            if (left == 0)
                return hipGetLastError();
        }
    }
    hipLaunchKernelGGL(k_budget<NM>, dim3(64), dim3(256), 0, s, Lg[cur], stats, ctx->d_status);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_closure_fp64(c3d_ctx_s *ctx, int64_t ncells, double *u, double *alpha,
                               const c3d_closure_opts &o, c3d_cell_stats *stats,
                               hipStream_t s, const DevBuf *ext)
{
    KParams P;
    P.ncells = ncells;
    P.nq = ctx->t.nq;
    P.nqPad = ctx->nqPad;
    P.maxIter = o.maxIter;
    P.unbatched = o.unbatched;
    P.tol = o.tol;
    P.tolGamma = o.tolGamma;
    P.condHMax = o.condHMax;
    P.m0 = ctx->cc.m0;
    P.isoAlpha0 = ctx->cc.isoAlpha0;
    P.minf = ctx->cc.minf;
    P.mixed = ctx->precision == C3D_PREC_MIXED;
    P.mixedMaxIter = ctx->mixedMaxIter;
    P.mixedCondMax = ctx->mixedCondMax;
    P.mixedSwitch = ctx->mixedSwitch;
    P.mixedFinishCond = ctx->mixedFinishCond;
    if (ncells <= 0)
        return hipSuccess;
    if (ncells > INT32_MAX / 2)
        return hipErrorInvalidValue;
    const int maxPasses = ctx->maxPasses;
    switch (ctx->t.nm) {
    case 4: return run_closure<4>(ctx, P, u, alpha, stats, maxPasses, s, ext);
    case 9: return run_closure<9>(ctx, P, u, alpha, stats, maxPasses, s, ext);
    case 16: return run_closure<16>(ctx, P, u, alpha, stats, maxPasses, s, ext);
    case 25: return run_closure<25>(ctx, P, u, alpha, stats, maxPasses, s, ext);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_dense_solve(int nm, int64_t n, const double *H, const double *g, double *d,
                              double *rcond, int *info, double *L, int lapackOrder, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    dim3 grid((unsigned)((n + kWavesPerBlock - 1) / kWavesPerBlock));
#define C3D_DS(NMV) hipLaunchKernelGGL(k_dense_solve<NMV>, grid, dim3(kBlock), 0, s, n, H, g, d, rcond, info, L, lapackOrder)
    switch (nm) {
    case 4: C3D_DS(4); break;
    case 9: C3D_DS(9); break;
    case 16: C3D_DS(16); break;
    case 25: C3D_DS(25); break;
    default: return hipErrorInvalidValue;
    }
#undef C3D_DS
    return hipGetLastError();
}

hipError_t launch_hessian_fp64(c3d_ctx_s *ctx, int64_t ncells, const double *alpha,
                               double *hs, hipStream_t s)
{
    if (ncells <= 0)
        return hipSuccess;
    dim3 grid((unsigned)((ncells + kWavesPerBlock - 1) / kWavesPerBlock));
    if (ctx->precision == C3D_PREC_MIXED) {
#define C3D_HC(NMV) hipLaunchKernelGGL(k_hessian_cols<NMV>, dim3((unsigned)ncells), dim3(kWave), 0, s, ncells, ctx->nqPad, ctx->d_M, ctx->d_wpad, (const _Float16 *)ctx->d_P16, ctx->d_R, alpha, hs)
        switch (ctx->t.nm) {
        case 4: C3D_HC(4); break;
        case 9: C3D_HC(9); break;
        case 16: C3D_HC(16); break;
        case 25: C3D_HC(25); break;
        default: return hipErrorInvalidValue;
        }
#undef C3D_HC
        return hipGetLastError();
    }
    switch (ctx->t.nm) {
    case 4: hipLaunchKernelGGL(k_hessian_fp64<4>, grid, dim3(kBlock), 0, s, ncells, ctx->nqPad, ctx->d_M, ctx->d_wpad, alpha, hs); break;
    case 9: hipLaunchKernelGGL(k_hessian_fp64<9>, grid, dim3(kBlock), 0, s, ncells, ctx->nqPad, ctx->d_M, ctx->d_wpad, alpha, hs); break;
    case 16: hipLaunchKernelGGL(k_hessian_fp64<16>, grid, dim3(kBlock), 0, s, ncells, ctx->nqPad, ctx->d_M, ctx->d_wpad, alpha, hs); break;
    case 25: hipLaunchKernelGGL(k_hessian_fp64<25>, grid, dim3(kBlock), 0, s, ncells, ctx->nqPad, ctx->d_M, ctx->d_wpad, alpha, hs); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace c3d

// ---------------------------------------------------------------- no-allocation closure
// c3d_closure_noalloc keeps all of a closure's scratch (Newton state, Hessians, worklists,
// warm-start features and proposals) in this fixed array; every device the library runs on
// holds it until the process ends.  Host code never passes the array's name to a kernel or a
// copy: it uses the device address looked up once per device at context creation.
static __device__ __align__(256) unsigned char scratch[C3D_NOALLOC_SCRATCH_BYTES];

namespace c3d {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMtx;
void *g_scratchPtr[kMaxDevices] = {};

void *scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return nullptr;
    std::lock_guard<std::mutex> lk(g_scratchMtx);
    return g_scratchPtr[device];
}
}  // namespace

hipError_t closure_scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratchMtx);
    if (g_scratchPtr[device])
        return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess)
        return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess)
        return e;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != device) {
        const hipError_t e2 = hipSetDevice(cur);
        if (e == hipSuccess)
            e = e2;
    }
    if (e == hipSuccess && !ptr)
        e = hipErrorInvalidSymbol;
    if (e == hipSuccess)
        g_scratchPtr[device] = ptr;
    return e;
}

}  // namespace c3d

extern "C" int c3d_closure_noalloc(c3d_ctx ctx, int64_t ncells, double *u, double *alpha,
                                   const c3d_closure_opts *opts, c3d_cell_stats *stats,
                                   void *stream)
{
    using namespace c3d;
    if (!ctx || !opts || ncells < 0 || (ncells > 0 && (!u || !alpha)) || opts->maxIter < 1)
        return C3D_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mtx);
    if (opts->useGaunt) {
        // served by the direct form only once c3d_closure / c3d_hessian has checked beta
        if (ctx->precision == C3D_PREC_FP64_REFORDER ||
            ctx->gauntDirect[opts->useGauntSparse ? 1 : 0] != 1)
            return C3D_ERR_UNSUPPORTED;
    }
    const bool warm = ctx->warm && ctx->mlp.have;
    if (closure_total_bytes((size_t)ncells, ctx->t.nm, warm) > (size_t)C3D_NOALLOC_SCRATCH_BYTES)
        return C3D_ERR_SCRATCH;
    void *ptr = scratch_ptr(ctx->device);
    if (!ptr)
        return C3D_ERR_HIP;
    DevBuf ext;
    ext.p = ptr;
    ext.bytes = C3D_NOALLOC_SCRATCH_BYTES;
    const hipError_t e =
        launch_closure_fp64(ctx, ncells, u, alpha, *opts, stats, (hipStream_t)stream, &ext);
    return e == hipSuccess ? C3D_OK : C3D_ERR_HIP;
}
