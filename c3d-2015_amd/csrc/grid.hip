// grid.hip -- the time-stepping kernels of the hot path on gfx950, FP64.
//
//   k_ghost_fill   periodic 2-layer face copy            comm.cpp:277-323 (serial branch)
//   k_ansatz       A[q][cell] = exp(alpha.m_q) (M_N) or alpha.m_q (P_N) on all cells
//                                                        flux_kinetic.cpp:20-37, 116-123
//   k_stencil      limited upwind differences, sum_q m_q (Fx/dx+Fy/dy+Fz/dz), optionally
//                  fused with the Euler stage             flux_kinetic.cpp:45-60, 126-196,
//                                                        update.cpp:21-37
//   k_flux_moment  P_N Fromm flux with half-range operators  flux_moment.cpp:30-104
//   k_euler        Euler stage with isotropic scattering update.cpp:21-37
//   k_rk2          uOld <- u / u <- (u + uOld)/2          update.cpp:206-212, 268-274
//   k_make_batch   synthetic closure batch (SplitMix64 counter RNG)
//
// Layout: the reference's AoS grid (src/global.h:15-16), cells (nx+4)(ny+4)(nz+4) with z
// fastest, nm moments contiguous per cell.  The ansatz is materialised ordinate-major,
// A[q*Ng + cell], so every stencil load of a wavefront is one contiguous segment.
// Every expression keeps the reference's evaluation order (no FMA contraction), so the
// results match the CPU code bit for bit whenever the exponentials agree.
#include <cfloat>
#include <cmath>

#include "c3d_detmath.h"
#include "c3d_internal.h"

#pragma clang fp contract(off)

namespace c3d {
namespace {

constexpr int kTB = 256;

struct Grid {
    int nx, ny, nz;
    int NX, NY, NZ;  // with ghosts
    __host__ __device__ size_t idx(int i, int j, int k) const
    {
        return (size_t)k + (size_t)NZ * ((size_t)j + (size_t)NY * (size_t)i);
    }
};

Grid make_grid(int nx, int ny, int nz) { return Grid{nx, ny, nz, nx + 4, ny + 4, nz + 4}; }

// ---------------------------------------------------------------- ghost fill
__global__ void k_ghost_fill(Grid G, int nm, double *u)
{
    // three slab families of 2 x 2 x (face area) cells, each copies nm moments
    const int64_t ax = (int64_t)2 * G.ny * G.nz, ay = (int64_t)2 * G.nx * G.nz,
                  az = (int64_t)2 * G.nx * G.ny;
    const int64_t total = (ax + ay + az) * 2 * nm;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int m = (int)(t % nm);
        int64_t c = t / nm;
        int side = (int)(c & 1);  // 0: low ghosts <- high interior, 1: high <- low
        c >>= 1;
        size_t dst, src;
        if (c < ax) {
            int l = (int)(c / ((int64_t)G.ny * G.nz));
            int64_t rem = c % ((int64_t)G.ny * G.nz);
            int j = 2 + (int)(rem / G.nz), k = 2 + (int)(rem % G.nz);
            if (side == 0) {
                dst = G.idx(l, j, k);
                src = G.idx(G.nx + l, j, k);
            } else {
                dst = G.idx(G.nx + 2 + l, j, k);
                src = G.idx(2 + l, j, k);
            }
        } else if (c < ax + ay) {
            c -= ax;
            int l = (int)(c / ((int64_t)G.nx * G.nz));
            int64_t rem = c % ((int64_t)G.nx * G.nz);
            int i = 2 + (int)(rem / G.nz), k = 2 + (int)(rem % G.nz);
            if (side == 0) {
                dst = G.idx(i, l, k);
                src = G.idx(i, G.ny + l, k);
            } else {
                dst = G.idx(i, G.ny + 2 + l, k);
                src = G.idx(i, 2 + l, k);
            }
        } else {
            c -= ax + ay;
            int l = (int)(c / ((int64_t)G.nx * G.ny));
            int64_t rem = c % ((int64_t)G.nx * G.ny);
            int i = 2 + (int)(rem / G.ny), j = 2 + (int)(rem % G.ny);
            if (side == 0) {
                dst = G.idx(i, j, l);
                src = G.idx(i, j, G.nz + l);
            } else {
                dst = G.idx(i, j, G.nz + 2 + l);
                src = G.idx(i, j, 2 + l);
            }
        }
        u[dst * nm + m] = u[src * nm + m];
    }
}

// ------------------------------------------------------------------- ansatz
// One thread per cell (all cells, ghosts included), M staged through LDS 64 ordinates at
// a time; the ansatz argument keeps flux_kinetic.cpp:23-24's sequential order.
template <int NM>
__global__ __launch_bounds__(kTB) void k_ansatz(int64_t ng, int nq, const double *__restrict__ M,
                                                 const double *__restrict__ alpha, double *A,
                                                 int isPn)
{
    __shared__ double sM[kChunk * NM];
    const int64_t cell = (int64_t)blockIdx.x * kTB + threadIdx.x;
    const bool have = cell < ng;
    double a[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m)
        a[m] = have ? alpha[cell * NM + m] : 0.0;
    for (int q0 = 0; q0 < nq; q0 += kChunk) {
        const int nqc = min(kChunk, nq - q0);
        __syncthreads();
        for (int e = threadIdx.x; e < nqc * NM; e += kTB)
            sM[e] = M[(size_t)q0 * NM + e];
        __syncthreads();
        if (have) {
            for (int ql = 0; ql < nqc; ++ql) {
                double kin = 0.0;
#pragma unroll
                for (int m = 0; m < NM; ++m)
                    kin += a[m] * sM[ql * NM + m];
                A[(size_t)(q0 + ql) * ng + cell] =
                    isPn == 1 ? kin : (isPn == 2 ? det_exp(kin) : exp(kin));
            }
        }
    }
}

__device__ __forceinline__ double sgn1(double x) { return x >= 0.0 ? 1.0 : -1.0; }
__device__ __forceinline__ double mmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double mmax(double a, double b) { return a > b ? a : b; }
// flux_kinetic.cpp:45-49
__device__ __forceinline__ double minmod(double x, double y)
{
    return sgn1(x) * mmax(0.0, mmin(fabs(x), y * sgn1(x)));
}
// flux_kinetic.cpp:55-60.  TH1: theta == 1 (the minmod limiter), where th * x == x exactly.
template <bool TH1>
__device__ __forceinline__ double slopefit(double l, double c, double r, double th)
{
    if constexpr (TH1)
        return minmod(r - c, minmod(0.5 * (r - l), c - l));
    else
        return minmod(th * (r - c), minmod(0.5 * (r - l), th * (c - l)));
}
// one axis of flux_kinetic.cpp:141-191; A at offsets -2..+2 along the axis
template <bool TH1>
__device__ __forceinline__ double axis_flux(double am2, double am1, double a0, double ap1,
                                            double ap2, double dir, double wq, double th)
{
    double f;
    if (dir > 0) {
        f = a0 + 0.5 * slopefit<TH1>(am1, a0, ap1, th) - am1 -
            0.5 * slopefit<TH1>(am2, am1, a0, th);
    } else {
        f = ap1 - 0.5 * slopefit<TH1>(a0, ap1, ap2, th) - a0 +
            0.5 * slopefit<TH1>(am1, a0, ap1, th);
    }
    return f * dir * wq;
}

struct StencilArgs {
    Grid G;
    int nq;
    double dx, dy, dz, theta, dt;
    int fuseEuler;
    double rdx, rdy, rdz;  // 1/dx, ... (used when every spacing is a power of two)
};

// x / d == x * (1/d) exactly, for every x, iff d is a power of two
bool exact_reciprocal(double d)
{
    int e = 0;
    return d > 0.0 && std::isfinite(d) && std::frexp(d, &e) == 0.5 && e > -1021 && e < 1022;
}

// P2: dx, dy, dz powers of two (division by them is multiplication by exact reciprocals)
template <int NM, bool P2, bool TH1>
__global__ __launch_bounds__(kTB) void k_stencil(StencilArgs S, const double *__restrict__ A,
                                                  const double *__restrict__ M,
                                                  const double *__restrict__ xi,
                                                  const double *__restrict__ eta,
                                                  const double *__restrict__ mu,
                                                  const double *__restrict__ w, double *flux,
                                                  double *u, const double *__restrict__ sigT,
                                                  const double *__restrict__ sigS)
{
    const Grid &G = S.G;
    const int64_t ni = (int64_t)G.nx * G.ny * G.nz;
    const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
    if (t >= ni)
        return;
    const int k = 2 + (int)(t % G.nz);
    const int j = 2 + (int)((t / G.nz) % G.ny);
    const int i = 2 + (int)(t / ((int64_t)G.nz * G.ny));
    const size_t ng = (size_t)G.NX * G.NY * G.NZ;
    const size_t c = G.idx(i, j, k);
    const size_t sx = (size_t)G.NY * G.NZ, sy = (size_t)G.NZ;
    double fl[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m)
        fl[m] = 0.0;
    // the 13 ansatz values of the cell's stencil at ordinate q: the centre, then -2, -1, +1,
    // +2 along x, y and z; the next ordinate's are in flight while this one's are used
    const ptrdiff_t off[13] = {0,
                               -2 * (ptrdiff_t)sx, -(ptrdiff_t)sx, (ptrdiff_t)sx, 2 * (ptrdiff_t)sx,
                               -2 * (ptrdiff_t)sy, -(ptrdiff_t)sy, (ptrdiff_t)sy, 2 * (ptrdiff_t)sy,
                               -2, -1, 1, 2};
    double v[13], nv[13];
#pragma unroll
    for (int e = 0; e < 13; ++e)
        v[e] = A[c + off[e]];
    for (int q = 0; q < S.nq; ++q) {
        if (q + 1 < S.nq) {
            const double *An = A + (size_t)(q + 1) * ng + c;
#pragma unroll
            for (int e = 0; e < 13; ++e)
                nv[e] = An[off[e]];
        }
        const double a0 = v[0];
        const double wq = w[q];
        const double fX = axis_flux<TH1>(v[1], v[2], a0, v[3], v[4], xi[q], wq, S.theta);
        const double fY = axis_flux<TH1>(v[5], v[6], a0, v[7], v[8], eta[q], wq, S.theta);
        const double fZ = axis_flux<TH1>(v[9], v[10], a0, v[11], v[12], mu[q], wq, S.theta);
#pragma unroll
        for (int e = 0; e < 13; ++e)
            v[e] = nv[e];
        double s;
        if constexpr (P2)
            s = fX * S.rdx + fY * S.rdy + fZ * S.rdz;
        else
            s = fX / S.dx + fY / S.dy + fZ / S.dz;
#pragma unroll
        for (int m = 0; m < NM; ++m)
            fl[m] += M[(size_t)q * NM + m] * s;
    }
    if (S.fuseEuler) {
        // update.cpp:28-35
        const double dt = S.dt, st = sigT[c], ss = sigS[c];
        double *uc = u + c * NM;
        uc[0] = uc[0] * (1.0 + dt * ss - dt * st) - dt * fl[0];
#pragma unroll
        for (int m = 1; m < NM; ++m)
            uc[m] = uc[m] * (1.0 - dt * st) - dt * fl[m];
    } else {
#pragma unroll
        for (int m = 0; m < NM; ++m)
            flux[c * NM + m] = fl[m];
    }
}

__global__ void k_zero_ghost_flux(Grid G, int nm, double *flux)
{
    const int64_t ng = (int64_t)G.NX * G.NY * G.NZ;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ng * nm;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = t / nm;
        int k = (int)(c % G.NZ), j = (int)((c / G.NZ) % G.NY), i = (int)(c / ((int64_t)G.NZ * G.NY));
        bool interior = i >= 2 && i < G.nx + 2 && j >= 2 && j < G.ny + 2 && k >= 2 && k < G.nz + 2;
        if (!interior)
            flux[t] = 0.0;
    }
}

// -------------------------------------------------------------- moment flux
template <int NM>
__global__ __launch_bounds__(kTB) void k_flux_moment(Grid G, const double *__restrict__ u,
                                                      double *flux,
                                                      const double *__restrict__ ops, double dx,
                                                      double dy, double dz)
{
    __shared__ double sO[6 * NM * NM];
    for (int e = threadIdx.x; e < 6 * NM * NM; e += kTB)
        sO[e] = ops[e];
    __syncthreads();
    const int64_t ni = (int64_t)G.nx * G.ny * G.nz;
    const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
    if (t >= ni)
        return;
    const int k = 2 + (int)(t % G.nz);
    const int j = 2 + (int)((t / G.nz) % G.ny);
    const int i = 2 + (int)(t / ((int64_t)G.nz * G.ny));
    const size_t c = G.idx(i, j, k);
    const size_t strides[3] = {(size_t)G.NY * G.NZ, (size_t)G.NZ, 1};
    const double hs[3] = {dx, dy, dz};
    double out[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m)
        out[m] = 0.0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const size_t s = strides[ax];
        const double *um2 = u + (c - 2 * s) * NM, *um1 = u + (c - s) * NM, *u0 = u + c * NM,
                     *up1 = u + (c + s) * NM, *up2 = u + (c + 2 * s) * NM;
        const double *Op = sO + (2 * ax) * NM * NM, *Om = sO + (2 * ax + 1) * NM * NM;
        double temp[NM], F[NM];
#pragma unroll
        for (int m = 0; m < NM; ++m)
            temp[m] = 0.25 * up1[m] + 0.75 * u0[m] - 1.25 * um1[m] + 0.25 * um2[m];
#pragma unroll
        for (int r = 0; r < NM; ++r) {
            double b = 0.0;
#pragma unroll
            for (int q = 0; q < NM; ++q)
                b += Op[r * NM + q] * temp[q];
            F[r] = b;
        }
#pragma unroll
        for (int m = 0; m < NM; ++m)
            temp[m] = -0.25 * up2[m] + 1.25 * up1[m] - 0.75 * u0[m] - 0.25 * um1[m];
#pragma unroll
        for (int r = 0; r < NM; ++r) {
            double b = F[r];
#pragma unroll
            for (int q = 0; q < NM; ++q)
                b += Om[r * NM + q] * temp[q];
            F[r] = b;
        }
        // flux = xF/dx + yF/dy + zF/dz, accumulated left to right
#pragma unroll
        for (int m = 0; m < NM; ++m)
            out[m] = (ax == 0) ? F[m] / hs[0] : out[m] + F[m] / hs[ax];
    }
#pragma unroll
    for (int m = 0; m < NM; ++m)
        flux[c * NM + m] = out[m];
}

// -------------------------------------------------------------------- euler
__global__ void k_euler(Grid G, int nm, double *u, const double *__restrict__ flux,
                        const double *__restrict__ sigT, const double *__restrict__ sigS,
                        double dt)
{
    const int64_t ni = (int64_t)G.nx * G.ny * G.nz;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ni)
        return;
    const int k = 2 + (int)(t % G.nz);
    const int j = 2 + (int)((t / G.nz) % G.ny);
    const int i = 2 + (int)(t / ((int64_t)G.nz * G.ny));
    const size_t c = G.idx(i, j, k);
    const double st = sigT[c], ss = sigS[c];
    double *uc = u + c * nm;
    const double *fc = flux + c * nm;
    uc[0] = uc[0] * (1.0 + dt * ss - dt * st) - dt * fc[0];
    for (int m = 1; m < nm; ++m)
        uc[m] = uc[m] * (1.0 - dt * st) - dt * fc[m];
}

// ---------------------------------------------------------------------- rk2
__global__ void k_rk2(Grid G, int nm, double *u, double *uOld, int kind)
{
    const int64_t n = (int64_t)G.nx * G.ny * G.nz * nm;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(t % nm);
        const int64_t cc = t / nm;
        const int k = 2 + (int)(cc % G.nz);
        const int j = 2 + (int)((cc / G.nz) % G.ny);
        const int i = 2 + (int)(cc / ((int64_t)G.nz * G.ny));
        const size_t e = G.idx(i, j, k) * nm + m;
        if (kind == 0)
            uOld[e] = u[e];
        else
            u[e] = 0.5 * (u[e] + uOld[e]);
    }
}

// ------------------------------------------------------------ synthetic batch
__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t idx, uint64_t stream)
{
    uint64_t key = splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
    uint64_t z = splitmix64(key + idx);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}
__device__ __forceinline__ double normal01(uint64_t seed, uint64_t idx, uint64_t stream)
{
    double u1 = uniform01(seed, idx, 2 * stream), u2 = uniform01(seed, idx, 2 * stream + 1);
    return sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
}

// Initial state over all cells including ghosts (init.cpp:480-506), cell centres at
// a + (i - 2 + 1/2) d with ghosts at their own unwrapped centres (init.cpp:483-485):
//   ic 0: the reference Gaussian, u0 = max(floor, 2 sqrt(pi) G_sigma(x)) with the normalised
//         G (init.cpp:27-34, 493-495);
//   ic 1: pulse rho = max(exp(-|x|^2/(2 sigma^2)), floor) (1 + noise U[-1,1]), u0 = 2 sqrt(pi)
//         rho (counter RNG stream 7 on the cell index);
//   ic 2: the smooth state of init.cpp:497-502.
// sigma_t = sigma_s = crossSection; alpha (if not null) from init.cpp:262-268.
__global__ __launch_bounds__(kTB) void k_init_grid(int nx, int ny, int nz, int nm, double ax,
                                                   double ay, double az, double dx, double dy,
                                                   double dz, int ic, double sigma, double floorv,
                                                   double noise, uint64_t seed, double cross,
                                                   double *u, double *alpha, double *sigmaT,
                                                   double *sigmaS)
{
    const int64_t NX = nx + 4, NY = ny + 4, NZ = nz + 4;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= NX * NY * NZ)
        return;
    const int k = (int)(c % NZ), j = (int)((c / NZ) % NY), i = (int)(c / (NZ * NY));
    const double x = ax + (i - 2) * dx + dx / 2.0;
    const double y = ay + (j - 2) * dy + dy / 2.0;
    const double z = az + (k - 2) * dz + dz / 2.0;
    sigmaT[c] = cross;
    sigmaS[c] = cross;
    double u0, u123 = 0.0;
    if (ic == 0) {
        const double f1 = 1.0 / (sigma * sqrt(2.0 * M_PI));
        const double g = f1 * f1 * f1 * exp(-(x * x + y * y + z * z) / (2.0 * sigma * sigma));
        u0 = fmax(floorv, g * 2.0 * sqrt(M_PI));
    } else if (ic == 1) {
        double rho = fmax(exp(-(x * x + y * y + z * z) / (2.0 * sigma * sigma)), floorv);
        if (noise != 0.0)
            rho = rho * (1.0 + noise * (2.0 * uniform01(seed, (uint64_t)c, 7) - 1.0));
        u0 = 2.0 * sqrt(M_PI) * rho;
    } else {
        u0 = (2.0 + cos(2 * M_PI * x) * cos(2 * M_PI * y) * cos(2 * M_PI * z)) * 2.0 * sqrt(M_PI);
        u123 = 0.3 * u0;
    }
    u[c * nm] = u0;
    for (int m = 1; m < nm; ++m)
        u[c * nm + m] = (m <= 3) ? u123 : 0.0;
    if (alpha) {
        alpha[c * nm] = sqrt(4.0 * M_PI) * log(u0 / sqrt(4.0 * M_PI));
        for (int m = 1; m < nm; ++m)
            alpha[c * nm + m] = 0.0;
    }
}

template <int NM>
__global__ __launch_bounds__(kTB) void k_make_batch(int64_t n, int nq, const double *__restrict__ M,
                                                     const double *__restrict__ w, uint64_t seed,
                                                     double smin, double smax, double *u,
                                                     double *alpha0, double *alphaStar)
{
    __shared__ double sM[kChunk * NM];
    __shared__ double sWt[kChunk];
    const int64_t c = (int64_t)blockIdx.x * kTB + threadIdx.x;
    const bool have = c < n;
    double as[NM], uc[NM];
    const double s = smin + (smax - smin) * uniform01(seed, (uint64_t)c, 1);
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        as[k] = s * normal01(seed, (uint64_t)c * 64 + k, 2);
        uc[k] = 0.0;
    }
    for (int q0 = 0; q0 < nq; q0 += kChunk) {
        const int nqc = min(kChunk, nq - q0);
        __syncthreads();
        for (int e = threadIdx.x; e < nqc * NM; e += kTB)
            sM[e] = M[(size_t)q0 * NM + e];
        if (threadIdx.x < nqc)
            sWt[threadIdx.x] = w[q0 + threadIdx.x];
        __syncthreads();
        if (have) {
            for (int ql = 0; ql < nqc; ++ql) {
                double e = 0.0;
#pragma unroll
                for (int k = 0; k < NM; ++k)
                    e = e + as[k] * sM[ql * NM + k];
                e = sWt[ql] * exp(e);
#pragma unroll
                for (int k = 0; k < NM; ++k)
                    uc[k] += e * sM[ql * NM + k];
            }
        }
    }
    if (!have)
        return;
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        u[c * NM + k] = uc[k];
        alpha0[c * NM + k] = (k == 0) ? sqrt(4.0 * M_PI) * log(uc[0] / sqrt(4.0 * M_PI)) : 0.0;
        if (alphaStar)
            alphaStar[c * NM + k] = as[k];
    }
}

unsigned blocks_for(int64_t n, int tb = kTB, int64_t cap = 1 << 20)
{
    int64_t b = (n + tb - 1) / tb;
    if (b > cap)
        b = cap;
    if (b < 1)
        b = 1;
    return (unsigned)b;
}

}  // namespace

hipError_t launch_ghost_fill(int nx, int ny, int nz, int nm, double *u, hipStream_t s)
{
    Grid G = make_grid(nx, ny, nz);
    int64_t total = ((int64_t)2 * ny * nz + (int64_t)2 * nx * nz + (int64_t)2 * nx * ny) * 2 * nm;
    hipLaunchKernelGGL(k_ghost_fill, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, s, G,
                       nm, u);
    return hipGetLastError();
}

hipError_t launch_ansatz(c3d_ctx_s *ctx, int nx, int ny, int nz, const double *alpha, double *A,
                         int isPn, hipStream_t s)
{
    // kind: 1 linear (P_N), 2 deterministic exp (reference-order mode), 0 device exp
    if (!isPn && ctx->precision == C3D_PREC_FP64_REFORDER)
        isPn = 2;
    const int64_t ng = (int64_t)(nx + 4) * (ny + 4) * (nz + 4);
    dim3 grid(blocks_for(ng, kTB, INT32_MAX));
    const int nq = ctx->t.nq;
    switch (ctx->t.nm) {
    case 4: hipLaunchKernelGGL(k_ansatz<4>, grid, dim3(kTB), 0, s, ng, nq, ctx->d_M, alpha, A, isPn); break;
    case 9: hipLaunchKernelGGL(k_ansatz<9>, grid, dim3(kTB), 0, s, ng, nq, ctx->d_M, alpha, A, isPn); break;
    case 16: hipLaunchKernelGGL(k_ansatz<16>, grid, dim3(kTB), 0, s, ng, nq, ctx->d_M, alpha, A, isPn); break;
    case 25: hipLaunchKernelGGL(k_ansatz<25>, grid, dim3(kTB), 0, s, ng, nq, ctx->d_M, alpha, A, isPn); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stencil(c3d_ctx_s *ctx, int nx, int ny, int nz, const double *A, double *flux,
                          double *u, const double *sigmaT, const double *sigmaS, double dt,
                          double dx, double dy, double dz, double theta, int fuseEuler,
                          int zeroGhosts, hipStream_t s)
{
    StencilArgs S{make_grid(nx, ny, nz), ctx->t.nq, dx, dy, dz, theta, dt, fuseEuler,
                  1.0 / dx, 1.0 / dy, 1.0 / dz};
    const bool p2 = exact_reciprocal(dx) && exact_reciprocal(dy) && exact_reciprocal(dz);
    const bool th1 = theta == 1.0;
    const int64_t ni = (int64_t)nx * ny * nz;
    dim3 grid(blocks_for(ni, kTB, INT32_MAX));
    if (zeroGhosts && !fuseEuler) {
        const int64_t ng = (int64_t)S.G.NX * S.G.NY * S.G.NZ;
        hipLaunchKernelGGL(k_zero_ghost_flux, dim3(blocks_for(ng * ctx->t.nm, 256, 65536)),
                           dim3(256), 0, s, S.G, ctx->t.nm, flux);
    }
#define C3D_STENCIL_V(NMV, P, T)                                                              \
    hipLaunchKernelGGL((k_stencil<NMV, P, T>), grid, dim3(kTB), 0, s, S, A, ctx->d_M,          \
                       ctx->d_xi, ctx->d_eta, ctx->d_mu, ctx->d_w, flux, u, sigmaT, sigmaS)
#define C3D_STENCIL(NMV)                                                                      \
    if (p2 && th1)                                                                             \
        C3D_STENCIL_V(NMV, true, true);                                                        \
    else if (p2)                                                                               \
        C3D_STENCIL_V(NMV, true, false);                                                       \
    else if (th1)                                                                              \
        C3D_STENCIL_V(NMV, false, true);                                                       \
    else                                                                                       \
        C3D_STENCIL_V(NMV, false, false)
    switch (ctx->t.nm) {
    case 4: C3D_STENCIL(4); break;
    case 9: C3D_STENCIL(9); break;
    case 16: C3D_STENCIL(16); break;
    case 25: C3D_STENCIL(25); break;
    default: return hipErrorInvalidValue;
    }
#undef C3D_STENCIL
#undef C3D_STENCIL_V
    return hipGetLastError();
}

hipError_t launch_flux_moment(c3d_ctx_s *ctx, int nx, int ny, int nz, const double *u,
                              double *flux, double dx, double dy, double dz, hipStream_t s)
{
    Grid G = make_grid(nx, ny, nz);
    const int64_t ni = (int64_t)nx * ny * nz;
    dim3 grid(blocks_for(ni, kTB, INT32_MAX));
    switch (ctx->t.nm) {
    case 4: hipLaunchKernelGGL(k_flux_moment<4>, grid, dim3(kTB), 0, s, G, u, flux, ctx->d_pnOps, dx, dy, dz); break;
    case 9: hipLaunchKernelGGL(k_flux_moment<9>, grid, dim3(kTB), 0, s, G, u, flux, ctx->d_pnOps, dx, dy, dz); break;
    case 16: hipLaunchKernelGGL(k_flux_moment<16>, grid, dim3(kTB), 0, s, G, u, flux, ctx->d_pnOps, dx, dy, dz); break;
    case 25: hipLaunchKernelGGL(k_flux_moment<25>, grid, dim3(kTB), 0, s, G, u, flux, ctx->d_pnOps, dx, dy, dz); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_euler(int nx, int ny, int nz, int nm, double *u, const double *flux,
                        const double *sigmaT, const double *sigmaS, double dt, hipStream_t s)
{
    Grid G = make_grid(nx, ny, nz);
    const int64_t ni = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_euler, dim3(blocks_for(ni, 256, INT32_MAX)), dim3(256), 0, s, G, nm, u,
                       flux, sigmaT, sigmaS, dt);
    return hipGetLastError();
}

hipError_t launch_rk2(int nx, int ny, int nz, int nm, double *u, double *uOld, int kind,
                      hipStream_t s)
{
    Grid G = make_grid(nx, ny, nz);
    const int64_t n = (int64_t)nx * ny * nz * nm;
    hipLaunchKernelGGL(k_rk2, dim3(blocks_for(n, 256, 1 << 20)), dim3(256), 0, s, G, nm, u, uOld,
                       kind);
    return hipGetLastError();
}

hipError_t launch_init_grid(int nx, int ny, int nz, int nm, const double *a, const double *d,
                            int ic, double sigma, double floorv, double noise, uint64_t seed,
                            double cross, double *u, double *alpha, double *sigmaT,
                            double *sigmaS, hipStream_t s)
{
    const int64_t ng = (int64_t)(nx + 4) * (ny + 4) * (nz + 4);
    hipLaunchKernelGGL(k_init_grid, dim3((unsigned)((ng + kTB - 1) / kTB)), dim3(kTB), 0, s, nx,
                       ny, nz, nm, a[0], a[1], a[2], d[0], d[1], d[2], ic, sigma, floorv, noise,
                       seed, cross, u, alpha, sigmaT, sigmaS);
    return hipGetLastError();
}

hipError_t launch_make_batch(c3d_ctx_s *ctx, int64_t n, uint64_t seed, double smin, double smax,
                             double *u, double *alpha0, double *alphaStar, hipStream_t s)
{
    dim3 grid(blocks_for(n, kTB, INT32_MAX));
    const int nq = ctx->t.nq;
    switch (ctx->t.nm) {
    case 4: hipLaunchKernelGGL(k_make_batch<4>, grid, dim3(kTB), 0, s, n, nq, ctx->d_M, ctx->d_w, seed, smin, smax, u, alpha0, alphaStar); break;
    case 9: hipLaunchKernelGGL(k_make_batch<9>, grid, dim3(kTB), 0, s, n, nq, ctx->d_M, ctx->d_w, seed, smin, smax, u, alpha0, alphaStar); break;
    case 16: hipLaunchKernelGGL(k_make_batch<16>, grid, dim3(kTB), 0, s, n, nq, ctx->d_M, ctx->d_w, seed, smin, smax, u, alpha0, alphaStar); break;
    case 25: hipLaunchKernelGGL(k_make_batch<25>, grid, dim3(kTB), 0, s, n, nq, ctx->d_M, ctx->d_w, seed, smin, smax, u, alpha0, alphaStar); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ statistics summary
// Per-block partial sums of the closure's per-cell outcomes, the quantities update.cpp:145-174
// folds: sums and maxima of totalIter and iterGamma, and the count of cells on each ladder
// rung (r within 1e-10 of the rung's value, opt.h:37-38).  Row b of `part` holds block b's
// [n, sumIter, sumGamma, maxIter, maxGamma, rung0..rung4]; the host adds the rows.  No
// atomics: a grid-stride loop, then a tree in LDS.
constexpr int kSumFields = 10;
__global__ __launch_bounds__(kTB) void k_stats_partials(int64_t n,
                                                        const c3d_cell_stats *__restrict__ st,
                                                        long long *part)
{
    __shared__ long long sh[kSumFields][kTB];
    long long v[kSumFields] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const double reg[5] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1};
    for (int64_t c = (int64_t)blockIdx.x * kTB + threadIdx.x; c < n;
         c += (int64_t)gridDim.x * kTB) {
        const c3d_cell_stats s = st[c];
        v[0] += 1;
        v[1] += s.totalIter;
        v[2] += s.iterGamma;
        v[3] = v[3] > s.totalIter ? v[3] : s.totalIter;
        v[4] = v[4] > s.iterGamma ? v[4] : s.iterGamma;
#pragma unroll
        for (int r = 0; r < 5; ++r)
            v[5 + r] += fabs(s.r - reg[r]) < 1e-10 ? 1 : 0;
    }
#pragma unroll
    for (int f = 0; f < kSumFields; ++f)
        sh[f][threadIdx.x] = v[f];
    __syncthreads();
    for (int w = kTB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int f = 0; f < kSumFields; ++f) {
                const long long a = sh[f][threadIdx.x], b = sh[f][threadIdx.x + w];
                sh[f][threadIdx.x] = (f == 3 || f == 4) ? (a > b ? a : b) : a + b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < kSumFields)
        part[(size_t)blockIdx.x * kSumFields + threadIdx.x] = sh[threadIdx.x][0];
}

hipError_t launch_stats_partials(int64_t n, const c3d_cell_stats *st, long long *part,
                                 int blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_stats_partials, dim3(blocks), dim3(kTB), 0, s, n, st, part);
    return hipGetLastError();
}

}  // namespace c3d
