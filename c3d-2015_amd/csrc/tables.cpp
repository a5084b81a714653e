// tables.cpp -- quadrature, real spherical harmonic basis, Gaunt coefficients and P_N
// half-range operators, built on the host once per context.
//
//   quadrature  init.cpp:423-431, 510-531: Q Gauss-Legendre nodes in mu (ascending, as
//               GSL's glfixed orders them, utils.cpp:52-57) x 2Q midpoints in phi;
//               w_q = (pi/Q) w_i, q = 2Q*i + j.
//   basis       utils.cpp:60-68 with GSL's sphPlm normalisation and Condon-Shortley phase,
//               ordered n = 0..N, m = -n..n (init.cpp:526-530); M2 the same to degree 2N.
//   Gaunt       init.cpp:287-337: beta_(ij) = argmin ||M2 beta - m_i m_j||_2 (QR least
//               squares), CSR of |beta| > 1e-13.
//   P_N ops     init.cpp:199-234.
#include <cmath>
#include <cstdint>
#include <vector>

#include "c3d_internal.h"

namespace c3d {
namespace {

// Legendre P_n and P_n' at x by the three-term recurrence.
void legendre(int n, double x, double &p, double &dp)
{
    double pm1 = 1.0, pc = x;
    if (n == 0) {
        p = 1.0;
        dp = 0.0;
        return;
    }
    for (int k = 2; k <= n; ++k) {
        double pn = ((2.0 * k - 1.0) * x * pc - (k - 1.0) * pm1) / k;
        pm1 = pc;
        pc = pn;
    }
    p = pc;
    dp = n * (x * pc - pm1) / (x * x - 1.0);
}

void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        // descending root i of P_n
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), p, dp;
        for (int it = 0; it < 100; ++it) {
            legendre(n, z, p, dp);
            double step = p / dp;
            z -= step;
            if (std::fabs(step) <= 1e-16 * std::fabs(z) || std::fabs(step) < 1e-300)
                break;
        }
        legendre(n, z, p, dp);
        double wt = 2.0 / ((1.0 - z * z) * dp * dp);
        x[n - 1 - i] = z;
        x[i] = -z;
        w[i] = w[n - 1 - i] = wt;
    }
    if (n & 1)
        x[n / 2] = 0.0;
}

// sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) P_l^m(x) with the (-1)^m phase, 0 <= m <= l.
double sph_plm(int l, int m, double x)
{
    double s = std::sqrt((1.0 - x) * (1.0 + x));
    double pmm = 1.0 / std::sqrt(4.0 * M_PI);
    for (int k = 1; k <= m; ++k)
        pmm *= -s * std::sqrt((2.0 * k + 1.0) / (2.0 * k));
    if (l == m)
        return pmm;
    double p1 = x * std::sqrt(2.0 * m + 3.0) * pmm;
    double p0 = pmm;
    for (int k = m + 2; k <= l; ++k) {
        double a = std::sqrt((4.0 * k * k - 1.0) / (double(k) * k - double(m) * m));
        double b = std::sqrt((double(k - 1) * (k - 1) - double(m) * m) /
                             (4.0 * double(k - 1) * (k - 1) - 1.0));
        double pk = a * (x * p1 - b * p0);
        p0 = p1;
        p1 = pk;
    }
    return p1;
}

double real_sh(int m, int n, double mu, double phi)
{
    if (m < 0)
        return M_SQRT2 * std::sin(m * phi) * sph_plm(n, -m, mu);
    if (m > 0)
        return M_SQRT2 * std::cos(m * phi) * sph_plm(n, m, mu);
    return sph_plm(n, 0, mu);
}

// Least-squares beta for every (i,j) by Householder QR of the nq x nm2 matrix M2.  With fewer
// ordinates than degree-2N functions (nq < nm2) the problem is underdetermined (GSL's QR solve
// of init.cpp:287-337 refuses it too): no coefficients, nnz = 0, and a context then declines
// Gaunt-form requests (its check that beta reproduces the basis products fails).
void gaunt(HostTables &t)
{
    const int nq = t.nq, nm = t.nm, nm2 = t.nm2;
    if (nq < nm2) {
        t.beta.assign((size_t)nm * nm * nm2, 0.0);
        t.betaRowPtr.assign(nm * nm + 1, 0);
        t.betaVal.clear();
        t.betaColInd.clear();
        t.nnz = 0;
        return;
    }
    std::vector<double> R((size_t)nq * nm2);  // column-major working copy
    for (int q = 0; q < nq; ++q)
        for (int k = 0; k < nm2; ++k)
            R[(size_t)k * nq + q] = t.M2[(size_t)q * nm2 + k];
    std::vector<double> tau(nm2), diag(nm2);
    for (int k = 0; k < nm2; ++k) {
        double *c = &R[(size_t)k * nq];
        double nrm2 = 0.0;
        for (int q = k; q < nq; ++q)
            nrm2 += c[q] * c[q];
        double nrm = std::sqrt(nrm2);
        double beta = (c[k] > 0.0) ? -nrm : nrm;
        double v0 = c[k] - beta;
        double vv = v0 * v0;
        for (int q = k + 1; q < nq; ++q)
            vv += c[q] * c[q];
        tau[k] = vv > 0.0 ? 2.0 * v0 * v0 / vv : 0.0;
        double inv = v0 != 0.0 ? 1.0 / v0 : 1.0;
        for (int q = k + 1; q < nq; ++q)
            c[q] *= inv;
        diag[k] = beta;
        for (int j = k + 1; j < nm2; ++j) {
            double *cj = &R[(size_t)j * nq];
            double s = cj[k];
            for (int q = k + 1; q < nq; ++q)
                s += c[q] * cj[q];
            s *= tau[k];
            cj[k] -= s;
            for (int q = k + 1; q < nq; ++q)
                cj[q] -= s * c[q];
        }
    }
    t.beta.assign((size_t)nm * nm * nm2, 0.0);
    std::vector<double> b(nq);
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j) {
            for (int q = 0; q < nq; ++q)
                b[q] = t.M2[(size_t)q * nm2 + i] * t.M2[(size_t)q * nm2 + j];
            for (int k = 0; k < nm2; ++k) {
                const double *c = &R[(size_t)k * nq];
                double s = b[k];
                for (int q = k + 1; q < nq; ++q)
                    s += c[q] * b[q];
                s *= tau[k];
                b[k] -= s;
                for (int q = k + 1; q < nq; ++q)
                    b[q] -= s * c[q];
            }
            double *x = &t.beta[(size_t)(i * nm + j) * nm2];
            for (int k = nm2 - 1; k >= 0; --k) {
                double s = b[k];
                for (int c = k + 1; c < nm2; ++c)
                    s -= R[(size_t)c * nq + k] * x[c];
                x[k] = s / diag[k];
            }
        }
    t.betaRowPtr.assign(nm * nm + 1, 0);
    t.betaVal.clear();
    t.betaColInd.clear();
    for (int ij = 0; ij < nm * nm; ++ij) {
        t.betaRowPtr[ij] = (int32_t)t.betaVal.size();
        for (int k = 0; k < nm2; ++k) {
            double v = t.beta[(size_t)ij * nm2 + k];
            if (std::fabs(v) > 1e-13) {
                t.betaVal.push_back(v);
                t.betaColInd.push_back(k);
            }
        }
    }
    t.betaRowPtr[nm * nm] = (int32_t)t.betaVal.size();
    t.nnz = (int)t.betaVal.size();
}

}  // namespace

void build_pn_operators(HostTables &t)
{
    const int nm = t.nm, nq = t.nq;
    const size_t s = (size_t)nm * nm;
    t.pnOps.assign(6 * s, 0.0);
    for (int i = 0; i < nm; ++i)
        for (int j = 0; j < nm; ++j) {
            double a[6] = {0, 0, 0, 0, 0, 0};
            for (int q = 0; q < nq; ++q) {
                double mi = t.M[(size_t)q * nm + i], mj = t.M[(size_t)q * nm + j];
                a[t.xi[q] > 0 ? 0 : 1] += t.xi[q] * mi * mj * t.w[q];
                a[t.eta[q] > 0 ? 2 : 3] += t.eta[q] * mi * mj * t.w[q];
                a[t.mu[q] > 0 ? 4 : 5] += t.mu[q] * mi * mj * t.w[q];
            }
            for (int o = 0; o < 6; ++o)
                t.pnOps[o * s + (size_t)i * nm + j] = a[o];
        }
}

void build_tables(int N, int Q, HostTables &t)
{
    t.N = N;
    t.Q = Q;
    t.nm = (N + 1) * (N + 1);
    t.nm2 = (2 * N + 1) * (2 * N + 1);
    t.nq = 2 * Q * Q;
    std::vector<double> gx, gw;
    gauss_legendre(Q, gx, gw);
    t.w.assign(t.nq, 0.0);
    t.xi.assign(t.nq, 0.0);
    t.eta.assign(t.nq, 0.0);
    t.mu.assign(t.nq, 0.0);
    t.M.assign((size_t)t.nq * t.nm, 0.0);
    t.M2.assign((size_t)t.nq * t.nm2, 0.0);
    for (int i = 0; i < Q; ++i)
        for (int j = 0; j < 2 * Q; ++j) {
            const int q = 2 * Q * i + j;
            const double phi = (j + 0.5) * M_PI / Q;
            const double st = std::sqrt(1.0 - gx[i] * gx[i]);
            t.w[q] = M_PI / Q * gw[i];
            t.xi[q] = st * std::cos(phi);
            t.eta[q] = st * std::sin(phi);
            t.mu[q] = gx[i];
            int k = 0;
            for (int n = 0; n <= 2 * N; ++n)
                for (int m = -n; m <= n; ++m, ++k) {
                    double v = real_sh(m, n, gx[i], phi);
                    t.M2[(size_t)q * t.nm2 + k] = v;
                    if (n <= N)
                        t.M[(size_t)q * t.nm + k] = v;
                }
        }
    gaunt(t);
    build_pn_operators(t);
}

}  // namespace c3d
