/*
 * c3d_oracle.c -- plain FP64 CPU restatement of the C3D-2015 hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see c3d_oracle.h).  Built with -ffp-contract=off so that
 * every expression rounds as the reference's own g++ -O3 build (no FMA) does.
 *
 * Parity pins (tests/ hold the checks):
 *   orc_solveH, orc_fobj, orc_linesearch, orc_flux_kinetic, orc_flux_moment,
 *   orc_ghost_fill            -> the reference's own sources compiled into oracle/_ref
 *   orc_gauss_legendre        -> numpy leggauss fixtures, sum w = 2, exactness
 *   orc_sph_plm / real_sh     -> scipy lpmv fixtures, closed forms for n <= 1,
 *                                orthonormality under the quadrature
 *   orc_gaunt                 -> numpy lstsq, exactness identity sum_k b_ijk m2_k = m_i m_j
 *   orc_dpotrf_l/dpocon_l/dpotrs_lapack -> scipy.linalg.lapack (LAPACK) live and fixtures
 *   orc_dpotrs_custom         -> numpy.linalg.solve
 *   orc_dnrm2                 -> numpy norm (1 ulp)
 *   orc_closure*              -> round trip u=<m e^{a*.m}> -> a*, isotropic closed form,
 *                                KKT (||g|| <= tol), forced ladder / restart cases,
 *                                per-iteration replay through the reference's fobj +
 *                                linesearch and scipy LAPACK
 *   orc_euler / orc_update    -> closed form on a uniform state, composition of pinned parts
 */
#include "c3d_oracle.h"

#include <alloca.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double orc_reg[ORC_NUM_REG] = {0.0, 1e-3, 1e-2, 1e-1, 5e-1}; /* opt.h:37-38 */

/* ---- exponential / logarithm used on the path ----
 * mode 0: the C library's exp/log, as the reference calls them (fobj.cpp:50,149;
 *         optBatch.cpp:139,314,336; flux_kinetic.cpp:35).
 * mode 1: a deterministic evaluation from + - * / only (Cody & Waite 1980 reduction with a
 *         split ln2 and a degree-13 Taylor polynomial for exp; the atanh series for log),
 *         which any IEEE binary64 machine evaluates to the same bits.  The device's
 *         reference-order mode uses the same method, so mode 1 makes the oracle comparable
 *         to it bit for bit. */
static int orc_exp_mode = 0;
void orc_set_exp_mode(int mode) { orc_exp_mode = mode; }

double orc_exp_det(double x)
{
    const double kInvLn2 = 0x1.71547652b82fep+0;
    const double kLn2Hi = 0x1.62e42fefa3800p-1;
    const double kLn2Lo = 0x1.ef35793c76730p-45;
    static const double c[14] = {
        1.0, 1.0, 0.5, 0x1.5555555555555p-3, 0x1.5555555555555p-5, 0x1.1111111111111p-7,
        0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-16,
        0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
        0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33};
    if (x != x)
        return x + x;
    if (x > 710.0)
        return HUGE_VAL;
    if (x < -746.0)
        return 0.0;
    double k = rint(x * kInvLn2);
    double r = (x - k * kLn2Hi) - k * kLn2Lo;
    double p = c[13];
    for (int n = 12; n >= 0; n--)
        p = p * r + c[n];
    return ldexp(p, (int)k);
}

double orc_log_det(double x)
{
    const double kLn2Hi = 0x1.62e42fefa3800p-1;
    const double kLn2Lo = 0x1.ef35793c76730p-45;
    if (!(x > 0.0))
        return (x == 0.0) ? -HUGE_VAL : NAN;
    if (x == HUGE_VAL)
        return x;
    int e = 0;
    double m = frexp(x, &e);
    if (m < 0x1.6a09e667f3bcdp-1) {
        m = m * 2.0;
        e = e - 1;
    }
    double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double p = 1.0 / 23.0;
    for (int k = 21; k >= 1; k -= 2)
        p = p * s2 + 1.0 / (double)k;
    double lm = 2.0 * s * p;
    double de = (double)e;
    return de * kLn2Hi + (lm + de * kLn2Lo);
}

static inline double orc_exp(double x) { return orc_exp_mode ? orc_exp_det(x) : exp(x); }
static inline double orc_log(double x) { return orc_exp_mode ? orc_log_det(x) : log(x); }

#define ORC_MIN(a, b) (((a) < (b)) ? (a) : (b)) /* utils.h:12 */
#define ORC_MAX(a, b) (((a) > (b)) ? (a) : (b)) /* utils.h:13 */
#define ORC_SIGN(a, b) ((b) >= 0.0 ? fabs(a) : -fabs(a)) /* utils.h:14 */

/* ===================================================================== tables */

/* Gauss-Legendre nodes on [-1,1], ascending (utils.cpp:52-57 via GSL glfixed, whose
 * points are ordered by increasing value).  Newton on P_n from the Tricomi guess. */
void orc_gauss_legendre(int n, double *x, double *w)
{
    for (int i = 0; i < (n + 1) / 2; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) <= 1e-16 * fabs(z) || fabs(z - z1) < 1e-300)
                break;
        }
        /* one more evaluation at the converged node for the weight */
        {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
        }
        double wi = 2.0 / ((1.0 - z * z) * pp * pp);
        /* z decreases with i: largest node first; store ascending */
        x[n - 1 - i] = z;
        x[i] = -z;
        w[n - 1 - i] = wi;
        w[i] = wi;
    }
    if (n % 2 == 1)
        x[n / 2] = 0.0;
}

/* Normalised associated Legendre function with the Condon-Shortley phase, the value of
 * GSL's gsl_sf_legendre_sphPlm(l, m, x) (utils.cpp:63-67):
 *   sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) P_l^m(x),  P_l^m including (-1)^m. */
double orc_sph_plm(int l, int m, double x)
{
    double pmm = sqrt(1.0 / (4.0 * M_PI));
    double somx2 = sqrt((1.0 - x) * (1.0 + x));
    for (int i = 1; i <= m; i++)
        pmm *= -somx2 * sqrt((2.0 * i + 1.0) / (2.0 * i));
    if (l == m)
        return pmm;
    double pmmp1 = x * sqrt(2.0 * m + 3.0) * pmm;
    if (l == m + 1)
        return pmmp1;
    double pll = 0.0;
    for (int ll = m + 2; ll <= l; ll++) {
        double a = sqrt((4.0 * ll * ll - 1.0) / ((double)ll * ll - (double)m * m));
        double b = sqrt(((ll - 1.0) * (ll - 1.0) - (double)m * m) /
                        (4.0 * (ll - 1.0) * (ll - 1.0) - 1.0));
        pll = a * (x * pmmp1 - b * pmm);
        pmm = pmmp1;
        pmmp1 = pll;
    }
    return pll;
}

/* Real spherical harmonic, utils.cpp:60-68. */
double orc_real_sh(int m, int n, double mu, double phi)
{
    if (m < 0)
        return M_SQRT2 * sin(m * phi) * orc_sph_plm(n, -m, mu);
    else if (m > 0)
        return M_SQRT2 * cos(m * phi) * orc_sph_plm(n, m, mu);
    return orc_sph_plm(n, 0, mu);
}

/* Quadrature and basis tables, init.cpp:423-548. */
void orc_quadrature(int N, int Q, double *w, double *xi, double *eta, double *mu,
                    double *M, double *M2)
{
    int nm = (N + 1) * (N + 1), nm2 = (2 * N + 1) * (2 * N + 1);
    double *wg = (double *)malloc(Q * sizeof(double));
    double *mg = (double *)malloc(Q * sizeof(double));
    double *phi = (double *)malloc(2 * Q * sizeof(double));
    orc_gauss_legendre(Q, mg, wg);
    for (int k = 0; k < 2 * Q; k++)
        phi[k] = (k + 0.5) * M_PI / Q; /* init.cpp:512-513 */
    for (int q1 = 0; q1 < Q; q1++) {
        for (int q2 = 0; q2 < 2 * Q; q2++) {
            int q = 2 * q1 * Q + q2; /* init.cpp:520 */
            w[q] = M_PI / Q * wg[q1];
            xi[q] = sqrt(1.0 - mg[q1] * mg[q1]) * cos(phi[q2]);
            eta[q] = sqrt(1.0 - mg[q1] * mg[q1]) * sin(phi[q2]);
            mu[q] = mg[q1];
            int k = 0;
            for (int n = 0; n < N + 1; n++)
                for (int m = -n; m <= n; m++)
                    M[q * nm + k++] = orc_real_sh(m, n, mg[q1], phi[q2]);
            if (M2) {
                k = 0;
                for (int n = 0; n < 2 * N + 1; n++)
                    for (int m = -n; m <= n; m++)
                        M2[q * nm2 + k++] = orc_real_sh(m, n, mg[q1], phi[q2]);
            }
        }
    }
    free(wg);
    free(mg);
    free(phi);
}

/* Gaunt coefficients by least squares, init.cpp:287-337: beta_(ij,:) solves
 * min || M2 beta - (m2_i * m2_j) ||_2 over the ordinates, via Householder QR as in
 * gsl_linalg_QR_decomp + QR_lssolve.  CSR keeps |beta| > 1e-13.  Returns nnz.
 * betaVal/betaColInd/betaRowPtr may be NULL (count only). */
int orc_gaunt(int nm, int nm2, int nq, const double *M2, double *beta, double *betaVal,
              int32_t *betaColInd, int32_t *betaRowPtr)
{
    /* fewer ordinates than degree-2N functions: underdetermined (GSL's QR least squares
     * refuses it); beta stays zero, and the CSR (if asked for) is empty */
    if (nq < nm2) {
        memset(beta, 0, (size_t)nm * nm * nm2 * sizeof(double));
        if (betaRowPtr)
            memset(betaRowPtr, 0, (size_t)(nm * nm + 1) * sizeof(int32_t));
        return 0;
    }
    /* A = M2 (nq x nm2), column-major copy for Householder */
    double *A = (double *)malloc((size_t)nq * nm2 * sizeof(double));
    double *tau = (double *)malloc(nm2 * sizeof(double));
    double *b = (double *)malloc(nq * sizeof(double));
    for (int q = 0; q < nq; q++)
        for (int k = 0; k < nm2; k++)
            A[q + (size_t)nq * k] = M2[q * nm2 + k];
    for (int k = 0; k < nm2; k++) {
        double *col = A + (size_t)nq * k;
        double nrm = 0.0;
        for (int q = k; q < nq; q++)
            nrm += col[q] * col[q];
        nrm = sqrt(nrm);
        double alpha = (col[k] > 0) ? -nrm : nrm;
        double v0 = col[k] - alpha;
        /* v = (v0, col[k+1:]) ; tau = -v0/alpha style normalisation: v <- v / v0 */
        double vnorm2 = v0 * v0;
        for (int q = k + 1; q < nq; q++)
            vnorm2 += col[q] * col[q];
        tau[k] = (vnorm2 > 0) ? 2.0 * v0 * v0 / vnorm2 : 0.0;
        for (int q = k + 1; q < nq; q++)
            col[q] /= (v0 != 0 ? v0 : 1.0);
        col[k] = alpha; /* R diagonal */
        /* apply H = I - tau v v^T to the remaining columns, v = (1, col[k+1:]) */
        for (int j = k + 1; j < nm2; j++) {
            double *cj = A + (size_t)nq * j;
            double s = cj[k];
            for (int q = k + 1; q < nq; q++)
                s += col[q] * cj[q];
            s *= tau[k];
            cj[k] -= s;
            for (int q = k + 1; q < nq; q++)
                cj[q] -= s * col[q];
        }
    }
    int nnz = 0;
    for (int i = 0; i < nm; i++) {
        for (int j = 0; j < nm; j++) {
            for (int q = 0; q < nq; q++)
                b[q] = M2[q * nm2 + i] * M2[q * nm2 + j];
            /* Q^T b */
            for (int k = 0; k < nm2; k++) {
                const double *col = A + (size_t)nq * k;
                double s = b[k];
                for (int q = k + 1; q < nq; q++)
                    s += col[q] * b[q];
                s *= tau[k];
                b[k] -= s;
                for (int q = k + 1; q < nq; q++)
                    b[q] -= s * col[q];
            }
            /* back substitution R x = (Q^T b)[0:nm2] */
            double *x = beta + ((size_t)(i * nm + j)) * nm2;
            for (int k = nm2 - 1; k >= 0; k--) {
                double s = b[k];
                for (int c = k + 1; c < nm2; c++)
                    s -= A[k + (size_t)nq * c] * x[c];
                x[k] = s / A[k + (size_t)nq * k];
            }
            for (int k = 0; k < nm2; k++)
                if (fabs(x[k]) > 1e-13)
                    nnz++;
        }
    }
    if (betaVal && betaColInd && betaRowPtr) {
        int ptr = 0;
        for (int ij = 0; ij < nm * nm; ij++) {
            betaRowPtr[ij] = ptr;
            for (int k = 0; k < nm2; k++) {
                double v = beta[(size_t)ij * nm2 + k];
                if (fabs(v) > 1e-13) {
                    betaVal[ptr] = v;
                    betaColInd[ptr] = k;
                    ptr++;
                }
            }
        }
        betaRowPtr[nm * nm] = ptr;
    }
    free(A);
    free(tau);
    free(b);
    return nnz;
}

/* P_N half-range flux operators, init.cpp:199-234.
 * ops = [XiPlus, XiMinus, EtaPlus, EtaMinus, MuPlus, MuMinus], each nm x nm row-major. */
void orc_pn_operators(int nm, int nq, const double *w, const double *xi, const double *eta,
                      const double *mu, const double *M, double *ops)
{
    size_t s = (size_t)nm * nm;
    for (int i = 0; i < nm; i++) {
        for (int j = 0; j < nm; j++) {
            double a[6] = {0, 0, 0, 0, 0, 0};
            for (int q = 0; q < nq; q++) {
                double mi = M[q * nm + i], mj = M[q * nm + j];
                if (xi[q] > 0)
                    a[0] += xi[q] * mi * mj * w[q];
                else
                    a[1] += xi[q] * mi * mj * w[q];
                if (eta[q] > 0)
                    a[2] += eta[q] * mi * mj * w[q];
                else
                    a[3] += eta[q] * mi * mj * w[q];
                if (mu[q] > 0)
                    a[4] += mu[q] * mi * mj * w[q];
                else
                    a[5] += mu[q] * mi * mj * w[q];
            }
            for (int o = 0; o < 6; o++)
                ops[o * s + i * nm + j] = a[o];
        }
    }
}

/* ===================================================== dense linear algebra */

/* Sequential dot, the order of reference BLAS ddot (left-to-right accumulation). */
static double orc_ddot_strided(int n, const double *x, int incx, const double *y, int incy)
{
    double s = 0.0;
    for (int i = 0; i < n; i++)
        s = s + x[i * incx] * y[i * incy];
    return s;
}

/* Cholesky 'L' of a column-major n x n matrix in place, LAPACK dpotf2 order
 * (the unblocked path dpotrf takes for n < 64; called at optBatch.cpp:199, opt.cpp:227).
 * Returns info (0, or j+1 at the first non-positive pivot). */
int orc_dpotrf_l(int n, double *A)
{
    for (int j = 0; j < n; j++) {
        double ajj = A[j + n * j] - orc_ddot_strided(j, A + j, n, A + j, n);
        if (ajj <= 0.0 || isnan(ajj)) {
            A[j + n * j] = ajj;
            return j + 1;
        }
        ajj = sqrt(ajj);
        A[j + n * j] = ajj;
        if (j < n - 1) {
            /* dgemv 'N': y(i) += (-x(jj)) * A(i,jj), columns jj ascending */
            for (int jj = 0; jj < j; jj++) {
                double temp = -A[j + n * jj];
                for (int i = j + 1; i < n; i++)
                    A[i + n * j] = A[i + n * j] + temp * A[i + n * jj];
            }
            double r = 1.0 / ajj; /* dscal */
            for (int i = j + 1; i < n; i++)
                A[i + n * j] = r * A[i + n * j];
        }
    }
    return 0;
}

static double orc_dasum(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; i++)
        s = s + fabs(x[i]);
    return s;
}

static int orc_idamax(int n, const double *x)
{
    int im = 0;
    double m = fabs(x[0]);
    for (int i = 1; i < n; i++)
        if (fabs(x[i]) > m) {
            m = fabs(x[i]);
            im = i;
        }
    return im;
}

/* dtrsv('L','N','N'): x <- L^{-1} x, column oriented (reference BLAS). */
static void orc_trsv_ln(int n, const double *L, double *x)
{
    for (int j = 0; j < n; j++) {
        if (x[j] != 0.0) {
            x[j] = x[j] / L[j + n * j];
            double temp = x[j];
            for (int i = j + 1; i < n; i++)
                x[i] = x[i] - temp * L[i + n * j];
        }
    }
}

/* dtrsv('L','T','N'): x <- L^{-T} x (reference BLAS). */
static void orc_trsv_lt(int n, const double *L, double *x)
{
    for (int j = n - 1; j >= 0; j--) {
        double temp = x[j];
        for (int i = n - 1; i > j; i--)
            temp = temp - L[i + n * j] * x[i];
        temp = temp / L[j + n * j];
        x[j] = temp;
    }
}

/* dlatrs grow bound (LAPACK dlatrs, NORMIN/TSCAL = 1 branch): returns 1 when the
 * Level-2 dtrsv path is taken.  Every well-scaled factor takes it; the careful scaled
 * path is not restated (its result equals dtrsv's whenever no rescaling is needed). */
static int orc_latrs_uses_trsv(int n, const double *L, const double *x, int trans)
{
    const double smlnum = DBL_MIN / DBL_EPSILON; /* dlamch(S)/dlamch(P) */
    double cnorm_max = 0.0;
    double *cn = (double *)alloca(n * sizeof(double));
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int i = j + 1; i < n; i++)
            s += fabs(L[i + n * j]);
        cn[j] = s;
        if (s > cnorm_max)
            cnorm_max = s;
    }
    if (cnorm_max > 1.0 / smlnum)
        return 0;
    double xmax = fabs(x[orc_idamax(n, x)]);
    double xbnd = xmax, grow;
    if (!trans) {
        grow = 1.0 / ORC_MAX(xbnd, smlnum);
        xbnd = grow;
        for (int j = 0; j < n; j++) {
            if (grow <= smlnum)
                break;
            double tjj = fabs(L[j + n * j]);
            xbnd = ORC_MIN(xbnd, ORC_MIN(1.0, tjj) * grow);
            if (tjj + cn[j] >= smlnum)
                grow = grow * (tjj / (tjj + cn[j]));
            else
                grow = 0.0;
        }
        grow = xbnd;
    } else {
        grow = 1.0 / ORC_MAX(xbnd, smlnum);
        xbnd = grow;
        for (int j = n - 1; j >= 0; j--) {
            if (grow <= smlnum)
                break;
            double xj = 1.0 + cn[j];
            grow = ORC_MIN(grow, xbnd / xj);
            double tjj = fabs(L[j + n * j]);
            if (xj > tjj)
                xbnd = xbnd * (tjj / xj);
        }
        grow = ORC_MIN(grow, xbnd);
    }
    return grow > smlnum;
}

/* dpocon('L'): reciprocal 1-norm condition estimate from the Cholesky factor
 * (optBatch.cpp:202, opt.cpp:230).  LAPACK's estimator: dlacn2 (Higham, ACM TOMS 674,
 * the Hager/Higham 1-norm power method) driving dlatrs(L) then dlatrs(L^T). */
double orc_dpocon_l(int n, const double *L, double anorm, int *info)
{
    *info = 0;
    if (n == 0)
        return 1.0;
    if (anorm == 0.0)
        return 0.0;
    double *v = (double *)alloca(n * sizeof(double));
    double *x = (double *)alloca(n * sizeof(double));
    int *isgn = (int *)alloca(n * sizeof(int));
    double est = 0.0;
    int kase = 0, isave1 = 0, isave2 = 0, isave3 = 0;
    const int itmax = 5;
    /* dlacn2 state machine; each kase != 0 return asks for x <- A^{-1} x */
    for (;;) {
        /* ---- dlacn2 ---- */
        if (kase == 0) {
            for (int i = 0; i < n; i++)
                x[i] = 1.0 / (double)n;
            kase = 1;
            isave1 = 1;
        } else {
            int jump = isave1;
            if (jump == 1) {
                if (n == 1) {
                    v[0] = x[0];
                    est = fabs(v[0]);
                    kase = 0;
                    goto done;
                }
                est = orc_dasum(n, x);
                for (int i = 0; i < n; i++) {
                    x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
                    isgn[i] = (int)lround(x[i]);
                }
                kase = 2;
                isave1 = 2;
            } else if (jump == 2) {
                isave2 = orc_idamax(n, x);
                isave3 = 2;
            L50:
                for (int i = 0; i < n; i++)
                    x[i] = 0.0;
                x[isave2] = 1.0;
                kase = 1;
                isave1 = 3;
            } else if (jump == 3) {
                memcpy(v, x, n * sizeof(double));
                double estold = est;
                est = orc_dasum(n, v);
                int converged = 1;
                for (int i = 0; i < n; i++) {
                    double xs = (x[i] >= 0.0) ? 1.0 : -1.0;
                    if ((int)lround(xs) != isgn[i]) {
                        converged = 0;
                        break;
                    }
                }
                if (converged || est <= estold)
                    goto L120;
                for (int i = 0; i < n; i++) {
                    x[i] = (x[i] >= 0.0) ? 1.0 : -1.0;
                    isgn[i] = (int)lround(x[i]);
                }
                kase = 2;
                isave1 = 4;
            } else if (jump == 4) {
                int jlast = isave2;
                isave2 = orc_idamax(n, x);
                if (x[jlast] != fabs(x[isave2]) && isave3 < itmax) {
                    isave3++;
                    goto L50;
                }
                goto L120;
            } else { /* jump == 5 */
                double temp = 2.0 * (orc_dasum(n, x) / (double)(3 * n));
                if (temp > est) {
                    memcpy(v, x, n * sizeof(double));
                    est = temp;
                }
                kase = 0;
                goto done;
            }
            goto apply;
        L120: {
            double altsgn = 1.0;
            for (int i = 0; i < n; i++) {
                x[i] = altsgn * (1.0 + (double)i / (double)(n - 1));
                altsgn = -altsgn;
            }
            kase = 1;
            isave1 = 5;
        }
        }
    apply:
        /* ---- x <- inv(L) x then inv(L^T) x (dlatrs -> dtrsv) ---- */
        if (!orc_latrs_uses_trsv(n, L, x, 0))
            *info = -100; /* careful dlatrs path would be taken (not restated) */
        orc_trsv_ln(n, L, x);
        if (!orc_latrs_uses_trsv(n, L, x, 1))
            *info = -100;
        orc_trsv_lt(n, L, x);
        continue;
    done:
        break;
    }
    if (est != 0.0)
        return (1.0 / est) / anorm;
    return 0.0;
}

/* Custom two-sided triangular solve used by the batched solver, utils.cpp:71-86. */
void orc_dpotrs_custom(int n, const double *L, double *d)
{
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < i; j++)
            d[i] -= L[i + n * j] * d[j];
        d[i] = d[i] / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
        for (int j = n - 1; j > i; j--)
            d[i] -= L[j + n * i] * d[j];
        d[i] = d[i] / L[i * n + i];
    }
}

/* LAPACK dpotrs('L', nrhs=1) as called by the unbatched solver (opt.cpp:285):
 * two dtrsm calls, reference BLAS order. */
void orc_dpotrs_lapack(int n, const double *L, double *d)
{
    /* dtrsm Left Lower NoTrans NonUnit: for k: if B(k)!=0: B(k)/=A(k,k); B(i)-=B(k)A(i,k) */
    for (int k = 0; k < n; k++) {
        if (d[k] != 0.0) {
            d[k] = d[k] / L[k + n * k];
            for (int i = k + 1; i < n; i++)
                d[i] = d[i] - d[k] * L[i + n * k];
        }
    }
    /* dtrsm Left Lower Trans NonUnit: for i=n..1: temp=B(i); for k=i+1..n: temp-=A(k,i)B(k) */
    for (int i = n - 1; i >= 0; i--) {
        double temp = d[i];
        for (int k = i + 1; k < n; k++)
            temp = temp - L[k + n * i] * d[k];
        temp = temp / L[i + n * i];
        d[i] = temp;
    }
}

/* dnrm2 (optBatch.cpp:261, opt.cpp:195,291), the scaled sum-of-squares of reference BLAS. */
double orc_dnrm2(int n, const double *x)
{
    if (n < 1)
        return 0.0;
    if (n == 1)
        return fabs(x[0]);
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        if (x[i] != 0.0) {
            double absxi = fabs(x[i]);
            if (scale < absxi) {
                double r = scale / absxi;
                ssq = 1.0 + ssq * (r * r);
                scale = absxi;
            } else {
                double r = absxi / scale;
                ssq = ssq + r * r;
            }
        }
    }
    return scale * sqrt(ssq);
}

/* utils_norm1, utils.cpp:35-49: maximum absolute row sum. */
double orc_norm1(int n, const double *A)
{
    double oneNorm = 0.0;
    for (int i = 0; i < n; i++) {
        double t = 0.0;
        for (int j = 0; j < n; j++)
            t += fabs(A[i * n + j]);
        if (t > oneNorm)
            oneNorm = t;
    }
    return oneNorm;
}

/* modcfl, optBatch.cpp:126-140 (== opt.cpp:117-132). */
double orc_modcfl(int n, const double *da)
{
    double da_1 = 0.0;
    for (int i = 0; i < n; i++)
        da_1 = da_1 + fabs(da[i]);
    int pnOrder = (int)floor(sqrt((double)n)) + 1;
    double m_inf = sqrt((2 * pnOrder + 1) / (2 * M_PI));
    return orc_exp(2.0 * da_1 * m_inf);
}

/* ========================================================== dual objective */

/* solveHMn, fobj.cpp:34-102. gExt: scratch of nm2 doubles. */
void orc_solveH(int nm, int nm2, int nq, const double *alpha, const orc_tables *t,
                int useGaunt, int useGauntSparse, double *h, double *gExt)
{
    const double *w = t->w, *p = t->M, *p2 = t->M2;
    for (int k = 0; k < nm2; k++)
        gExt[k] = 0.0;
    for (int i = 0; i < nm * nm; i++)
        h[i] = 0.0;
    for (int q = 0; q < nq; q++) {
        double temp = 0;
        for (int i = 0; i < nm; i++)
            temp = temp + alpha[i] * p[q * nm + i];
        temp = orc_exp(temp);
        if (useGaunt) {
            for (int k = 0; k < nm2; k++)
                gExt[k] += w[q] * p2[q * nm2 + k] * temp;
        } else {
            for (int i = 0; i < nm; i++) {
                double pi = p[q * nm + i];
                for (int j = 0; j <= i; j++) {
                    double pj = p[q * nm + j];
                    h[i * nm + j] += w[q] * pi * pj * temp;
                }
            }
        }
    }
    if (useGaunt && !useGauntSparse) {
        for (int i = 0; i < nm; i++)
            for (int j = i; j < nm; j++) {
                double h_ij = 0.0;
                for (int k = 0; k < nm2; k++)
                    h_ij += gExt[k] * t->beta[(size_t)(i * nm + j) * nm2 + k];
                h[i * nm + j] = h[j * nm + i] = h_ij;
            }
    } else if (useGaunt && useGauntSparse) {
        for (int i = 0; i < nm; i++)
            for (int j = i; j < nm; j++) {
                double h_ij = 0.0;
                int index = i * nm + j;
                for (int k = t->betaRowPtr[index]; k < t->betaRowPtr[index + 1]; k++)
                    h_ij += gExt[t->betaColInd[k]] * t->betaVal[k];
                h[i * nm + j] = h[j * nm + i] = h_ij;
            }
    } else {
        for (int i = 0; i < nm; i++)
            for (int j = i + 1; j < nm; j++)
                h[i * nm + j] = h[j * nm + i];
    }
}

/* fobjMn without h, fobj.cpp:135-164. g may be NULL. */
double orc_fobj(int nm, int nq, const double *alpha, const double *u, const double *w,
                const double *M, double *g)
{
    if (g)
        for (int i = 0; i < nm; i++)
            g[i] = 0.0;
    double f = 0.0;
    for (int q = 0; q < nq; q++) {
        double temp = 0;
        for (int i = 0; i < nm; i++)
            temp = temp + alpha[i] * M[q * nm + i];
        temp = orc_exp(temp);
        if (g)
            for (int i = 0; i < nm; i++)
                g[i] += w[q] * M[q * nm + i] * temp;
        f = f + w[q] * temp;
    }
    if (g)
        for (int i = 0; i < nm; i++)
            g[i] = g[i] - u[i];
    for (int i = 0; i < nm; i++)
        f = f - alpha[i] * u[i];
    return f;
}

/* fobjMnUseH, fobj.cpp:171-185 (and the h != NULL branch of fobjMn, :118-131). */
double orc_fobj_useH(int nm, const double *alpha, const double *u, double m0, double *g,
                     const double *h)
{
    if (g)
        for (int i = 0; i < nm; i++)
            g[i] = h[i] / m0 - u[i];
    double f = h[0] / (m0 * m0);
    for (int i = 0; i < nm; i++)
        f = f - alpha[i] * u[i];
    return f;
}

/* Backtracking Armijo line search, linesearch.cpp:29-81. Returns t (0 on failure,
 * alpha restored). */
double orc_linesearch(int nm, int nq, double *alpha, const double *d, double f,
                      const double *g, const double *u, const double *w, const double *M)
{
    double t = 1.0, fNew, delta = 1e-3, beta = .5, maxFactor = 0, temp;
    double alphaOld[64];
    for (int i = 0; i < nm; i++)
        alphaOld[i] = alpha[i];
    for (int i = 0; i < nm; i++) {
        temp = fabs(d[i]) / (fabs(alphaOld[i]) + 1e-100);
        if (temp > maxFactor)
            maxFactor = temp;
    }
    while (t * maxFactor > DBL_EPSILON / 2) {
        for (int i = 0; i < nm; i++)
            alpha[i] = alphaOld[i] + t * d[i];
        fNew = orc_fobj(nm, nq, alpha, u, w, M, NULL);
        temp = 0.0;
        for (int i = 0; i < nm; i++)
            temp += g[i] * d[i];
        if (fNew <= f + delta * t * temp)
            return t;
        t = beta * t;
    }
    for (int i = 0; i < nm; i++)
        alpha[i] = alphaOld[i];
    return 0.0;
}

/* ================================================================== closure */

static void orc_isotropic_alpha(int n, double *alpha) /* optBatch.cpp:115-121 */
{
    alpha[0] = 2.0 * sqrt(M_PI) * log(0.5 / sqrt(M_PI));
    for (int i = 1; i < n; i++)
        alpha[i] = 0.0;
}

/* One cell with the batched semantics: Optimization::preProcess (optBatch.cpp:312-331),
 * the H phase (solveHMn, :421) alternating with solveTheRest (:143-282) until it reports
 * finished, then postProcess (:334-344).  Unbatched semantics: opt() (opt.cpp:330-355)
 * around optScaled (opt.cpp:139-324).  Returns 0, or the ORC_FLAG_* error bit. */
int orc_closure_cell(int nm, int nm2, int nq, double *u, double *alpha, const orc_opts *o,
                     const orc_tables *t, orc_stats *s)
{
    double h[36 * 36], chol[36 * 36], g[36], d[36], alpha0[36], uOrig[36], gExt[128];
    double m0 = t->M[0];
    int flags = 0;
    double u0 = u[0];
    if (u0 <= 0) {
        /* the reference aborts (optBatch.cpp:319-323); leave the cell as given, report */
        if (s) {
            memset(s, 0, sizeof(*s));
            s->flags = ORC_FLAG_U0_NONPOS;
        }
        return ORC_FLAG_U0_NONPOS;
    }
    double alphaScale = orc_log(u0) * 2.0 * sqrt(M_PI);
    alpha[0] = alpha[0] - alphaScale;
    for (int k = 0; k < nm; k++)
        u[k] = u[k] / u0;
    memcpy(alpha0, alpha, nm * sizeof(double));
    memcpy(uOrig, u, nm * sizeof(double));

    int iter = 0, iterGamma = 0, totalIter = 0, regIndex = 0;
    double tstep = 1.0, r = 0.0, err = 0.0, gamma = 0.0;
    int maxIter = o->maxIter;
    const int unb = o->unbatched;
    double tolGamma = o->tolGamma; /* opt.cpp:149; update() passes 1.1 when unbatched */
    double rtol = unb ? orc_dnrm2(nm, u) * o->tol : o->tol;

    for (;;) {
        /* H phase */
        orc_solveH(nm, nm2, nq, alpha, t, o->useGaunt, o->useGauntSparse, h, gExt);
        /* ---- solveTheRest / one optScaled iteration ---- */
        int maxIt = maxIter;
        if (!unb && regIndex == ORC_NUM_REG - 1)
            maxIt = 10 * maxIter; /* optBatch.cpp:183-184 */
        iter++;
        totalIter++;
        double f = orc_fobj_useH(nm, alpha, u, m0, g, h);
        memcpy(chol, h, (size_t)nm * nm * sizeof(double));
        int cholErr = orc_dpotrf_l(nm, chol);
        double norm1 = orc_norm1(nm, h);
        double condest = 0.0;
        if (cholErr == 0) {
            int cinfo;
            double rc = orc_dpocon_l(nm, chol, norm1, &cinfo);
            condest = 1.0 / rc;
        }
        if (totalIter == 1 && (cholErr != 0 || condest > o->condHMax)) {
            orc_isotropic_alpha(nm, alpha0);
            if (unb)
                rtol = orc_dnrm2(nm, u) * o->tol;
            memcpy(alpha, alpha0, nm * sizeof(double));
            iter = 0;
            iterGamma = 0;
            tstep = 1;
            flags |= ORC_FLAG_ISO_RESTART;
            continue;
        }
        if (iter >= maxIt || cholErr != 0 || tstep == 0 || condest > o->condHMax) {
            if (regIndex < ORC_NUM_REG - 1) {
                regIndex++;
                r = orc_reg[regIndex];
                u[0] = 1.0;
                for (int k = 1; k < nm; k++)
                    u[k] = uOrig[k] * (1.0 - r);
            } else {
                flags |= ORC_FLAG_LADDER;
                break;
            }
            if (unb)
                rtol = orc_dnrm2(nm, u) * o->tol;
            memcpy(alpha, alpha0, nm * sizeof(double));
            iterGamma = 0;
            iter = 0;
            tstep = 1;
            if (unb && regIndex == ORC_NUM_REG - 1)
                maxIter = 10 * maxIter; /* opt.cpp:275-276 */
            continue;
        }
        for (int i = 0; i < nm; i++)
            d[i] = g[i];
        if (unb)
            orc_dpotrs_lapack(nm, chol, d);
        else
            orc_dpotrs_custom(nm, chol, d);
        for (int i = 0; i < nm; i++)
            d[i] = -d[i];
        err = orc_dnrm2(nm, g);
        gamma = orc_modcfl(nm, d);
        if (err <= rtol && gamma <= tolGamma)
            break;
        else if (err <= rtol && gamma > tolGamma)
            iterGamma++;
        tstep = orc_linesearch(nm, nq, alpha, d, f, g, u, t->w, t->M);
    }
    /* postProcess (optBatch.cpp:334-344) / opt() rescale (opt.cpp:351-354) */
    double alphaScale2 = orc_log(u0) * 2.0 * sqrt(M_PI);
    for (int k = 0; k < nm; k++)
        u[k] = u[k] * u0;
    alpha[0] = alpha[0] + alphaScale2;
    if (s) {
        s->r = r;
        s->err = err;
        s->gamma = gamma;
        s->totalIter = totalIter;
        s->iterGamma = iterGamma;
        s->regIndex = regIndex;
        s->flags = flags;
    }
    return flags & ORC_FLAG_LADDER;
}

/* The closure over a batch of cells (update.cpp:105-178 with Optimization::opt; or
 * update.cpp:43-99 with opt()).  Cells are independent, so the stack interleaving of
 * optBatch.cpp:371-468 does not change any cell's result.  Returns the OR of the
 * error flags. */
int orc_closure(int64_t ncells, int nm, int nm2, int nq, double *u, double *alpha,
                const orc_opts *o, const orc_tables *t, orc_stats *stats, int nthreads)
{
    int errs = 0;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads) reduction(| : errs)
#endif
    for (int64_t c = 0; c < ncells; c++) {
        errs |= orc_closure_cell(nm, nm2, nq, u + c * nm, alpha + c * nm, o, t,
                                 stats ? stats + c : NULL);
    }
    (void)nthreads;
    return errs;
}

/* ============================================================ grid kernels */

#define NG (2 * ORC_NUM_GHOST)
#define I3(i, j, k) ((size_t)(k) + (size_t)(nz + NG) * ((size_t)(j) + (size_t)(ny + NG) * (size_t)(i)))

/* Serial periodic ghost fill, comm.cpp:281-323 (faces only, 2 layers). */
void orc_ghost_fill(int nx, int ny, int nz, int nm, double *u)
{
    int gX1 = 2, gX2 = nx + 1, gY1 = 2, gY2 = ny + 1, gZ1 = 2, gZ2 = nz + 1;
    int x0 = 0, x1 = gX1, x2 = gX2 + 1 - 2, x3 = nx + 3 + 1 - 2;
    int y0 = 0, y1 = gY1, y2 = gY2 + 1 - 2, y3 = ny + 3 + 1 - 2;
    int z0 = 0, z1 = gZ1, z2 = gZ2 + 1 - 2, z3 = nz + 3 + 1 - 2;
    for (int i = 0; i < 2; i++)
        for (int j = gY1; j <= gY2; j++)
            for (int k = gZ1; k <= gZ2; k++)
                for (int m = 0; m < nm; m++) {
                    u[I3(x0 + i, j, k) * nm + m] = u[I3(x2 + i, j, k) * nm + m];
                    u[I3(x3 + i, j, k) * nm + m] = u[I3(x1 + i, j, k) * nm + m];
                }
    for (int i = gX1; i <= gX2; i++)
        for (int j = 0; j < 2; j++)
            for (int k = gZ1; k <= gZ2; k++)
                for (int m = 0; m < nm; m++) {
                    u[I3(i, y0 + j, k) * nm + m] = u[I3(i, y2 + j, k) * nm + m];
                    u[I3(i, y3 + j, k) * nm + m] = u[I3(i, y1 + j, k) * nm + m];
                }
    for (int i = gX1; i <= gX2; i++)
        for (int j = gY1; j <= gY2; j++)
            for (int k = 0; k < 2; k++)
                for (int m = 0; m < nm; m++) {
                    u[I3(i, j, z0 + k) * nm + m] = u[I3(i, j, z2 + k) * nm + m];
                    u[I3(i, j, z3 + k) * nm + m] = u[I3(i, j, z1 + k) * nm + m];
                }
}

static inline double orc_minmod(double x, double y) /* flux_kinetic.cpp:45-49 */
{
    return ORC_SIGN(1.0, x) * ORC_MAX(0.0, ORC_MIN(fabs(x), y * ORC_SIGN(1.0, x)));
}

static inline double orc_slopefit(double l, double c, double r, double theta) /* :55-60 */
{
    return orc_minmod(theta * (r - c), orc_minmod(0.5 * (r - l), theta * (c - l)));
}

/* Kinetic flux, flux_kinetic.cpp:66-198 (CPU branch).  ansatz: scratch of N_g doubles.
 * flux is written on interior cells only. */
void orc_flux_kinetic(int nx, int ny, int nz, int nm, int nq, const double *alpha,
                      double *flux, const orc_tables *t, double dx, double dy, double dz,
                      double theta, int isPn, double *ansatz, int nthreads)
{
    const double *M = t->M, *xi = t->xi, *eta = t->eta, *mu = t->mu, *w = t->w;
    int NX = nx + NG, NY = ny + NG, NZ = nz + NG;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#endif
    (void)nthreads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads)
#endif
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 2; j <= ny + 1; j++)
            for (int k = 2; k <= nz + 1; k++)
                for (int m = 0; m < nm; m++)
                    flux[I3(i, j, k) * nm + m] = 0.0;
    for (int q = 0; q < nq; q++) {
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads)
#endif
        for (int i = 0; i < NX; i++)
            for (int j = 0; j < NY; j++)
                for (int k = 0; k < NZ; k++) {
                    const double *a = alpha + I3(i, j, k) * nm;
                    double kin = 0.0;
                    for (int m = 0; m < nm; m++)
                        kin += a[m] * M[q * nm + m];
                    ansatz[I3(i, j, k)] = isPn ? kin : orc_exp(kin);
                }
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthreads)
#endif
        for (int i = 2; i <= nx + 1; i++)
            for (int j = 2; j <= ny + 1; j++)
                for (int k = 2; k <= nz + 1; k++) {
                    double k1, k2, k3, k4, fX, fY, fZ;
                    if (xi[q] > 0) {
                        k1 = ansatz[I3(i - 2, j, k)];
                        k2 = ansatz[I3(i - 1, j, k)];
                        k3 = ansatz[I3(i, j, k)];
                        k4 = ansatz[I3(i + 1, j, k)];
                        fX = k3 + 0.5 * orc_slopefit(k2, k3, k4, theta) - k2 -
                             0.5 * orc_slopefit(k1, k2, k3, theta);
                    } else {
                        k1 = ansatz[I3(i - 1, j, k)];
                        k2 = ansatz[I3(i, j, k)];
                        k3 = ansatz[I3(i + 1, j, k)];
                        k4 = ansatz[I3(i + 2, j, k)];
                        fX = k3 - 0.5 * orc_slopefit(k2, k3, k4, theta) - k2 +
                             0.5 * orc_slopefit(k1, k2, k3, theta);
                    }
                    fX = fX * xi[q] * w[q];
                    if (eta[q] > 0) {
                        k1 = ansatz[I3(i, j - 2, k)];
                        k2 = ansatz[I3(i, j - 1, k)];
                        k3 = ansatz[I3(i, j, k)];
                        k4 = ansatz[I3(i, j + 1, k)];
                        fY = k3 + 0.5 * orc_slopefit(k2, k3, k4, theta) - k2 -
                             0.5 * orc_slopefit(k1, k2, k3, theta);
                    } else {
                        k1 = ansatz[I3(i, j - 1, k)];
                        k2 = ansatz[I3(i, j, k)];
                        k3 = ansatz[I3(i, j + 1, k)];
                        k4 = ansatz[I3(i, j + 2, k)];
                        fY = k3 - 0.5 * orc_slopefit(k2, k3, k4, theta) - k2 +
                             0.5 * orc_slopefit(k1, k2, k3, theta);
                    }
                    fY = fY * eta[q] * w[q];
                    if (mu[q] > 0) {
                        k1 = ansatz[I3(i, j, k - 2)];
                        k2 = ansatz[I3(i, j, k - 1)];
                        k3 = ansatz[I3(i, j, k)];
                        k4 = ansatz[I3(i, j, k + 1)];
                        fZ = k3 + 0.5 * orc_slopefit(k2, k3, k4, theta) - k2 -
                             0.5 * orc_slopefit(k1, k2, k3, theta);
                    } else {
                        k1 = ansatz[I3(i, j, k - 1)];
                        k2 = ansatz[I3(i, j, k)];
                        k3 = ansatz[I3(i, j, k + 1)];
                        k4 = ansatz[I3(i, j, k + 2)];
                        fZ = k3 - 0.5 * orc_slopefit(k2, k3, k4, theta) - k2 +
                             0.5 * orc_slopefit(k1, k2, k3, theta);
                    }
                    fZ = fZ * mu[q] * w[q];
                    double *fl = flux + I3(i, j, k) * nm;
                    for (int m = 0; m < nm; m++)
                        fl[m] += M[q * nm + m] * (fX / dx + fY / dy + fZ / dz);
                }
    }
}

static void orc_mv(int n, double beta, const double *A, const double *x, double *b)
{ /* flux_moment.cpp:15-24 */
    for (int i = 0; i < n; i++)
        b[i] = beta * b[i];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            b[i] += A[i * n + j] * x[j];
}

/* P_N moment flux, flux_moment.cpp:30-104. ops as orc_pn_operators. */
void orc_flux_moment(int nx, int ny, int nz, int nm, const double *u, double *flux,
                     const double *ops, double dx, double dy, double dz, int nthreads)
{
    size_t s = (size_t)nm * nm;
#ifdef _OPENMP
    if (nthreads <= 0)
        nthreads = omp_get_max_threads();
#pragma omp parallel for num_threads(nthreads)
#endif
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 2; j <= ny + 1; j++)
            for (int k = 2; k <= nz + 1; k++) {
                double temp[64], xF[64], yF[64], zF[64];
#define U(a, b, c, m) u[I3(a, b, c) * nm + (m)]
                for (int m = 0; m < nm; m++)
                    temp[m] = 0.25 * U(i + 1, j, k, m) + 0.75 * U(i, j, k, m) -
                              1.25 * U(i - 1, j, k, m) + 0.25 * U(i - 2, j, k, m);
                orc_mv(nm, 0.0, ops + 0 * s, temp, xF);
                for (int m = 0; m < nm; m++)
                    temp[m] = -0.25 * U(i + 2, j, k, m) + 1.25 * U(i + 1, j, k, m) -
                              0.75 * U(i, j, k, m) - 0.25 * U(i - 1, j, k, m);
                orc_mv(nm, 1.0, ops + 1 * s, temp, xF);
                for (int m = 0; m < nm; m++)
                    temp[m] = 0.25 * U(i, j + 1, k, m) + 0.75 * U(i, j, k, m) -
                              1.25 * U(i, j - 1, k, m) + 0.25 * U(i, j - 2, k, m);
                orc_mv(nm, 0.0, ops + 2 * s, temp, yF);
                for (int m = 0; m < nm; m++)
                    temp[m] = -0.25 * U(i, j + 2, k, m) + 1.25 * U(i, j + 1, k, m) -
                              0.75 * U(i, j, k, m) - 0.25 * U(i, j - 1, k, m);
                orc_mv(nm, 1.0, ops + 3 * s, temp, yF);
                for (int m = 0; m < nm; m++)
                    temp[m] = 0.25 * U(i, j, k + 1, m) + 0.75 * U(i, j, k, m) -
                              1.25 * U(i, j, k - 1, m) + 0.25 * U(i, j, k - 2, m);
                orc_mv(nm, 0.0, ops + 4 * s, temp, zF);
                for (int m = 0; m < nm; m++)
                    temp[m] = -0.25 * U(i, j, k + 2, m) + 1.25 * U(i, j, k + 1, m) -
                              0.75 * U(i, j, k, m) - 0.25 * U(i, j, k - 1, m);
                orc_mv(nm, 1.0, ops + 5 * s, temp, zF);
#undef U
                for (int m = 0; m < nm; m++)
                    flux[I3(i, j, k) * nm + m] = xF[m] / dx + yF[m] / dy + zF[m] / dz;
            }
    (void)nthreads;
}

/* Euler stage with isotropic scattering, update.cpp:21-37. */
void orc_euler(int nx, int ny, int nz, int nm, double *u, const double *flux,
               const double *sigmaT, const double *sigmaS, double dt)
{
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 2; j <= ny + 1; j++)
            for (int k = 2; k <= nz + 1; k++) {
                size_t c = I3(i, j, k);
                u[c * nm] = u[c * nm] * (1.0 + dt * sigmaS[c] - dt * sigmaT[c]) -
                            dt * flux[c * nm];
                for (int m = 1; m < nm; m++)
                    u[c * nm + m] = u[c * nm + m] * (1.0 - dt * sigmaT[c]) -
                                    dt * flux[c * nm + m];
            }
}

/* One SSP-RK2 step, update.cpp:185-275: uOld <- u (interior); two stages of
 * [ghost fill, closure over all cells (M_N), flux, Euler]; u <- (u + uOld)/2. */
int orc_update(int nx, int ny, int nz, int nm, int nm2, int nq, int solverMn,
               int pnUseKinetic, double *u, double *alpha, double *uOld, double *flux,
               double *ansatz, const double *sigmaT, const double *sigmaS,
               const orc_tables *t, const double *ops, const orc_opts *o, double dx,
               double dy, double dz, double theta, double dt, orc_stats *stats, int nthreads)
{
    int errs = 0;
    int64_t ng = (int64_t)(nx + NG) * (ny + NG) * (nz + NG);
    orc_opts oc = *o;
    if (oc.unbatched)
        oc.tolGamma = 1.1; /* solveOptimization passes 1.1 to opt(), update.cpp:57 */
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 2; j <= ny + 1; j++)
            for (int k = 2; k <= nz + 1; k++)
                for (int m = 0; m < nm; m++)
                    uOld[I3(i, j, k) * nm + m] = u[I3(i, j, k) * nm + m];
    for (int st = 0; st < 2; st++) {
        orc_ghost_fill(nx, ny, nz, nm, u);
        if (solverMn) {
            errs |= orc_closure(ng, nm, nm2, nq, u, alpha, &oc, t,
                                stats ? stats + st * ng : NULL, nthreads);
            orc_flux_kinetic(nx, ny, nz, nm, nq, alpha, flux, t, dx, dy, dz, theta, 0, ansatz,
                             nthreads);
        } else if (!pnUseKinetic) {
            orc_flux_moment(nx, ny, nz, nm, u, flux, ops, dx, dy, dz, nthreads);
        } else {
            orc_flux_kinetic(nx, ny, nz, nm, nq, u, flux, t, dx, dy, dz, theta, 1, ansatz,
                             nthreads);
        }
        orc_euler(nx, ny, nz, nm, u, flux, sigmaT, sigmaS, dt);
    }
    for (int i = 2; i <= nx + 1; i++)
        for (int j = 2; j <= ny + 1; j++)
            for (int k = 2; k <= nz + 1; k++)
                for (int m = 0; m < nm; m++)
                    u[I3(i, j, k) * nm + m] = 0.5 * (u[I3(i, j, k) * nm + m] +
                                                     uOld[I3(i, j, k) * nm + m]);
    return errs;
}

/* ====================================================== initial conditions */

/* SplitMix64 (Steele, Lea, Flood 2014): the counter RNG behind every synthetic input. */
uint64_t orc_splitmix64(uint64_t x)
{
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform in [0,1): 53 high bits of splitmix64(splitmix64(seed ^ stream-salt) + idx) */
double orc_uniform(uint64_t seed, uint64_t idx, uint64_t stream)
{
    uint64_t key = orc_splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
    uint64_t z = orc_splitmix64(key + idx);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* standard normal by Box-Muller on streams 2s and 2s+1 */
double orc_normal(uint64_t seed, uint64_t idx, uint64_t stream)
{
    double u1 = orc_uniform(seed, idx, 2 * stream);
    double u2 = orc_uniform(seed, idx, 2 * stream + 1);
    return sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
}

/* Initial state over all cells including ghosts, init.cpp:480-506.
 *   icKind 0: reference Gaussian, u0 = MAX(floor, G_sigma(x) * 2 sqrt(pi)) with the
 *             normalised G (init.cpp:27-34, 493-495)
 *   icKind 1: pulse, rho = max(exp(-|x|^2/(2 sigma^2)), floor) * (1 + noise * U[-1,1]),
 *             u0 = 2 sqrt(pi) rho   (the time-stepping workloads)
 *   icKind 2: smooth (init.cpp:497-502)
 *   icKind 3: funky (init.cpp:42-66, 503-504) on the box [a, a + n d]
 * Cell centres follow init.cpp:483-485 (ghosts at their own, unwrapped centres). */
void orc_init_grid(int nx, int ny, int nz, int nm, double ax, double ay, double az,
                   double dx, double dy, double dz, int icKind, double sigma, double floorv,
                   double noise, uint64_t seed, double crossSection, double *u,
                   double *sigmaT, double *sigmaS)
{
    for (int i = 0; i < nx + NG; i++)
        for (int j = 0; j < ny + NG; j++)
            for (int k = 0; k < nz + NG; k++) {
                double x = ax + (i - 2) * dx + dx / 2.0;
                double y = ay + (j - 2) * dy + dy / 2.0;
                double z = az + (k - 2) * dz + dz / 2.0;
                size_t c = I3(i, j, k);
                sigmaT[c] = crossSection;
                sigmaS[c] = crossSection;
                for (int m = 1; m < nm; m++)
                    u[c * nm + m] = 0.0;
                if (icKind == 0) {
                    double s2 = sigma * sigma;
                    double f1 = 1.0 / (sigma * sqrt(2.0 * M_PI));
                    double f3 = f1 * f1 * f1;
                    double gsn = f3 * exp(-(x * x + y * y + z * z) / (2.0 * s2));
                    u[c * nm] = ORC_MAX(floorv, gsn * 2.0 * sqrt(M_PI));
                } else if (icKind == 1) {
                    double rho = exp(-(x * x + y * y + z * z) / (2.0 * sigma * sigma));
                    rho = ORC_MAX(rho, floorv);
                    if (noise != 0.0)
                        rho = rho * (1.0 + noise * (2.0 * orc_uniform(seed, c, 7) - 1.0));
                    u[c * nm] = 2.0 * sqrt(M_PI) * rho;
                } else if (icKind == 3) {
                    /* init.cpp:42-66 with the box [a, a + n d]: 125 Gaussians centred on the
                       odd tenths of the box, plus the smooth state */
                    double wx = nx * dx, wy = ny * dy, wz = nz * dz, sum = 0.0;
                    double s2 = sigma * sigma;
                    double f1 = 1.0 / (sigma * sqrt(2.0 * M_PI));
                    double f3 = f1 * f1 * f1;
                    for (int gi = 1; gi < 10; gi += 2)
                        for (int gj = 1; gj < 10; gj += 2)
                            for (int gk = 1; gk < 10; gk += 2) {
                                double px = x - (ax + gi * wx / 10.0);
                                double py = y - (ay + gj * wy / 10.0);
                                double pz = z - (az + gk * wz / 10.0);
                                sum += f3 * exp(-(px * px + py * py + pz * pz) / (2.0 * s2));
                            }
                    u[c * nm] = (2.0 + cos(2 * M_PI * x) * cos(2 * M_PI * y) *
                                           cos(2 * M_PI * z) + sum) * 2.0 * sqrt(M_PI);
                } else {
                    u[c * nm] = (2.0 + cos(2 * M_PI * x) * cos(2 * M_PI * y) *
                                           cos(2 * M_PI * z)) * 2.0 * sqrt(M_PI);
                    if (nm > 3) {
                        u[c * nm + 1] = 0.3 * u[c * nm];
                        u[c * nm + 2] = 0.3 * u[c * nm];
                        u[c * nm + 3] = 0.3 * u[c * nm];
                    }
                }
            }
}

/* alpha initialisation, init.cpp:262-268. */
void orc_init_alpha(int64_t ncells, int nm, const double *u, double *alpha)
{
    for (int64_t c = 0; c < ncells; c++) {
        alpha[c * nm] = sqrt(4.0 * M_PI) * log(u[c * nm] / sqrt(4.0 * M_PI));
        for (int m = 1; m < nm; m++)
            alpha[c * nm + m] = 0.0;
    }
}

/* Closure-only batch: per cell s ~ U[smin,smax], alpha*_k ~ N(0, s^2),
 * u = sum_q w_q m_q exp(alpha*.m_q), Newton start alpha0 from init.cpp:265. */
void orc_closure_batch_input(int64_t ncells, int nm, int nq, const double *w, const double *M,
                             uint64_t seed, double smin, double smax, double *alphaStar,
                             double *u, double *alpha0)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int64_t c = 0; c < ncells; c++) {
        double s = smin + (smax - smin) * orc_uniform(seed, (uint64_t)c, 1);
        double *as = alphaStar + c * nm;
        for (int k = 0; k < nm; k++)
            as[k] = s * orc_normal(seed, (uint64_t)c * 64 + k, 2);
        double *uc = u + c * nm;
        for (int k = 0; k < nm; k++)
            uc[k] = 0.0;
        for (int q = 0; q < nq; q++) {
            double e = 0.0;
            for (int k = 0; k < nm; k++)
                e = e + as[k] * M[q * nm + k];
            e = w[q] * exp(e);
            for (int k = 0; k < nm; k++)
                uc[k] += e * M[q * nm + k];
        }
        alpha0[c * nm] = sqrt(4.0 * M_PI) * log(uc[0] / sqrt(4.0 * M_PI));
        for (int k = 1; k < nm; k++)
            alpha0[c * nm + k] = 0.0;
    }
}
