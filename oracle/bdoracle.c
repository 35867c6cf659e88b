/*
 * bdoracle.c -- CPU FP64 restatement of BROWNPAK's Brownian-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the shared
 * library built from this file.
 *
 * Build: see oracle/Makefile.  It is compiled with -ffp-contract=off so that,
 * like the reference's x86-64 gfortran -O3 build (build/Makefile:18-24, no
 * -march, hence no FMA), every product and sum is separately rounded.
 *
 * Each routine cites the reference routine it restates.  Dense LAPACK
 * operations the reference delegates to MKL (dpotrf, dstevd, dsbevd, dgeqp3,
 * dorgqr) are restated from their published algorithms:
 *   - dpotrf('U'): unblocked upper Cholesky (LAPACK dpotf2 semantics);
 *   - dgeqp3/dorgqr: Householder QR with column pivoting and partial column
 *     norm downdating (LAPACK dlaqp2/dlarfg/dorg2r, Quintana-Orti et al. 1998);
 *   - dstevd/dsbevd: the band matrix is expanded to dense and diagonalised by
 *     Householder tridiagonalisation followed by implicit QL with Wilkinson
 *     shifts (Bowdler, Martin, Reinsch & Wilkinson 1968).  The Lanczos result
 *     P*sqrt(L)*P^T*E1*R does not depend on eigenvector signs or on the basis
 *     chosen inside a degenerate eigenspace, so any exact eigensolver yields
 *     the same B.xi up to rounding.
 */
#include "bdoracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

static double nrm2(int n, const double *x)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

static double dot(int n, const double *x, const double *y)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

/* ===================================================================== */
/* Mobility                                                               */
/* ===================================================================== */

/* Host threads for the two O(n^2) routines below (orc_rpy, orc_symv_lower).
 * 1 (the default, and what the tests use) runs the serial code; more runs
 * an OpenMP version for timing the restatement on all host cores: the same
 * RPY entries, and a symv whose sums are added in another order. */
static int g_threads = 1;
void orc_set_threads(int t) { g_threads = t > 1 ? t : 1; }

/* calc_rpy_tensor, bd_solver_m.f90:458-522.  RPY tensor for bead radius 1,
 * overlap branch for r < 2, unit diagonal.  Full n x n column-major. */
void orc_rpy(int nb, const double *crd, double *mob)
{
    const int n = 3 * nb;
    memset(mob, 0, sizeof(double) * (size_t)n * n);
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1)
    for (int j = 1; j < nb; j++) {
        const double *rj = crd + 3 * j;
        for (int i = 0; i < j; i++) {
            const double *ri = crd + 3 * i;
            double rij[3] = {rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2]};
            double rijm = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
            double irijm = 1.0 / rijm;
            double irijm2 = irijm * irijm;
            double c1, c2, consij;
            if (rijm >= 2.0) {
                c1 = 1.0 + (2.0 / 3.0) * irijm2;
                c2 = 1.0 - 2.0 * irijm2;
                consij = 0.75 * irijm;
            } else {
                c1 = 1.0 - 9.0 * rijm / 32.0;
                c2 = 3.0 * rijm / 32.0;
                consij = 1.0;
            }
            for (int b = 0; b < 3; b++) {
                for (int a = 0; a < 3; a++) {
                    double v = c2 * rij[a] * rij[b] * irijm2;
                    if (a == b) v = c1 + v;
                    mob[IDX(3 * i + a, 3 * j + b, n)] = consij * v;
                }
            }
        }
    }
    /* strictly-upper -> strictly-lower copy, bd_solver_m.f90:511-516 */
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads) if (g_threads > 1)
    for (int j = 1; j < n; j++)
        for (int i = 0; i < j; i++) mob[IDX(j, i, n)] = mob[IDX(i, j, n)];
    for (int j = 0; j < n; j++) mob[IDX(j, j, n)] = 1.0;
}

/* dsymv('L') as called at bd_solver_m.f90:446 and brown_m.f90:232:
 * y := alpha*A*x + beta*y reading only the lower triangle of A. */
/* The smallest j in [0, n] whose columns 0..j-1 hold at least the fraction
 * frac of the lower triangle (j (2n - j + 1) / 2 of n (n + 1) / 2 entries). */
static int tri_split(int n, double frac)
{
    const double target = frac * 0.5 * (double)n * (n + 1.0);
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (0.5 * mid * (2.0 * n - mid + 1.0) >= target) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

static void symv_lower_threads(int n, double alpha, const double *a, const double *x, double beta, double *y)
{
    /* each thread takes a range of columns of about equal triangle area and
     * accumulates its own copy of the row sums; the copies are then added */
    const int nt = g_threads;
    double *t = (double *)calloc((size_t)n * nt, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        const int id = omp_get_thread_num(), m = omp_get_num_threads();
        const int j0 = tri_split(n, (double)id / m), j1 = id == m - 1 ? n : tri_split(n, (double)(id + 1) / m);
        double *tt = t + (size_t)id * n;
        for (int j = j0; j < j1; j++) {
            const double *col = a + (size_t)j * n;
            const double xj = x[j];
            double acc = col[j] * xj;
            for (int i = j + 1; i < n; i++) {
                tt[i] += col[i] * xj;
                acc += col[i] * x[i];
            }
            tt[j] += acc;
        }
#pragma omp barrier
#pragma omp for schedule(static)
        for (int i = 0; i < n; i++) {
            double v = 0.0;
            for (int q = 0; q < m; q++) v += t[(size_t)q * n + i];
            y[i] = alpha * v + (beta == 0.0 ? 0.0 : beta * y[i]);
        }
    }
    free(t);
}

void orc_symv_lower(int n, double alpha, const double *a, const double *x,
                    double beta, double *y)
{
    if (g_threads > 1) {
        symv_lower_threads(n, alpha, a, x, beta, y);
        return;
    }
    double *t = (double *)calloc((size_t)n, sizeof(double));
    for (int j = 0; j < n; j++) {
        const double *col = a + (size_t)j * n;
        double xj = x[j];
        double acc = col[j] * xj;          /* diagonal */
        for (int i = j + 1; i < n; i++) {  /* column j below diagonal */
            t[i] += col[i] * xj;           /* A(i,j) x_j into row i   */
            acc += col[i] * x[i];          /* A(j,i) = A(i,j)          */
        }
        t[j] += acc;
    }
    for (int i = 0; i < n; i++) y[i] = alpha * t[i] + (beta == 0.0 ? 0.0 : beta * y[i]);
    free(t);
}

/* ===================================================================== */
/* Small dense kernels with LAPACK semantics                              */
/* ===================================================================== */

/* dlarfg: generate H = I - tau*[1;v][1;v]^T with H*[alpha;x] = [beta;0]. */
static void larfg(int n, double *alpha, double *x, int incx, double *tau)
{
    if (n <= 1) { *tau = 0.0; return; }
    double xn = 0.0;
    for (int i = 0; i < n - 1; i++) xn += x[i * incx] * x[i * incx];
    xn = sqrt(xn);
    if (xn == 0.0) { *tau = 0.0; return; }
    double beta = -copysign(hypot(*alpha, xn), *alpha);
    *tau = (beta - *alpha) / beta;
    double sc = 1.0 / (*alpha - beta);
    for (int i = 0; i < n - 1; i++) x[i * incx] *= sc;
    *alpha = beta;
}

/* dgeqp3 (unblocked path, dlaqp2): QR with column pivoting of the m x n
 * column-major a; jpvt returns 1-based pivots (all columns free on entry, as
 * brown_m.f90:358,410 zero iwork). */
void orc_geqp3(int m, int n, double *a, int *jpvt, double *tau)
{
    const double tol3z = sqrt(DBL_EPSILON);
    double *vn1 = (double *)malloc(sizeof(double) * n);
    double *vn2 = (double *)malloc(sizeof(double) * n);
    double *work = (double *)malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) {
        jpvt[j] = j + 1;
        vn1[j] = nrm2(m, a + IDX(0, j, m));
        vn2[j] = vn1[j];
    }
    int k = m < n ? m : n;
    for (int i = 0; i < k; i++) {
        int pvt = i;
        for (int j = i + 1; j < n; j++)
            if (fabs(vn1[j]) > fabs(vn1[pvt])) pvt = j;
        if (pvt != i) {
            for (int r = 0; r < m; r++) {
                double t = a[IDX(r, pvt, m)];
                a[IDX(r, pvt, m)] = a[IDX(r, i, m)];
                a[IDX(r, i, m)] = t;
            }
            int it = jpvt[pvt]; jpvt[pvt] = jpvt[i]; jpvt[i] = it;
            vn1[pvt] = vn1[i];
            vn2[pvt] = vn2[i];
        }
        /* reflector for a(i:m, i) */
        larfg(m - i, &a[IDX(i, i, m)], &a[IDX(i + 1, i, m)], 1, &tau[i]);
        if (i < n - 1) {
            double aii = a[IDX(i, i, m)];
            a[IDX(i, i, m)] = 1.0;
            /* apply H^T from the left to a(i:m, i+1:n) */
            for (int j = i + 1; j < n; j++) {
                double s = 0.0;
                for (int r = i; r < m; r++) s += a[IDX(r, i, m)] * a[IDX(r, j, m)];
                work[j] = s;
            }
            for (int j = i + 1; j < n; j++) {
                double t = tau[i] * work[j];
                for (int r = i; r < m; r++) a[IDX(r, j, m)] -= t * a[IDX(r, i, m)];
            }
            a[IDX(i, i, m)] = aii;
        }
        /* update partial column norms */
        for (int j = i + 1; j < n; j++) {
            if (vn1[j] != 0.0) {
                double temp = fabs(a[IDX(i, j, m)]) / vn1[j];
                temp = 1.0 - temp * temp;
                if (temp < 0.0) temp = 0.0;
                double r = vn1[j] / vn2[j];
                double temp2 = temp * r * r;
                if (temp2 <= tol3z) {
                    if (i < m - 1) {
                        vn1[j] = nrm2(m - i - 1, a + IDX(i + 1, j, m));
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] *= sqrt(temp);
                }
            }
        }
    }
    free(vn1); free(vn2); free(work);
}

/* dorgqr (dorg2r semantics): form the m x n Q from the reflectors. */
void orc_orgqr(int m, int n, double *a, const double *tau)
{
    double *work = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    for (int j = n - 1; j >= 0; j--) {
        if (j < n - 1) {
            a[IDX(j, j, m)] = 1.0;
            for (int c = j + 1; c < n; c++) {
                double s = 0.0;
                for (int r = j; r < m; r++) s += a[IDX(r, j, m)] * a[IDX(r, c, m)];
                work[c] = s;
            }
            for (int c = j + 1; c < n; c++) {
                double t = tau[j] * work[c];
                for (int r = j; r < m; r++) a[IDX(r, c, m)] -= t * a[IDX(r, j, m)];
            }
        }
        if (j < m - 1)
            for (int r = j + 1; r < m; r++) a[IDX(r, j, m)] *= -tau[j];
        a[IDX(j, j, m)] = 1.0 - tau[j];
        for (int r = 0; r < j; r++) a[IDX(r, j, m)] = 0.0;
    }
    free(work);
}

/* Implicit QL with Wilkinson shifts on the symmetric tridiagonal (d, e),
 * e[i] couples i and i+1.  z (n x n col-major) is multiplied by the
 * rotations.  Returns 0 on success, k>0 if eigenvalue k failed to converge. */
static int tql(int n, double *d, double *e, double *z)
{
    if (n == 1) return 0;
    e[n - 1] = 0.0;
    for (int l = 0; l < n; l++) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= DBL_EPSILON * dd) break;
            }
            if (m != l) {
                if (iter++ == 90) return l + 1;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + copysign(r, g));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                int early = 0;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i];
                    double b = c * e[i];
                    r = hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        early = 1;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    double *zi = z + (size_t)i * n, *zi1 = z + (size_t)(i + 1) * n;
                    for (int k = 0; k < n; k++) {
                        double t = zi1[k];
                        zi1[k] = s * zi[k] + c * t;
                        zi[k] = c * zi[k] - s * t;
                    }
                }
                if (early) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
    return 0;
}

/* Dense symmetric eigendecomposition: a (n x n col-major, both triangles
 * valid) is replaced by the eigenvectors, w receives the eigenvalues.
 * Householder tridiagonalisation then implicit QL. */
int orc_syev(int n, double *a, double *w)
{
    double *v = (double *)calloc((size_t)n * n, sizeof(double)); /* reflectors */
    double *vn = (double *)calloc((size_t)n, sizeof(double));
    double *p = (double *)calloc((size_t)n, sizeof(double));
    double *e = (double *)calloc((size_t)n, sizeof(double));
    int *has = (int *)calloc((size_t)n, sizeof(int));
    for (int k = 0; k + 2 < n; k++) {
        int len = n - k - 1;
        double *x = vn; /* x = a(k+1:n, k) */
        for (int i = 0; i < len; i++) x[i] = a[IDX(k + 1 + i, k, n)];
        double xnorm = nrm2(len, x);
        double tail = nrm2(len - 1, x + 1);
        if (tail == 0.0) continue; /* already tridiagonal in this column */
        double alpha = -copysign(xnorm, x[0]);
        x[0] -= alpha;
        double vnorm = nrm2(len, x);
        for (int i = 0; i < len; i++) x[i] /= vnorm;
        double *vk = v + (size_t)k * n;
        for (int i = 0; i < len; i++) vk[k + 1 + i] = x[i];
        has[k] = 1;
        /* A22 <- H A22 H, H = I - 2 v v^T on indices k+1..n-1 */
        for (int i = k + 1; i < n; i++) {
            double s = 0.0;
            for (int j = k + 1; j < n; j++) s += a[IDX(i, j, n)] * vk[j];
            p[i] = s;
        }
        double K = 0.0;
        for (int i = k + 1; i < n; i++) K += vk[i] * p[i];
        for (int i = k + 1; i < n; i++) p[i] -= K * vk[i]; /* q = p - (v.p) v */
        for (int j = k + 1; j < n; j++)
            for (int i = k + 1; i < n; i++)
                a[IDX(i, j, n)] -= 2.0 * (vk[i] * p[j] + p[i] * vk[j]);
        a[IDX(k + 1, k, n)] = alpha;
        a[IDX(k, k + 1, n)] = alpha;
        for (int i = k + 2; i < n; i++) {
            a[IDX(i, k, n)] = 0.0;
            a[IDX(k, i, n)] = 0.0;
        }
    }
    for (int i = 0; i < n; i++) w[i] = a[IDX(i, i, n)];
    for (int i = 0; i + 1 < n; i++) e[i] = a[IDX(i + 1, i, n)];
    /* Q = H_0 H_1 ... H_{n-3}: apply from the left in reverse order to I */
    for (size_t i = 0; i < (size_t)n * n; i++) a[i] = 0.0;
    for (int i = 0; i < n; i++) a[IDX(i, i, n)] = 1.0;
    for (int k = n - 3; k >= 0; k--) {
        if (!has[k]) continue;
        const double *vk = v + (size_t)k * n;
        for (int j = 0; j < n; j++) {
            double s = 0.0;
            for (int i = k + 1; i < n; i++) s += vk[i] * a[IDX(i, j, n)];
            if (s != 0.0)
                for (int i = k + 1; i < n; i++) a[IDX(i, j, n)] -= 2.0 * s * vk[i];
        }
    }
    int info = tql(n, w, e, a);
    free(v); free(vn); free(p); free(e); free(has);
    return info;
}

/* ===================================================================== */
/* Brownian displacement                                                  */
/* ===================================================================== */

/* calc_bdw_cholesky, brown_m.f90:516-554: dpotrf('U') then bdw = U*bdw.
 * The upper triangle of mob (including the diagonal) receives U; the strictly
 * lower triangle is not touched.  Returns the dpotrf info (0 = ok). */
int orc_chol_bdw(int n, int s, double *mob, double *bdw)
{
    for (int j = 0; j < n; j++) {
        double ajj = mob[IDX(j, j, n)];
        for (int k = 0; k < j; k++) ajj -= mob[IDX(k, j, n)] * mob[IDX(k, j, n)];
        if (!(ajj > 0.0)) {
            mob[IDX(j, j, n)] = ajj;
            return j + 1;
        }
        ajj = sqrt(ajj);
        mob[IDX(j, j, n)] = ajj;
        for (int i = j + 1; i < n; i++) {
            double t = mob[IDX(j, i, n)];
            for (int k = 0; k < j; k++) t -= mob[IDX(k, j, n)] * mob[IDX(k, i, n)];
            mob[IDX(j, i, n)] = t / ajj;
        }
    }
    /* dtrmv/dtrmm('U','N','N'): x := U x, row i uses U(i, i:n) */
    for (int c = 0; c < s; c++) {
        double *x = bdw + (size_t)c * n;
        for (int i = 0; i < n; i++) {
            double t = 0.0;
            for (int k = i; k < n; k++) t += mob[IDX(i, k, n)] * x[k];
            x[i] = t;
        }
    }
    return 0;
}

/* calc_bdw_lanc, brown_m.f90:185-288 (s = 1). */
/* Full reorthogonalisation of w (n x s) against V(:, 0:ncol), classical
 * Gram-Schmidt applied twice.  Not in the reference; optional. */
static void reorth(int n, int s, const double *V, int ncol, double *w)
{
    double *c = (double *)malloc(sizeof(double) * (size_t)(ncol > 0 ? ncol : 1));
    for (int pass = 0; pass < 2; pass++)
        for (int k = 0; k < s; k++) {
            double *wk = w + (size_t)k * n;
            for (int m = 0; m < ncol; m++) c[m] = dot(n, V + (size_t)m * n, wk);
            for (int m = 0; m < ncol; m++) {
                const double *vm = V + (size_t)m * n;
                for (int i = 0; i < n; i++) wk[i] = wk[i] - c[m] * vm[i];
            }
        }
    free(c);
}

int orc_lanczos(int n, const double *mob, double *bdw, int lanc_mxitr,
                double lanc_tol, int *lconv, int *iters)
{
    return orc_lanczos_ex(n, mob, bdw, lanc_mxitr, lanc_tol, 0, lconv, iters);
}

int orc_lanczos_ex(int n, const double *mob, double *bdw, int lanc_mxitr,
                   double lanc_tol, int do_reorth, int *lconv, int *iters)
{
    const int dimk_min = 2;
    const int dimk_max = (lanc_mxitr > dimk_min ? lanc_mxitr : dimk_min) + 1;
    double *V = (double *)calloc((size_t)n * (dimk_max + 1), sizeof(double));
    double *w = (double *)calloc((size_t)n, sizeof(double));
    double *old = (double *)calloc((size_t)n, sizeof(double));
    double *hd = (double *)calloc((size_t)dimk_max + 1, sizeof(double));
    double *hsd = (double *)calloc((size_t)dimk_max + 1, sizeof(double));
    double *ev = (double *)calloc((size_t)dimk_max, sizeof(double));
    double *ee = (double *)calloc((size_t)dimk_max, sizeof(double));
    double *P = (double *)calloc((size_t)dimk_max * dimk_max, sizeof(double));
    double *t1 = (double *)calloc((size_t)dimk_max, sizeof(double));
    double *t2 = (double *)calloc((size_t)dimk_max, sizeof(double));
    int ierr = 0;
    *lconv = 0;
    *iters = 0;

    double normz = nrm2(n, bdw);
    for (int i = 0; i < n; i++) V[i] = bdw[i] / normz;

    for (int j = 1; j <= dimk_max; j++) {
        (*iters)++;
        double *vj = V + (size_t)(j - 1) * n;
        orc_symv_lower(n, 1.0, mob, vj, 0.0, w);
        if (j > 1) {
            double *vjm = V + (size_t)(j - 2) * n;
            for (int i = 0; i < n; i++) w[i] = w[i] - hsd[j - 1] * vjm[i];
        }
        hd[j] = dot(n, w, vj);
        if (j < dimk_max) {
            for (int i = 0; i < n; i++) w[i] = w[i] - hd[j] * vj[i];
            if (do_reorth) reorth(n, 1, V, j, w);
            hsd[j] = nrm2(n, w);
            double *vjp = V + (size_t)j * n;
            for (int i = 0; i < n; i++) vjp[i] = w[i] / hsd[j];
        }
        if (j >= dimk_min) {
            for (int i = 0; i < j; i++) { ev[i] = hd[i + 1]; ee[i] = hsd[i + 1]; }
            memset(P, 0, sizeof(double) * (size_t)j * j);
            for (int i = 0; i < j; i++) P[IDX(i, i, j)] = 1.0;
            if (tql(j, ev, ee, P) != 0) { ierr = 1; break; }
            for (int i = 0; i < j; i++) ev[i] = sqrt(ev[i]);
            for (int i = 0; i < j; i++) t1[i] = ev[i] * P[IDX(0, i, j)];
            for (int r = 0; r < j; r++) {
                double t = 0.0;
                for (int c = 0; c < j; c++) t += P[IDX(r, c, j)] * t1[c];
                t2[r] = t;
            }
            for (int i = 0; i < n; i++) {
                double t = 0.0;
                for (int c = 0; c < j; c++) t += V[IDX(i, c, n)] * t2[c];
                bdw[i] = normz * t;
            }
            if (j > dimk_min) {
                double num = 0.0;
                for (int i = 0; i < n; i++) num += (bdw[i] - old[i]) * (bdw[i] - old[i]);
                double err = sqrt(num) / nrm2(n, old);
                if (err <= lanc_tol) { *lconv = 1; break; }
            }
            memcpy(old, bdw, sizeof(double) * n);
        }
    }
    free(V); free(w); free(old); free(hd); free(hsd); free(ev); free(ee);
    free(P); free(t1); free(t2);
    return ierr;
}

/* calc_bdw_blanc, brown_m.f90:292-512 (s > 1). */
int orc_block_lanczos(int n, int s, const double *mob, double *bdw,
                      int lanc_mxitr, double lanc_tol, int *lconv, int *iters)
{
    return orc_block_lanczos_ex(n, s, mob, bdw, lanc_mxitr, lanc_tol, 0, lconv, iters);
}

int orc_block_lanczos_ex(int n, int s, const double *mob, double *bdw,
                         int lanc_mxitr, double lanc_tol, int do_reorth,
                         int *lconv, int *iters)
{
    const int dimk_min = 2;
    const int dimk_max = (lanc_mxitr > dimk_min ? lanc_mxitr : dimk_min) + 1;
    const int ncol = s * dimk_max;
    double *V = (double *)calloc((size_t)n * ncol, sizeof(double));
    double *W = (double *)calloc((size_t)n * s, sizeof(double));
    double *old = (double *)calloc((size_t)n * s, sizeof(double));
    double *hbd = (double *)calloc((size_t)s * ncol, sizeof(double));
    double *hbsd = (double *)calloc((size_t)s * ncol, sizeof(double));
    double *matr = (double *)calloc((size_t)s * s, sizeof(double));
    double *tau = (double *)calloc((size_t)s, sizeof(double));
    int *jpvt = (int *)calloc((size_t)s, sizeof(int));
    double *H = (double *)calloc((size_t)ncol * ncol, sizeof(double));
    double *lam = (double *)calloc((size_t)ncol, sizeof(double));
    double *t1 = (double *)calloc((size_t)ncol * s, sizeof(double));
    double *t2 = (double *)calloc((size_t)ncol * s, sizeof(double));
    double *a2 = (double *)calloc((size_t)s * s, sizeof(double));
    int ierr = 0;
    *lconv = 0;
    *iters = 0;

    /* xi = Q R P^T; R P^T -> matr, Q -> V_1 (brown_m.f90:358-380) */
    orc_geqp3(n, s, bdw, jpvt, tau);
    for (int jc = 0; jc < s; jc++)
        for (int r = 0; r <= jc; r++) matr[IDX(r, jpvt[jc] - 1, s)] = bdw[IDX(r, jc, n)];
    orc_orgqr(n, s, bdw, tau);
    memcpy(V, bdw, sizeof(double) * (size_t)n * s);

    for (int j = 1; j <= dimk_max; j++) {
        (*iters)++;
        double *Vj = V + (size_t)(j - 1) * s * n;
        for (int c = 0; c < s; c++)
            orc_symv_lower(n, 1.0, mob, Vj + (size_t)c * n, 0.0, W + (size_t)c * n);
        if (j > 1) {
            /* W -= V_{j-1} * Hsd_{j-1}^T (brown_m.f90:393) */
            double *Vm = V + (size_t)(j - 2) * s * n;
            double *B = hbsd + (size_t)(j - 2) * s * s;
            for (int c = 0; c < s; c++)
                for (int i = 0; i < n; i++) {
                    double t = 0.0;
                    for (int k = 0; k < s; k++) t += Vm[IDX(i, k, n)] * B[IDX(c, k, s)];
                    W[IDX(i, c, n)] = W[IDX(i, c, n)] - t;
                }
        }
        /* H_jj = V_j^T W (brown_m.f90:399) */
        double *A = hbd + (size_t)(j - 1) * s * s;
        for (int c = 0; c < s; c++)
            for (int r = 0; r < s; r++)
                A[IDX(r, c, s)] = dot(n, Vj + (size_t)r * n, W + (size_t)c * n);
        if (j < dimk_max) {
            /* W -= V_j * sym_L(H_jj) (brown_m.f90:406) */
            for (int c = 0; c < s; c++)
                for (int r = 0; r < s; r++)
                    a2[IDX(r, c, s)] = (r >= c) ? A[IDX(r, c, s)] : A[IDX(c, r, s)];
            for (int c = 0; c < s; c++)
                for (int i = 0; i < n; i++) {
                    double t = 0.0;
                    for (int k = 0; k < s; k++) t += Vj[IDX(i, k, n)] * a2[IDX(k, c, s)];
                    W[IDX(i, c, n)] = W[IDX(i, c, n)] - t;
                }
            if (do_reorth) reorth(n, s, V, j * s, W);
            orc_geqp3(n, s, W, jpvt, tau);
            double *B = hbsd + (size_t)(j - 1) * s * s;
            for (int jc = 0; jc < s; jc++)
                for (int r = 0; r <= jc; r++) B[IDX(r, jpvt[jc] - 1, s)] = W[IDX(r, jc, n)];
            orc_orgqr(n, s, W, tau);
            memcpy(V + (size_t)j * s * n, W, sizeof(double) * (size_t)n * s);
        }
        if (j >= dimk_min) {
            const int js = j * s;
            /* dense copy of the band matrix packed at brown_m.f90:442-463 */
            memset(H, 0, sizeof(double) * (size_t)js * js);
            for (int jb = 0; jb < j; jb++) {
                const double *Ab = hbd + (size_t)jb * s * s;
                for (int c = 0; c < s; c++)
                    for (int r = c; r < s; r++) {
                        H[IDX(jb * s + r, jb * s + c, js)] = Ab[IDX(r, c, s)];
                        H[IDX(jb * s + c, jb * s + r, js)] = Ab[IDX(r, c, s)];
                    }
            }
            for (int jb = 0; jb + 1 < j; jb++) {
                const double *Bb = hbsd + (size_t)jb * s * s;
                for (int c = 0; c < s; c++)
                    for (int r = 0; r < s; r++) {
                        H[IDX((jb + 1) * s + r, jb * s + c, js)] = Bb[IDX(r, c, s)];
                        H[IDX(jb * s + c, (jb + 1) * s + r, js)] = Bb[IDX(r, c, s)];
                    }
            }
            if (orc_syev(js, H, lam) != 0) { ierr = 1; break; }
            for (int i = 0; i < js; i++) lam[i] = sqrt(lam[i]);
            /* t1 = P(1:s,:)^T matr; scale rows by sqrt(lambda); t2 = P t1 */
            for (int c = 0; c < s; c++)
                for (int i = 0; i < js; i++) {
                    double t = 0.0;
                    for (int k = 0; k < s; k++) t += H[IDX(k, i, js)] * matr[IDX(k, c, s)];
                    t1[IDX(i, c, js)] = lam[i] * t;
                }
            for (int c = 0; c < s; c++)
                for (int i = 0; i < js; i++) {
                    double t = 0.0;
                    for (int k = 0; k < js; k++) t += H[IDX(i, k, js)] * t1[IDX(k, c, js)];
                    t2[IDX(i, c, js)] = t;
                }
            for (int c = 0; c < s; c++)
                for (int i = 0; i < n; i++) {
                    double t = 0.0;
                    for (int k = 0; k < js; k++) t += V[IDX(i, k, n)] * t2[IDX(k, c, js)];
                    bdw[IDX(i, c, n)] = t;
                }
            if (j > dimk_min) {
                double num = 0.0, den = 0.0;
                for (int c = 0; c < s; c++) {
                    double a = 0.0, b = 0.0;
                    for (int i = 0; i < n; i++) {
                        a += fabs(bdw[IDX(i, c, n)] - old[IDX(i, c, n)]);
                        b += fabs(old[IDX(i, c, n)]);
                    }
                    if (a > num) num = a;
                    if (b > den) den = b;
                }
                if (num / den <= lanc_tol) { *lconv = 1; *iters = j; break; }
            }
            memcpy(old, bdw, sizeof(double) * (size_t)n * s);
        }
    }
    free(V); free(W); free(old); free(hbd); free(hbsd); free(matr); free(tau);
    free(jpvt); free(H); free(lam); free(t1); free(t2); free(a2);
    return ierr;
}

/* brn_calc_bdw, brown_m.f90:147-181.  method: 0 CHOL (U.xi, as the
 * reference), 1 LANC, 2 CHOL with L.xi = U^T.xi (covariance D). */
int orc_brn_calc_bdw(int method, int n, int s, double *mob, double *bdw,
                     int lanc_mxitr, double lanc_tol, int *lconv, int *iters)
{
    return orc_brn_calc_bdw_ex(method, n, s, mob, bdw, lanc_mxitr, lanc_tol, 0, lconv, iters);
}

int orc_brn_calc_bdw_ex(int method, int n, int s, double *mob, double *bdw,
                        int lanc_mxitr, double lanc_tol, int do_reorth,
                        int *lconv, int *iters)
{
    *lconv = 1;
    *iters = 0;
    if (method == 0) return orc_chol_bdw(n, s, mob, bdw) ? 1 : 0;
    if (method == 2) {
        int info = orc_chol_bdw(n, 0, mob, bdw);
        if (info) return 1;
        for (int c = 0; c < s; c++) { /* x := U^T x, row i uses U(1:i, i) */
            double *x = bdw + (size_t)c * n;
            for (int i = n - 1; i >= 0; i--) {
                double t = 0.0;
                for (int k = 0; k <= i; k++) t += mob[IDX(k, i, n)] * x[k];
                x[i] = t;
            }
        }
        return 0;
    }
    if (s == 1) return orc_lanczos_ex(n, mob, bdw, lanc_mxitr, lanc_tol, do_reorth, lconv, iters);
    return orc_block_lanczos_ex(n, s, mob, bdw, lanc_mxitr, lanc_tol, do_reorth, lconv, iters);
}

/* ===================================================================== */
/* Simulation box                                                         */
/* ===================================================================== */

/* Fortran nint: round half away from zero. */
static double fnint(double x) { return round(x); }

/* smbx_get_image, simbox_m.f90:106-124 (aligned box). */
void orc_get_image(const double *box, double *r)
{
    for (int d = 0; d < 3; d++) r[d] = r[d] - box[d] * fnint(r[d] / box[d]);
}

/* smbx_wrap_all, simbox_m.f90:128-153 (aligned box). */
void orc_wrap_all(const double *box, int nb, double *crd)
{
    for (int i = 0; i < nb; i++)
        for (int d = 0; d < 3; d++)
            crd[3 * i + d] = crd[3 * i + d] - box[d] * floor(crd[3 * i + d] / box[d]);
}

/* smbx_to_origin, simbox_m.f90:182-202. */
void orc_to_origin(int nb, double *crd, double *com)
{
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < nb; i++)
        for (int d = 0; d < 3; d++) c[d] += crd[3 * i + d];
    for (int d = 0; d < 3; d++) c[d] = c[d] / nb;
    for (int i = 0; i < nb; i++)
        for (int d = 0; d < 3; d++) crd[3 * i + d] = crd[3 * i + d] - c[d];
    if (com)
        for (int d = 0; d < 3; d++) com[d] = c[d];
}

/* compute_ic0 rgsq, stats_m.f90:441-446. */
double orc_rgsq(int nb, const double *crd)
{
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < nb; i++)
        for (int d = 0; d < 3; d++) c[d] += crd[3 * i + d];
    for (int d = 0; d < 3; d++) c[d] = c[d] / nb;
    double rg = 0.0;
    for (int i = 0; i < nb; i++) {
        double s = 0.0;
        for (int d = 0; d < 3; d++) {
            double t = crd[3 * i + d] - c[d];
            s += t * t;
        }
        rg += s;
    }
    return rg / nb;
}

/* ===================================================================== */
/* Potentials (ia_bond_m.f90, ia_vdw_m.f90)                               */
/* ===================================================================== */

#define MATH_SXRT2 1.122462048309373
#define MATH_PI 3.1415926535897931 /* constants_m.f90:31 */

/* ia_bond_setup / ia_vdw_setup: derived parameter slots (ia_bond_m.f90:158-
 * 190, 227-265; ia_vdw_m.f90:107-151). */
void orc_setup_params(orc_system *sys)
{
    for (int t = 0; t < sys->nbond_types; t++) {
        double *p = sys->bond_params + (size_t)t * ORC_MXPARAM;
        if (sys->bond_styles[t] == 2) p[3] = p[1] * p[1];
        if (sys->bond_styles[t] == 3) {
            p[4] = p[1] * p[1];
            p[5] = MATH_SXRT2 * p[3];
        }
    }
    /* ia_vdw_setup, ia_vdw_m.f90:21-47: the setters' derived slots */
    for (int t = 0; t < sys->nvdw_types; t++) {
        double *p = sys->vdw_params + (size_t)t * ORC_MXPARAM;
        int sty = sys->vdw_styles[t];
        if (sty == 1 || sty == 4 || sty == 5) { /* vdw_lj_set 85-151, 270-306, 356-389 */
            double sir = p[1] / p[2];
            double sir2 = sir * sir, sir6 = sir2 * sir2 * sir2, sir12 = sir6 * sir6;
            double pot_rcut = 4 * p[0] * (sir12 - sir6);
            if (sty == 1) {
                p[3] = pot_rcut;
                p[4] = -24 * p[0] * (2 * sir12 - sir6) / p[2];
            } else if (sty == 4) {
                p[6] = pot_rcut;
                p[7] = (p[3] > 0.0) ? p[4] * exp(-p[5] * p[3]) / p[3] : 0.0;
            } else {
                p[5] = pot_rcut;
                p[6] = (p[3] > 0.0) ? p[4] / p[3] : 0.0;
            }
        } else if (sty == 2) { /* vdw_gaussian_set 185-206 */
            p[3] = p[0] * exp(-p[1] * (p[2] * p[2]));
            p[4] = -2.0 * p[0] * p[1] * p[2] * exp(-p[1] * (p[2] * p[2]));
        }
    }
}

/* ia_get_bond_force, ia_bond_m.f90:85-362.  Returns ierr. */
static int bond_force(double r, int sty, const double *p, double *enrg, double *frc)
{
    *enrg = 0.0;
    *frc = 0.0;
    switch (sty) {
    case 1: /* harmonic, 143-154 */
        *enrg = 0.5 * p[0] * (r - p[1]) * (r - p[1]);
        *frc = p[0] * (r - p[1]);
        return 0;
    case 2: { /* FENE with r0 offset, 194-223 */
        double k = p[0], rmax = p[1], r0 = p[2], rmaxsq = p[3];
        double extn = r - r0, extnsq = extn * extn;
        if (r >= rmax + r0) return 1;
        *enrg = -0.5 * k * rmaxsq * log(1.0 - extnsq / rmaxsq);
        *frc = k * extn / (1.0 - extnsq / rmaxsq);
        return 0;
    }
    case 3: { /* Kremer-Grest, 269-306 */
        double k = p[0], rmax = p[1], eps = p[2], sigma = p[3];
        double rmaxsq = p[4], rcut = p[5], rsq = r * r;
        if (r >= rmax) return 1;
        if (r >= rcut) {
            *enrg = -0.5 * k * rmaxsq * log(1.0 - rsq / rmaxsq);
            *frc = k * r / (1.0 - rsq / rmaxsq);
        } else {
            double sir = sigma / r, sir2 = sir * sir, sir6 = sir2 * sir2 * sir2;
            double sir12 = sir6 * sir6;
            *enrg = -0.5 * k * rmaxsq * log(1.0 - rsq / rmaxsq) + 4 * eps * (sir12 - sir6) + eps;
            *frc = k * r / (1.0 - rsq / rmaxsq) - 24 * eps * (2 * sir12 - sir6) / r;
        }
        return 0;
    }
    case 4: { /* Marko-Siggia, 334-362 */
        double lp = p[0], rmax = p[1];
        if (r >= rmax) return 1;
        double rt = r / rmax, rrt = 1.0 / (1.0 - rt);
        *enrg = (-0.5 * rt * rt + 0.25 * rrt * rrt + 0.25 * rt) * (rmax / lp);
        *frc = (-rt - 0.25 * rrt * rrt + 0.25) / lp;
        return 0;
    }
    default:
        return 0;
    }
}

/* ia_get_vdw_force, ia_vdw_m.f90:51-81, every style; qiqj = qi*qj. */
static void vdw_force(double r, double qiqj, int sty, const double *p, double *enrg, double *frc)
{
    *enrg = 0.0;
    *frc = 0.0;
    switch (sty) {
    case 1: { /* vdw_lj, 155-180 */
        double eps = p[0], sigma = p[1], rcut = p[2], vrc = p[3], dvrc = p[4];
        if (r < rcut) {
            double sir = sigma / r, sir2 = sir * sir, sir6 = sir2 * sir2 * sir2;
            double sir12 = sir6 * sir6;
            *enrg = 4 * eps * (sir12 - sir6) - vrc - (r - rcut) * dvrc;
            *frc = -24 * eps * (2 * sir12 - sir6) / r - dvrc;
        }
        return;
    }
    case 2: { /* vdw_gaussian, 210-229 */
        double A = p[0], B = p[1], rcut = p[2];
        if (r < rcut) {
            double exrs = exp(-B * r * r);
            *enrg = A * exrs - p[3] - (r - rcut) * p[4];
            *frc = -2 * A * B * r * exrs - p[4];
        }
        return;
    }
    case 3: { /* vdw_cosine, 249-266 */
        double A = p[0], rcut = p[1];
        double pf = MATH_PI / rcut, pr = pf * r;
        if (r < rcut) {
            *enrg = A * (1.0 + cos(pr));
            *frc = -A * pf * sin(pr);
        }
        return;
    }
    case 4:   /* vdw_lj_coul_debye, 310-352 */
    case 5: { /* vdw_lj_coul, 391-432 */
        double eps = p[0], sigma = p[1], rcut = p[2], rcut_coul = p[3], C = p[4];
        double kappa = (sty == 4) ? p[5] : 0.0;
        double pot_rcut = (sty == 4) ? p[6] : p[5];
        double pot_rcut_coul = (sty == 4) ? p[7] : p[6];
        double ekr, fc;
        if (sty == 4) {
            ekr = C * qiqj * exp(-kappa * r) / r;
            fc = -ekr * (1 + kappa * r) / r;
        } else {
            ekr = C * qiqj / r;
            fc = -ekr / r;
        }
        if (rcut_coul > 0.0) {
            if (r < rcut_coul) {
                *enrg = ekr - qiqj * pot_rcut_coul;
                *frc = fc;
            }
        } else {
            *enrg = ekr;
            *frc = fc;
        }
        if (r < rcut) {
            double sir = sigma / r, sir2 = sir * sir, sir6 = sir2 * sir2 * sir2;
            double sir12 = sir6 * sir6;
            *enrg = *enrg + 4 * eps * (sir12 - sir6) - pot_rcut;
            *frc = *frc - 24 * eps * (2 * sir12 - sir6) / r;
        }
        return;
    }
    case 6: { /* vdw_dpd, 454-470 */
        double A = p[0], rcut = p[1], ir = 1.0 - (r / rcut);
        if (r < rcut) {
            *enrg = 0.5 * A * rcut * ir * ir;
            *frc = -A * ir;
        }
        return;
    }
    default:
        return;
    }
}

/* ia_get_angle_force, ia_angle_m.f90:46-79 (style 1). */
static void angle_force(const double *q1, const double *q2, int sty, const double *p, double *enrg,
                        double *fim1, double *fi, double *fip1)
{
    *enrg = 0.0;
    for (int d = 0; d < 3; d++) fim1[d] = fi[d] = fip1[d] = 0.0;
    double q1mag = sqrt(q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2]);
    double q2mag = sqrt(q2[0] * q2[0] + q2[1] * q2[1] + q2[2] * q2[2]);
    double q1hat[3], q2hat[3];
    for (int d = 0; d < 3; d++) {
        q1hat[d] = q1[d] / q1mag;
        q2hat[d] = q2[d] / q2mag;
    }
    double ctheta = q1hat[0] * q2hat[0] + q1hat[1] * q2hat[1] + q1hat[2] * q2hat[2];
    if (ctheta > 1.0) ctheta = 1.0;
    if (ctheta < -1.0) ctheta = -1.0;
    if (sty != 1) return;
    double kang = p[0];
    *enrg = kang * (1.0 - ctheta);
    for (int d = 0; d < 3; d++) {
        fim1[d] = kang * (-q2hat[d] + ctheta * q1hat[d]) / q1mag;
        fip1[d] = kang * (q1hat[d] - ctheta * q2hat[d]) / q2mag;
        fi[d] = -(fim1[d] + fip1[d]);
    }
}

/* ia_get_tether_force, ia_tether_m.f90:53-80, teth_rigid 110-128, teth_harm
 * 146-158.  Returns ierr. */
static int tether_force(double r, int sty, const double *p, double *enrg, double *frc)
{
    *enrg = 0.0;
    *frc = 0.0;
    if (sty == 1) return (fabs(r - p[0]) < p[1]) ? 0 : 1;
    if (sty == 2) {
        *enrg = 0.5 * p[0] * (r - p[1]) * (r - p[1]);
        *frc = p[0] * (r - p[1]);
    }
    return 0;
}

static void add_stress(double *st, const double *rij, const double *fi)
{
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) st[a + 3 * b] = st[a + 3 * b] - rij[a] * fi[b];
}

/* ia_calc_forces, interaction_m.f90:84-147, with ia_add_bond_forces
 * (219-272) and ia_add_vdw_forces (151-215) over the given pair table. */
int orc_forces(const orc_system *sys, const double *crd, const int *pt_ptr,
               const int *pt_col, double *frc, orc_energies *en)
{
    const int nb = sys->nbeads;
    memset(frc, 0, sizeof(double) * 3 * (size_t)nb);
    memset(en, 0, sizeof(*en));
    if (sys->nbonds > 0) {
        en->bndlen = 0.0;
        en->bndlen_min = DBL_MAX;
        en->bndlen_max = 0.0;
        for (int b = 0; b < sys->nbonds; b++) {
            int typ = sys->bonds[3 * b] - 1;
            int i = sys->bonds[3 * b + 1] - 1, j = sys->bonds[3 * b + 2] - 1;
            double rij[3];
            for (int d = 0; d < 3; d++) rij[d] = crd[3 * j + d] - crd[3 * i + d];
            if (sys->imcon) orc_get_image(sys->box, rij);
            double r = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
            en->bndlen += r;
            if (r < en->bndlen_min) en->bndlen_min = r;
            if (r > en->bndlen_max) en->bndlen_max = r;
            double e, f;
            if (bond_force(r, sys->bond_styles[typ], sys->bond_params + (size_t)typ * ORC_MXPARAM, &e, &f))
                return 1;
            en->energy_bond += e;
            double fi[3];
            for (int d = 0; d < 3; d++) fi[d] = f * rij[d] / r;
            for (int d = 0; d < 3; d++) {
                frc[3 * i + d] += fi[d];
                frc[3 * j + d] -= fi[d];
            }
            add_stress(en->stress, rij, fi);
        }
        en->bndlen /= sys->nbonds;
    }
    /* ia_add_angle_forces, interaction_m.f90:276-325 */
    for (int a = 0; a < sys->nangles; a++) {
        int typ = sys->angles[4 * a] - 1;
        int im1 = sys->angles[4 * a + 1] - 1, i = sys->angles[4 * a + 2] - 1, ip1 = sys->angles[4 * a + 3] - 1;
        double q1[3], q2[3], fim1[3], fi[3], fip1[3], e;
        for (int d = 0; d < 3; d++) {
            q1[d] = crd[3 * i + d] - crd[3 * im1 + d];
            q2[d] = crd[3 * ip1 + d] - crd[3 * i + d];
        }
        if (sys->imcon) {
            orc_get_image(sys->box, q1);
            orc_get_image(sys->box, q2);
        }
        angle_force(q1, q2, sys->angle_styles[typ], sys->angle_params + (size_t)typ * ORC_MXPARAM, &e, fim1, fi,
                    fip1);
        en->energy_angle += e;
        for (int d = 0; d < 3; d++) {
            frc[3 * im1 + d] += fim1[d];
            frc[3 * i + d] += fi[d];
            frc[3 * ip1 + d] += fip1[d];
        }
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++)
                en->stress[c + 3 * b] = en->stress[c + 3 * b] - q1[c] * fim1[b] + q2[c] * fip1[b];
    }
    /* ia_add_tether_forces, interaction_m.f90:375-418 (no imaging) */
    for (int k = 0; k < sys->ntethers; k++) {
        int typ = sys->tethers[2 * k] - 1, ia = sys->tethers[2 * k + 1] - 1;
        double q[3];
        for (int d = 0; d < 3; d++) q[d] = crd[3 * ia + d] - sys->tether_points[3 * k + d];
        double qmag = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
        double e, f;
        if (tether_force(qmag, sys->tether_styles[typ], sys->tether_params + (size_t)typ * ORC_MXPARAM, &e, &f))
            return 1;
        double fj[3];
        for (int d = 0; d < 3; d++) fj[d] = -f * q[d] / qmag;
        en->energy_tether += e;
        for (int d = 0; d < 3; d++) frc[3 * ia + d] += fj[d];
        for (int b = 0; b < 3; b++)
            for (int c = 0; c < 3; c++) en->stress[c + 3 * b] += q[c] * fj[b];
    }
    if (sys->lvdw && sys->nvdw_types > 0 && pt_ptr) {
        const int nt = sys->natom_types;
        for (int i = 0; i < nb; i++) {
            int ati = sys->atom_types[i];
            for (int k = pt_ptr[i]; k < pt_ptr[i + 1]; k++) {
                int j = pt_col[k] - 1;
                int atj = sys->atom_types[j];
                int typ = (ati < atj) ? atj + (2 * nt - ati) * (ati - 1) / 2
                                      : ati + (2 * nt - atj) * (atj - 1) / 2;
                typ -= 1;
                double rij[3];
                for (int d = 0; d < 3; d++) rij[d] = crd[3 * j + d] - crd[3 * i + d];
                if (sys->imcon) orc_get_image(sys->box, rij);
                double r = sqrt(rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2]);
                double e, f;
                double qq = sys->charge ? sys->charge[i] * sys->charge[j] : 0.0;
                vdw_force(r, qq, sys->vdw_styles[typ], sys->vdw_params + (size_t)typ * ORC_MXPARAM, &e, &f);
                en->energy_vdw += e;
                double fi[3];
                for (int d = 0; d < 3; d++) fi[d] = f * rij[d] / r;
                for (int d = 0; d < 3; d++) {
                    frc[3 * i + d] += fi[d];
                    frc[3 * j + d] -= fi[d];
                }
                add_stress(en->stress, rij, fi);
            }
        }
    }
    /* ia_add_external_forces, ia_external_m.f90:61-107 */
    for (int e = 0; e < sys->nexternals; e++) {
        const double *p = sys->external_params + (size_t)e * ORC_MXPARAM;
        if (sys->external_styles[e] == 1) {
            int ia = (int)p[0] - 1;
            en->energy_external += -p[1] * (crd[3 * ia] - crd[0]);
            frc[3 * ia] += p[1];
        } else if (sys->external_styles[e] == 2) {
            int m = (int)p[0] - 1;
            for (int i = 0; i < nb; i++) {
                double x = crd[3 * i + m];
                if (p[2] > 0.0 ? x < p[1] : x > p[1]) return 1;
            }
        }
    }
    en->energy_tot = en->energy_bond + en->energy_angle + en->energy_dihedral + en->energy_vdw
                     + en->energy_tether + en->energy_external;
    if (sys->imcon) {
        double vol = sys->box[0] * sys->box[1] * sys->box[2];
        for (int k = 0; k < 9; k++) en->stress[k] /= vol;
    }
    return 0;
}

/* ===================================================================== */
/* Exclusions and pair tables                                             */
/* ===================================================================== */

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

static int sort_unique(int *v, int n)
{
    if (n == 0) return 0;
    qsort(v, (size_t)n, sizeof(int), cmp_int);
    int m = 1;
    for (int i = 1; i < n; i++)
        if (v[i] != v[m - 1]) v[m++] = v[i];
    return m;
}

/* atbo_build + atat_build + exat_build, connectivity_m.f90:43-90, 202-317.
 * Returns the number of entries written, or -1 if cap is exceeded. */
int orc_exclusions(int nb, int nbonds, const int *bonds, int excl,
                   int *row_ptr, int *col, int cap)
{
    /* 1-ring: atoms bonded to each atom (atat table), sorted & unique */
    int *deg = (int *)calloc((size_t)nb + 1, sizeof(int));
    for (int b = 0; b < nbonds; b++) {
        deg[bonds[3 * b + 1] - 1]++;
        deg[bonds[3 * b + 2] - 1]++;
    }
    int *aptr = (int *)calloc((size_t)nb + 1, sizeof(int));
    for (int i = 0; i < nb; i++) aptr[i + 1] = aptr[i] + deg[i];
    int *acol = (int *)malloc(sizeof(int) * (size_t)(aptr[nb] > 0 ? aptr[nb] : 1));
    int *fill = (int *)calloc((size_t)nb, sizeof(int));
    for (int b = 0; b < nbonds; b++) {
        int i = bonds[3 * b + 1], j = bonds[3 * b + 2];
        if (j != i) acol[aptr[i - 1] + fill[i - 1]++] = j;
        if (i != j) acol[aptr[j - 1] + fill[j - 1]++] = i;
    }
    int *alen = (int *)calloc((size_t)nb, sizeof(int));
    for (int i = 0; i < nb; i++) alen[i] = sort_unique(acol + aptr[i], fill[i]);

    int total = 0;
    int bufcap = 64;
    int *buf = (int *)malloc(sizeof(int) * bufcap);
    row_ptr[0] = 0;
    for (int i = 0; i < nb; i++) {
        int m = 0;
        if (excl > 0) {
#define PUSH(v)                                                      \
    do {                                                             \
        if (m == bufcap) {                                           \
            bufcap *= 2;                                             \
            buf = (int *)realloc(buf, sizeof(int) * bufcap);         \
        }                                                            \
        buf[m++] = (v);                                              \
    } while (0)
            PUSH(i + 1);
            for (int a = 0; a < alen[i]; a++) {
                int j = acol[aptr[i] + a];
                PUSH(j);
                if (excl > 1) {
                    for (int a2 = 0; a2 < alen[j - 1]; a2++) {
                        int j2 = acol[aptr[j - 1] + a2];
                        PUSH(j2);
                        if (excl > 2)
                            for (int a3 = 0; a3 < alen[j2 - 1]; a3++) PUSH(acol[aptr[j2 - 1] + a3]);
                    }
                }
            }
#undef PUSH
        }
        m = sort_unique(buf, m);
        if (total + m > cap) { total = -1; break; }
        memcpy(col + total, buf, sizeof(int) * m);
        total += m;
        row_ptr[i + 1] = total;
    }
    free(buf); free(deg); free(aptr); free(acol); free(fill); free(alen);
    return total;
}

/* itbl_is_in, table_m.f90:129-147 (row i, 0-based bead, value 1-based). */
static int ex_has(const int *ex_ptr, const int *ex_col, int i, int v)
{
    for (int k = ex_ptr[i]; k < ex_ptr[i + 1]; k++)
        if (ex_col[k] == v) return 1;
    return 0;
}

/* build_pt_n2, pairtab_m.fpp:152-190 (rc_sq = rcutoff^2, DIR). */
int orc_pairs_n2(const orc_system *sys, const int *ex_ptr, const int *ex_col,
                 const double *crd, double rc_sq, int *row_ptr, int *col, int cap)
{
    const int nb = sys->nbeads;
    int np = 0;
    row_ptr[0] = 0;
    for (int i = 0; i < nb; i++) {
        if (i < nb - 1) {
            for (int j = i + 1; j < nb; j++) {
                if (ex_has(ex_ptr, ex_col, i, j + 1)) continue;
                double rij[3];
                for (int d = 0; d < 3; d++) rij[d] = crd[3 * j + d] - crd[3 * i + d];
                if (sys->imcon) orc_get_image(sys->box, rij);
                double rsq = rij[0] * rij[0] + rij[1] * rij[1] + rij[2] * rij[2];
                if (rsq < rc_sq) {
                    if (np >= cap) return -1;
                    col[np++] = j + 1;
                }
            }
        }
        row_ptr[i + 1] = np;
    }
    return np;
}

/* build_pt_verlet, pairtab_m.fpp:194-252. */
int orc_pairs_verlet(const orc_system *sys, const int *ex_ptr, const int *ex_col,
                     const double *crd, orc_verlet *vs, int *row_ptr, int *col,
                     int cap, int *rebuilt)
{
    const int nb = sys->nbeads;
    *rebuilt = 0;
    if (!vs->first_call) {
        double drmax = 0.0;
        for (int i = 0; i < nb; i++) {
            double dr[3];
            for (int d = 0; d < 3; d++) dr[d] = crd[3 * i + d] - vs->crd_save[3 * i + d];
            if (sys->imcon) orc_get_image(sys->box, dr);
            double s = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            if (s > drmax) drmax = s;
        }
        if (4 * drmax < vs->tskin_sq) return -2; /* keep the current list */
    }
    vs->first_call = 0;
    int np = orc_pairs_n2(sys, ex_ptr, ex_col, crd, vs->rskin_sq, row_ptr, col, cap);
    memcpy(vs->crd_save, crd, sizeof(double) * 3 * (size_t)nb);
    *rebuilt = 1;
    return np;
}

/* build_pt_cell_list, pairtab_m.fpp:305-385, with cl_set_cell_size,
 * cl_build_cell_nbrs and cl_build (cell_list_m.f90:91-228).  Cells are of
 * size L/floor(L/rcutoff); candidate pairs are not distance filtered.
 * Returns -3 for a negative coordinate (the reference stops there). */
int orc_pairs_cell_list(const orc_system *sys, const int *ex_ptr,
                        const int *ex_col, const double *crd, int *row_ptr,
                        int *col, int cap)
{
    static const int dstn[13][3] = {{1, 0, 0},  {1, 1, 0},  {-1, 1, 0}, {0, 1, 0},  {0, 0, 1},
                                    {-1, 0, 1}, {1, 0, 1},  {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                                    {-1, 1, 1}, {0, 1, 1},  {1, 1, 1}};
    const int nb = sys->nbeads;
    int nc[3];
    double cs[3];
    for (int d = 0; d < 3; d++) {
        nc[d] = (int)floor(sys->box[d] / sys->rcutoff);
        cs[d] = sys->box[d] / nc[d];
    }
    const int nct = nc[0] * nc[1] * nc[2];
    int *host = (int *)malloc(sizeof(int) * nb);
    int *pop = (int *)calloc((size_t)nct, sizeof(int));
    int *pos = (int *)calloc((size_t)nct + 1, sizeof(int));
    int *cells = (int *)malloc(sizeof(int) * nb);
    for (int i = 0; i < nb; i++) {
        int ic3[3];
        for (int d = 0; d < 3; d++) {
            double q = crd[3 * i + d] / cs[d];
            if (q <= -1.0 || q != q) { free(host); free(pop); free(pos); free(cells); return -3; }
            ic3[d] = (int)q; /* Fortran int(): truncation toward zero */
            if (ic3[d] < 0) { free(host); free(pop); free(pos); free(cells); return -3; }
            if (ic3[d] > nc[d] - 1) ic3[d] = nc[d] - 1;
        }
        int ic = ic3[2] * nc[0] * nc[1] + ic3[1] * nc[0] + ic3[0];
        host[i] = ic;
        pop[ic]++;
    }
    for (int ic = 0; ic < nct; ic++) pos[ic + 1] = pos[ic] + pop[ic];
    int *fillp = (int *)malloc(sizeof(int) * (nct + 1));
    memcpy(fillp, pos, sizeof(int) * (nct + 1));
    for (int i = 0; i < nb; i++) cells[fillp[host[i]]++] = i + 1;
    free(fillp);

    /* ptaov: per-atom growable lists, then flattened in atom order */
    int *cnt = (int *)calloc((size_t)nb, sizeof(int));
    int *lcap = (int *)calloc((size_t)nb, sizeof(int));
    int **lst = (int **)calloc((size_t)nb, sizeof(int *));
#define APPEND(ia, v)                                                          \
    do {                                                                       \
        if (cnt[ia] == lcap[ia]) {                                             \
            lcap[ia] = lcap[ia] ? 2 * lcap[ia] : 16;                           \
            lst[ia] = (int *)realloc(lst[ia], sizeof(int) * lcap[ia]);         \
        }                                                                      \
        lst[ia][cnt[ia]++] = (v);                                              \
    } while (0)
    for (int ic = 0; ic < nct; ic++) {
        int icx = ic % nc[0], icy = (ic / nc[0]) % nc[1], icz = ic / (nc[0] * nc[1]);
        for (int a = pos[ic]; a < pos[ic + 1] - 1; a++) {
            int ia = cells[a];
            for (int b = a + 1; b < pos[ic + 1]; b++) {
                int ja = cells[b];
                if (ex_has(ex_ptr, ex_col, ia - 1, ja)) continue;
                APPEND(ia - 1, ja);
            }
        }
        for (int a = pos[ic]; a < pos[ic + 1]; a++) {
            int ia = cells[a];
            for (int k = 0; k < 13; k++) {
                int jx = ((icx + dstn[k][0]) % nc[0] + nc[0]) % nc[0];
                int jy = ((icy + dstn[k][1]) % nc[1] + nc[1]) % nc[1];
                int jz = ((icz + dstn[k][2]) % nc[2] + nc[2]) % nc[2];
                int jc = jz * nc[0] * nc[1] + jy * nc[0] + jx;
                for (int b = pos[jc]; b < pos[jc + 1]; b++) {
                    int ja = cells[b];
                    if (ex_has(ex_ptr, ex_col, ia - 1, ja)) continue;
                    APPEND(ia - 1, ja);
                }
            }
        }
    }
#undef APPEND
    int np = 0;
    row_ptr[0] = 0;
    for (int i = 0; i < nb; i++) {
        if (np + cnt[i] > cap) { np = -1; break; }
        if (cnt[i]) memcpy(col + np, lst[i], sizeof(int) * cnt[i]);
        np += cnt[i];
        row_ptr[i + 1] = np;
    }
    for (int i = 0; i < nb; i++) free(lst[i]);
    free(lst); free(cnt); free(lcap);
    free(host); free(pop); free(pos); free(cells);
    return np;
}

/* ===================================================================== */
/* BD time loop                                                           */
/* ===================================================================== */

/* bds_run (bd_solver_m.f90:128-212) with integrate_em (216-238),
 * calc_diffusion (355-403) and calc_drift (407-454).  Statistics, dumps and
 * logging are not part of the step.  xi replaces brn_calc_dw
 * (brown_m.f90:558-573). */
/* calc_drift (bd_solver_m.f90:407-454) at crd: pair table, forces, flow
 * field, D.F (or F) and the factor dt, into drift.  Returns 0, -1 (bond
 * overstretched) or -12 (pair table capacity). */
static int drift_eval(const orc_system *sys, const orc_bdopts *opt, const double *mob, const int *ex_ptr,
                      const int *ex_col, orc_verlet *vsp, int *pt_ptr, int *pt_col, int ptcap, const double *crd,
                      double *frc, double *drift, orc_energies *en)
{
    const int nb = sys->nbeads, n = 3 * nb;
    memset(drift, 0, sizeof(double) * n);
    if (sys->lvdw && sys->nvdw_types > 0) {
        int np;
        if (sys->mth_ptgen == 1) {
            int rebuilt;
            np = orc_pairs_verlet(sys, ex_ptr, ex_col, crd, vsp, pt_ptr, pt_col, ptcap, &rebuilt);
            if (np == -2) np = 0;
        } else if (sys->mth_ptgen == 2) {
            np = orc_pairs_cell_list(sys, ex_ptr, ex_col, crd, pt_ptr, pt_col, ptcap);
        } else {
            np = orc_pairs_n2(sys, ex_ptr, ex_col, crd, sys->rcutoff * sys->rcutoff, pt_ptr, pt_col, ptcap);
        }
        if (np < 0) return -12;
    }
    if (orc_forces(sys, crd, (sys->lvdw && sys->nvdw_types > 0) ? pt_ptr : NULL, pt_col, frc, en)) return -1;
    switch (opt->flow_style) {
    case 1:
        for (int i = 0; i < nb; i++) drift[3 * i] = opt->flow_rate * crd[3 * i + 1];
        break;
    case 2:
        for (int i = 0; i < nb; i++) {
            drift[3 * i] = opt->flow_rate * crd[3 * i];
            drift[3 * i + 1] = -opt->flow_rate * crd[3 * i + 1];
        }
        break;
    case 3:
        for (int i = 0; i < nb; i++) {
            drift[3 * i] = opt->flow_rate * crd[3 * i];
            drift[3 * i + 1] = -0.5 * opt->flow_rate * crd[3 * i + 1];
            drift[3 * i + 2] = -0.5 * opt->flow_rate * crd[3 * i + 2];
        }
        break;
    default:
        break;
    }
    if (opt->lhdia)
        orc_symv_lower(n, 1.0, mob, frc, 1.0, drift);
    else
        for (int i = 0; i < n; i++) drift[i] = drift[i] + frc[i];
    for (int i = 0; i < n; i++) drift[i] = opt->tim_stp * drift[i];
    return 0;
}

/* se_fval (bd_solver_m.f90:312-332): f = x - drift(x) - crd0 - diffusion. */
typedef struct {
    const orc_system *sys;
    const orc_bdopts *opt;
    const double *mob;
    const int *ex_ptr, *ex_col;
    orc_verlet *vs;
    int *pt_ptr, *pt_col;
    int ptcap;
    double *frc, *drift;
    orc_energies *en;
    const double *crd0, *dcol;
    int err;
} se_ctx;

static int se_fval(void *p, int n, const double *x, double *f)
{
    se_ctx *c = (se_ctx *)p;
    int e = drift_eval(c->sys, c->opt, c->mob, c->ex_ptr, c->ex_col, c->vs, c->pt_ptr, c->pt_col, c->ptcap, x,
                       c->frc, c->drift, c->en);
    if (e) {
        c->err = e;
        return 1;
    }
    for (int i = 0; i < n; i++) f[i] = x[i] - c->drift[i] - c->crd0[i] - c->dcol[i];
    return 0;
}

long long orc_bd_run(const orc_system *sys, const orc_bdopts *opt, double *crd,
                     int nsteps, const double *xi, long long nxi, orc_runinfo *info)
{
    return orc_bd_run_ex(sys, opt, crd, nsteps, xi, nxi, NULL, info);
}

/* disp (n, nsteps), if given, replaces the Brownian term of every step
 * (a recorded displacement replayed with the noise held fixed). */
long long orc_bd_run_ex(const orc_system *sys, const orc_bdopts *opt, double *crd,
                        int nsteps, const double *xi, long long nxi,
                        const double *disp, orc_runinfo *info)
{
    const int nb = sys->nbeads, n = 3 * nb;
    const int s = opt->lhdia ? opt->nts_mobsam : 1;
    const double sqrt_two_dt = sqrt(2.0 * opt->tim_stp);
    long long used = 0;
    double *mob = opt->lhdia ? (double *)malloc(sizeof(double) * (size_t)n * n) : NULL;
    double *diff = (double *)calloc((size_t)n * s, sizeof(double));
    double *drift = (double *)calloc((size_t)n, sizeof(double));
    double *frc = (double *)calloc((size_t)n, sizeof(double));
    double *rpred = opt->integrator == 1 ? (double *)calloc((size_t)n, sizeof(double)) : NULL;
    double *dpred = opt->integrator == 1 ? (double *)calloc((size_t)n, sizeof(double)) : NULL;
    int excap = nb * 64 + 16;
    int *ex_ptr = (int *)calloc((size_t)nb + 1, sizeof(int));
    int *ex_col = (int *)malloc(sizeof(int) * excap);
    int ptcap = 0;
    int *pt_ptr = (int *)calloc((size_t)nb + 1, sizeof(int));
    int *pt_col = NULL;
    orc_verlet vs = {1, NULL, 0.0, 0.0};
    int cntr = 0;
    const int limud = (sys->imcon == 0) && (sys->ntethers == 0);
    orc_energies en;
    long long ret = 0;

    memset(info, 0, sizeof(*info));
    if (sys->lvdw && sys->nvdw_types > 0) {
        if (orc_exclusions(nb, sys->nbonds, sys->bonds, sys->excluded_atoms, ex_ptr, ex_col, excap) < 0) {
            ret = -10;
            goto done;
        }
        ptcap = nb * 256 + 1024;
        pt_col = (int *)malloc(sizeof(int) * (size_t)ptcap);
        if (sys->mth_ptgen == 1) {
            vs.crd_save = (double *)calloc((size_t)n, sizeof(double));
            vs.rskin_sq = (sys->rcutoff + sys->tskin) * (sys->rcutoff + sys->tskin);
            vs.tskin_sq = sys->tskin * sys->tskin;
        }
    }
    if (limud) orc_to_origin(nb, crd, NULL);

    for (int step = 0; step < nsteps; step++) {
        /* ---- calc_diffusion ----
         * The reference zeroes the whole diffusion block on every call
         * (bd_solver_m.f90:364), so with s > 1 the steps that reuse the
         * mobility get no Brownian term.  keep_reuse_noise keeps the block. */
        if (!opt->lhdia || !opt->keep_reuse_noise || cntr == 0)
            memset(diff, 0, sizeof(double) * (size_t)n * s);
        if (disp) {
            if (opt->lhdia) {
                if (cntr == 0) {
                    orc_rpy(nb, crd, mob);
                    info->n_mob_updates++;
                }
                cntr++;
            }
        } else if (!opt->lhdia) {
            if (used + n > nxi) { ret = -11; goto done; }
            for (int i = 0; i < n; i++) diff[i] = sqrt_two_dt * xi[used + i];
            used += n;
        } else {
            if (cntr == 0) {
                orc_rpy(nb, crd, mob);
                if (used + (long long)n * s > nxi) { ret = -11; goto done; }
                memcpy(diff, xi + used, sizeof(double) * (size_t)n * s);
                used += (long long)n * s;
                int lconv = 1, iters = 0;
                int meth = opt->mob_fctr == 1 ? 1 : (opt->mob_fctr == 2 ? 2 : 0);
                int ierr = orc_brn_calc_bdw_ex(meth, n, s, mob, diff, opt->lanc_mxitr, opt->lanc_tol,
                                               opt->lanc_reorth, &lconv, &iters);
                if (ierr) { info->ierr = 2; ret = -2; goto done; }
                if (meth != 1)
                    for (int j = 0; j < n; j++) mob[IDX(j, j, n)] = 1.0;
                else {
                    info->last_lanczos_iters = iters;
                    info->total_lanczos_iters += iters;
                    if (!lconv) info->n_not_converged++;
                }
                info->n_mob_updates++;
                for (size_t k = 0; k < (size_t)n * s; k++) diff[k] = sqrt_two_dt * diff[k];
            }
            cntr++;
        }
        /* ---- integrate_se (bd_solver_m.f90:242-308): solve
         * x - drift(x) - crd0 - diffusion = 0 by NITSOL from x = crd0, with
         * input(1) = se_nlmxitr, input(4) = se_kdmax, input(7) = 1 ---- */
        if (opt->integrator == 2) {
            const double *dcol = disp ? disp + (size_t)step * n : diff + (size_t)(opt->lhdia ? cntr - 1 : 0) * n;
            double *crd0 = (double *)malloc(sizeof(double) * n);
            memcpy(crd0, crd, sizeof(double) * n);
            se_ctx sc = {sys, opt, mob, ex_ptr, ex_col, &vs, pt_ptr, pt_col, ptcap, frc, drift, &en, crd0, dcol, 0};
            int input[10] = {opt->se_nlmxitr, 0, 0, opt->se_kdmax, 0, 0, 1, 0, 0, 0};
            int nit[6];
            int iterm = orc_nitsol(n, crd, se_fval, &sc, opt->se_ftol, opt->se_stptol, input, nit);
            free(crd0);
            info->se_nfe += nit[0];
            info->se_nli += nit[3];
            info->se_nni += nit[4];
            info->se_nbt += nit[5];
            info->se_iterm = iterm;
            if (iterm != 0) { ret = -13; goto done; }
            if (opt->lhdia && cntr == s) cntr = 0;
            info->nts++;
            if (limud) {
                double com[3];
                orc_to_origin(nb, crd, com);
                for (int d = 0; d < 3; d++) info->molecule_com[d] += com[d];
            }
            if (sys->imcon) orc_wrap_all(sys->box, nb, crd);
            continue;
        }
        /* ---- calc_drift ---- */
        {
            int e = drift_eval(sys, opt, mob, ex_ptr, ex_col, &vs, pt_ptr, pt_col, ptcap, crd, frc, drift, &en);
            if (e == -1) { info->ierr = 1; ret = -1; goto done; }
            if (e) { ret = e; goto done; }
        }
        /* ---- predictor-corrector (not in the reference): r* = r + d(r) + w,
         * then the drift is the mean of d(r) and d(r*); D and w are those of
         * the step (the mobility is held over nts_mobsam steps anyway) ---- */
        if (opt->integrator == 1) {
            const double *dcol = disp ? disp + (size_t)step * n : diff + (size_t)(opt->lhdia ? cntr - 1 : 0) * n;
            for (int i = 0; i < n; i++) rpred[i] = crd[i] + drift[i] + dcol[i];
            if (sys->imcon) orc_wrap_all(sys->box, nb, rpred);
            int e = drift_eval(sys, opt, mob, ex_ptr, ex_col, &vs, pt_ptr, pt_col, ptcap, rpred, frc, dpred, &en);
            if (e == -1) { info->ierr = 1; ret = -1; goto done; }
            if (e) { ret = e; goto done; }
            for (int i = 0; i < n; i++) drift[i] = 0.5 * (drift[i] + dpred[i]);
        }
        /* ---- update ---- */
        {
            const double *dcol = disp ? disp + (size_t)step * n : diff + (size_t)(opt->lhdia ? cntr - 1 : 0) * n;
            for (int i = 0; i < n; i++) crd[i] = crd[i] + drift[i] + dcol[i];
        }
        if (opt->lhdia && cntr == s) cntr = 0;
        info->nts++;
        if (limud) {
            double com[3];
            orc_to_origin(nb, crd, com);
            for (int d = 0; d < 3; d++) info->molecule_com[d] += com[d];
        }
        if (sys->imcon) orc_wrap_all(sys->box, nb, crd);
    }
    ret = used;
done:
    free(mob); free(diff); free(drift); free(frc); free(rpred); free(dpred); free(ex_ptr); free(ex_col);
    free(pt_ptr); free(pt_col); free(vs.crd_save);
    return ret;
}
