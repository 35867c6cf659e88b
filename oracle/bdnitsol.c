/* bdnitsol.c -- CPU restatement of NITSOL v0.3 (src/nitsol/, vendored in the
 * reference) for the semi-implicit Euler step (integrate_se,
 * bd_solver_m.f90:242-308).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device's semi-implicit step.
 * Only tests/ and bench.py's cpu_baseline leg may use it.
 *
 * Restated: the inexact Newton driver with backtracking (nitdrv.f, nitbt.f),
 * the "simpler GMRES" Krylov solver with restarts (nitgm.f), finite-difference
 * Jacobian-vector products of order 1, 2 and 4 (nitfd in nitjv.f), the input
 * defaults of nitsol.f:400-520 and the parameters of nitbd.f/nitdflts.h.
 * Not restated (integrate_se never selects them): BiCGSTAB and TFQMR
 * (input(3) = 1, 2), analytic J*v (input(2) = 1), right preconditioning
 * (input(5) = 1); such inputs return -3, -2, -5 as illegal values.
 *
 * Two routines come from BLAS/LAPACK, which the reference takes from MKL
 * (not in the reference tree):
 *  - dnrm2/ddot: the Euclidean norm and inner product (plain sums here);
 *  - dlaic1 (LAPACK incremental condition estimation): given x, |x| = 1,
 *    with |L x| = sest for a lower-triangular L, it returns the extreme
 *    singular value sestpr of [L 0; w^T gamma] along [s x; c], i.e. the
 *    largest (job 1) or smallest (job 2) singular value of the 2 x 2 matrix
 *    M = [[sest, 0], [x.w, gamma]] and its right singular vector (s, c).
 *    Restated here from that definition: sigma_max from the eigenvalues of
 *    M^T M, sigma_min = |det M| / sigma_max.
 *  - dlamch('e') is the unit roundoff 2^-53, so epsmach = 2 dlamch('e') is
 *    DBL_EPSILON.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bdoracle.h"

/* nitdflts.h / nitbd.f */
#define NIT_CHOICE1_EXP (0.5 * (1.0 + 2.23606797749978981))
#define NIT_CHOICE2_EXP 2.0
#define NIT_CHOICE2_COEF 1.0
#define NIT_ETA_CUTOFF 0.1
#define NIT_ETA_MAX (1.0 - 1.0e-4)
#define NIT_THMIN 0.1
#define NIT_THMAX 0.5
#define NIT_ETA_FIXED 0.1
#define NIT_EPSMACH DBL_EPSILON

static double vdot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

static double vnrm2(int n, const double *a) { return sqrt(vdot(n, a, a)); }

static void vaxpy(int n, double alpha, const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] += alpha * x[i];
}

static void vscal(int n, double alpha, double *x)
{
    for (int i = 0; i < n; i++) x[i] *= alpha;
}

void orc_dlaic1(int job, int j, const double *x, double sest, const double *w, double gamma, double *sestpr,
                double *s, double *c)
{
    double alpha = vdot(j, x, w);
    /* M^T M = [[sest^2 + alpha^2, alpha gamma], [alpha gamma, gamma^2]] */
    double a = sest * sest + alpha * alpha, b = alpha * gamma, d = gamma * gamma;
    double tr = a + d, h = 0.5 * (a - d);
    double disc = sqrt(h * h + b * b);
    double lmax = 0.5 * tr + disc;
    double smax = sqrt(lmax);
    /* eigenvector of the larger eigenvalue: (b, lmax - a) or (lmax - d, b) */
    double v1, v2;
    if (a >= d) {
        v1 = lmax - d;
        v2 = b;
    } else {
        v1 = b;
        v2 = lmax - a;
    }
    double nv = hypot(v1, v2);
    if (nv == 0.0) {
        v1 = 1.0;
        v2 = 0.0;
    } else {
        v1 /= nv;
        v2 /= nv;
    }
    if (job == 1) {
        *sestpr = smax;
        *s = v1;
        *c = v2;
    } else {
        *sestpr = (smax > 0.0) ? fabs(sest * gamma) / smax : 0.0;
        *s = -v2;
        *c = v1;
    }
}

typedef struct {
    int n;
    orc_fn fn;
    void *ctx;
    int *nfe;
} nit_f;

static int call_f(nit_f *F, const double *x, double *f)
{
    int r = F->fn(F->ctx, F->n, x, f);
    return r;
}

/* nitfd (nitjv.f): z = J v by differences; v is overwritten (scratch).
 * order 1 within a GMRES cycle (ijacv 0), else ifdord. */
static int nit_fd(nit_f *F, const double *xcur, const double *fcur, int ijacv, int ifdord, double *v, double *z,
                  double *work)
{
    const int n = F->n;
    double eps = vnrm2(n, v);
    if (eps == 0.0) return 1;
    if (ijacv == 0 || ifdord == 1) {
        eps = sqrt((1.0 + vnrm2(n, xcur)) * NIT_EPSMACH) / eps;
        for (int i = 0; i < n; i++) v[i] = xcur[i] + eps * v[i];
        if (call_f(F, v, z)) return 1;
        (*F->nfe)++;
        for (int i = 0; i < n; i++) z[i] = (z[i] - fcur[i]) / eps;
        return 0;
    }
    if (ifdord == 2) {
        eps = pow((1.0 + vnrm2(n, xcur)) * NIT_EPSMACH, 1.0 / 3.0) / eps;
        for (int i = 0; i < n; i++) work[i] = xcur[i] + eps * v[i];
        if (call_f(F, work, z)) return 1;
        (*F->nfe)++;
        for (int i = 0; i < n; i++) work[i] = xcur[i] - eps * v[i];
        if (call_f(F, work, v)) return 1;   /* v now holds f(x - eps v) */
        (*F->nfe)++;
        double t = 2.0 * eps;
        for (int i = 0; i < n; i++) z[i] = (z[i] - v[i]) / t;
        return 0;
    }
    /* ifdord 4: (8 f(x + e/2 v) - 8 f(x - e/2 v) - f(x + e v) + f(x - e v)) / (6 e);
     * the reference shifts xcur in place and back, here a copy is shifted */
    eps = pow((1.0 + vnrm2(n, xcur)) * NIT_EPSMACH, 1.0 / 5.0) / eps;
    double *xs = (double *)calloc((size_t)n, sizeof(double)), *fv = (double *)calloc((size_t)n, sizeof(double));
    int rc = 1;
    for (int i = 0; i < n; i++) xs[i] = xcur[i] + eps * v[i];
    if (call_f(F, xs, z)) goto out;
    (*F->nfe)++;
    for (int i = 0; i < n; i++) xs[i] = xcur[i] - eps * v[i];
    if (call_f(F, xs, fv)) goto out;
    (*F->nfe)++;
    for (int i = 0; i < n; i++) z[i] = fv[i] - z[i];
    for (int i = 0; i < n; i++) xs[i] = xcur[i] - 0.5 * eps * v[i];
    if (call_f(F, xs, fv)) goto out;
    (*F->nfe)++;
    vaxpy(n, -8.0, fv, z);
    for (int i = 0; i < n; i++) xs[i] = xcur[i] + 0.5 * eps * v[i];
    if (call_f(F, xs, fv)) goto out;
    (*F->nfe)++;
    vaxpy(n, 8.0, fv, z);
    vscal(n, 1.0 / (6.0 * eps), z);
    rc = 0;
out:
    free(xs);
    free(fv);
    (void)work;
    return rc;
}

/* nitjv for itask 0 without preconditioning: z = J v (v unchanged). */
static int nit_jv(nit_f *F, const double *xcur, const double *fcur, int ijacv, int ifdord, const double *v,
                  double *z, double *w1, double *w2, int *njve)
{
    memcpy(w1, v, sizeof(double) * F->n);
    if (nit_fd(F, xcur, fcur, ijacv, ifdord, w1, z, w2)) return 1;
    (*njve)++;
    return 0;
}

/* nitgm.f: the simpler GMRES of Walker and Zhou with restarts.  vv
 * (n, kdmax+1) column-major; vv(:,1) ends as the linear residual. */
static int nit_gm(nit_f *F, const double *xcur, const double *fcur, double fcnrm, double *step, double eta,
                  int iksmax, int iresup, int ifdord, int *njve, int *nli, int kdmax, double *vv, double *rr,
                  double *svbig, double *svsml, double *w, double *rwork, double *scr, double *rsnrm_out)
{
    const int n = F->n;
    const double cndmax = 1.0 / (100.0 * NIT_EPSMACH);
    int itrmks = 0, igm = 0, kd = 0, kdp1 = 0, ijacv = 0;
    double big = 0.0, small = 0.0, rsnrm = 1.0;
#define VV(k) (vv + (size_t)(k) * n) /* 0-based column */
#define RR(i, j) rr[(size_t)(j) * kdmax + (i)]
    memset(step, 0, sizeof(double) * n);
    double rsnrm0 = fcnrm;
    const double abstol = eta * rsnrm0;
    memcpy(VV(0), fcur, sizeof(double) * n);
    vscal(n, -1.0 / fcnrm, VV(0));
    for (;;) {
        /* label 100: a new cycle */
        kd = 0;
        rsnrm = 1.0;
        for (;;) {
            /* label 200 */
            kd++;
            kdp1 = kd + 1;
            (*nli)++;
            igm++;
            if (nit_jv(F, xcur, fcur, ijacv, ifdord, VV(kd - 1), VV(kdp1 - 1), rwork, scr, njve)) {
                itrmks = 1;
                goto done;
            }
            for (int i = 2; i <= kd; i++) {
                RR(i - 2, kd - 1) = vdot(n, VV(i - 1), VV(kdp1 - 1));
                vaxpy(n, -RR(i - 2, kd - 1), VV(i - 1), VV(kdp1 - 1));
            }
            RR(kd - 1, kd - 1) = vnrm2(n, VV(kdp1 - 1));
            if (kd == 1) {
                big = RR(0, 0);
                small = big;
                svbig[0] = 1.0;
                svsml[0] = 1.0;
            } else {
                double sestpr, sn, cs;
                orc_dlaic1(1, kd - 1, svbig, big, &RR(0, kd - 1), RR(kd - 1, kd - 1), &sestpr, &sn, &cs);
                big = sestpr;
                vscal(kd - 1, sn, svbig);
                svbig[kd - 1] = cs;
                orc_dlaic1(2, kd - 1, svsml, small, &RR(0, kd - 1), RR(kd - 1, kd - 1), &sestpr, &sn, &cs);
                small = sestpr;
                vscal(kd - 1, sn, svsml);
                svsml[kd - 1] = cs;
            }
            if (big >= small * cndmax) {
                if (kd == 1) {
                    itrmks = 5;
                    goto done;
                }
                kdp1 = kd;
                kd = kd - 1;
                vaxpy(n, w[kd - 1], VV(kdp1 - 1), VV(0));
                break;   /* label 300 */
            }
            vscal(n, 1.0 / RR(kd - 1, kd - 1), VV(kdp1 - 1));
            w[kd - 1] = vdot(n, VV(0), VV(kdp1 - 1));
            double t = fmax(fmin(w[kd - 1] / rsnrm, 1.0), -1.0);
            rsnrm = rsnrm * sin(acos(t));
            if (rsnrm0 * rsnrm <= abstol || kd == kdmax || igm >= iksmax) break;
            vaxpy(n, -w[kd - 1], VV(kdp1 - 1), VV(0));
        }
        /* label 300: the step of this cycle */
        memcpy(svbig, w, sizeof(double) * kd);
        for (int i = kd; i >= 1; i--) {
            w[i - 1] = w[i - 1] / RR(i - 1, i - 1);
            if (i > 1) vaxpy(i - 1, -w[i - 1], &RR(0, i - 1), w);
        }
        memcpy(rwork, VV(0), sizeof(double) * n);
        vscal(n, w[0], rwork);
        if (kd > 1) {
            vaxpy(kd - 1, w[0], svbig, w + 1);
            for (int i = 2; i <= kd; i++) vaxpy(n, w[i - 1], VV(i - 1), rwork);
        }
        if (iresup == 0) vaxpy(n, -svbig[kd - 1], VV(kdp1 - 1), VV(0));
        vaxpy(n, rsnrm0, rwork, step);
        if (iresup == 1) {
            /* J*step with the restart order (ijacv -1 selects ifdord) */
            if (nit_jv(F, xcur, fcur, -1, ifdord, step, VV(0), rwork, VV(kdp1 - 1), njve)) {
                itrmks = 1;
                goto done;
            }
            for (int i = 0; i < n; i++) VV(0)[i] = -fcur[i] - VV(0)[i];
        }
        if (rsnrm0 * rsnrm <= abstol) {
            itrmks = 0;
            goto done;
        }
        if (igm >= iksmax) {
            itrmks = 3;
            goto done;
        }
        if (big >= small * cndmax) {
            itrmks = 5;
            goto done;
        }
        {
            double t = (double)kd * log(abstol / (rsnrm0 * rsnrm)) / log(rsnrm / (1.0 + 10.0 * NIT_EPSMACH));
            if (t >= 1000.0 * (double)(iksmax - igm)) {
                itrmks = 4;
                goto done;
            }
        }
        double t;
        if (iresup == 0) {
            rsnrm0 = rsnrm0 * rsnrm;
            t = 1.0 / rsnrm;
        } else {
            rsnrm0 = vnrm2(n, VV(0));
            t = 1.0 / rsnrm0;
        }
        vscal(n, t, VV(0));
    }
done:
    if (itrmks != 1 && itrmks != 2) {
        if (iresup == 0) {
            vscal(n, rsnrm0, VV(0));
            rsnrm = rsnrm0 * rsnrm;
        } else {
            rsnrm = vnrm2(n, VV(0));
        }
    }
    *rsnrm_out = rsnrm;
    return itrmks;
#undef VV
#undef RR
}

/* nitbt.f: backtracking along step; returns itrmbt. */
static int nit_bt(nit_f *F, const double *xcur, double fcnrm, double *step, double *eta, double *xpls,
                  double *fpls, double *fpnrm, double oftjs, double *redfac, int *ibt, int ibtmax)
{
    const int n = F->n;
    const double t = 1.0e-4;
    *ibt = 0;
    *redfac = 1.0;
    for (;;) {
        for (int i = 0; i < n; i++) xpls[i] = xcur[i] + step[i];
        if (call_f(F, xpls, fpls)) return 2;
        (*F->nfe)++;
        *fpnrm = vnrm2(n, fpls);
        if (*fpnrm <= (1.0 - t * (1.0 - *eta)) * fcnrm || ibtmax == -1) return 0;
        (*ibt)++;
        if (*ibt > ibtmax) return 1;
        double theta = -(oftjs * *redfac) / ((*fpnrm) * (*fpnrm) - fcnrm * fcnrm - 2.0 * oftjs * *redfac);
        if (theta < NIT_THMIN) theta = NIT_THMIN;
        if (theta > NIT_THMAX) theta = NIT_THMAX;
        vscal(n, theta, step);
        *eta = 1.0 - theta * (1.0 - *eta);
        *redfac = theta * *redfac;
    }
}

int orc_nitsol(int n, double *x, orc_fn fn, void *ctx, double ftol, double stptol, const int *input, int *info)
{
    int nnimax, ijacv, ikrysl, kdmax = 20, irpre, iksmax, iresup = 0, ifdord = 1, ibtmax, ieta;
    memset(info, 0, sizeof(int) * 6);
    /* nitsol.f: input defaults and checks */
    if (input[0] == 0) nnimax = 200;
    else if (input[0] > 0) nnimax = input[0];
    else return -1;
    if (input[1] != 0) return -2;   /* only finite-difference J*v */
    ijacv = 0;
    if (input[2] != 0) return -3;   /* only GMRES */
    ikrysl = 0;
    if (input[3] == 0) kdmax = 20;
    else if (input[3] > 0) kdmax = input[3];
    else return -4;
    if (input[4] != 0) return -5;   /* no right preconditioning */
    irpre = 0;
    if (input[5] == 0) iksmax = 1000;
    else if (input[5] > 0) iksmax = input[5];
    else return -6;
    if (input[6] == 0 || input[6] == 1) iresup = input[6];
    else return -7;
    if (input[7] == 0) ifdord = 2;
    else if (input[7] == 1 || input[7] == 2 || input[7] == 4) ifdord = input[7];
    else return -8;
    if (input[8] == 0) ibtmax = 10;
    else if (input[8] > 0 || input[8] == -1) ibtmax = input[8];
    else return -9;
    if (input[9] >= 0 && input[9] <= 3) ieta = input[9];
    else return -10;
    (void)ijacv;
    (void)ikrysl;
    (void)irpre;

    double *fcur = (double *)malloc(sizeof(double) * n), *xpls = (double *)malloc(sizeof(double) * n);
    double *fpls = (double *)malloc(sizeof(double) * n), *step = (double *)malloc(sizeof(double) * n);
    double *scr = (double *)malloc(sizeof(double) * n);
    double *vv = (double *)malloc(sizeof(double) * (size_t)n * (kdmax + 1));
    double *rr = (double *)calloc((size_t)kdmax * kdmax, sizeof(double));
    double *svbig = (double *)malloc(sizeof(double) * kdmax), *svsml = (double *)malloc(sizeof(double) * kdmax);
    double *w = (double *)malloc(sizeof(double) * kdmax);
    int nfe = 0, njve = 0, nli = 0, nni = 0, nbt = 0, iterm = 0, itrmks = 0;
    nit_f F = {n, fn, ctx, &nfe};
    double alpha = 0.0, gamma = 0.0, eta, etamin = 0.0, fcnrm, stpnrm = 0.0, rsnrm = 0.0;

    /* nitdrv.f */
    if (ieta == 0) alpha = NIT_CHOICE1_EXP;
    if (ieta == 2) {
        alpha = NIT_CHOICE2_EXP;
        gamma = NIT_CHOICE2_COEF;
    }
    eta = (ieta == 3) ? NIT_ETA_FIXED : 0.5;
    if (call_f(&F, x, fcur)) {
        iterm = 2;
        goto out;
    }
    nfe++;
    fcnrm = vnrm2(n, fcur);
    for (;;) {
        if (fcnrm <= ftol) {
            iterm = 0;
            break;
        }
        if (nni > 0 && stpnrm <= stptol && itrmks == 0) {
            iterm = 0;
            break;
        }
        if (nni >= nnimax) {
            iterm = 1;
            break;
        }
        itrmks = nit_gm(&F, x, fcur, fcnrm, step, eta, iksmax, iresup, ifdord, &njve, &nli, kdmax, vv, rr, svbig,
                        svsml, w, fpls, scr, &rsnrm);
        if (itrmks == 1 || itrmks == 2) {
            iterm = itrmks + 2;
            break;
        }
        if (itrmks >= 3) {
            if (rsnrm / fcnrm >= 1.0) {
                iterm = 5;
                break;
            }
            double t = log(ftol / fcnrm) / log(rsnrm / ((1.0 + 10.0 * NIT_EPSMACH) * fcnrm));
            if (t > 1000.0 * (double)(nnimax - nni)) {
                iterm = 5;
                break;
            }
        }
        double oftlm = -vdot(n, fcur, vv);
        double oftjs = oftlm - fcnrm * fcnrm;
        double fpnrm, redfac;
        int ibt;
        int itrmbt = nit_bt(&F, x, fcnrm, step, &eta, xpls, fpls, &fpnrm, oftjs, &redfac, &ibt, ibtmax);
        if (itrmbt == 1) {
            iterm = 6;
            break;
        }
        if (itrmbt == 2) {
            iterm = 2;
            break;
        }
        nbt += ibt;
        if (ieta == 0) {
            etamin = pow(eta, alpha);
            double t = 1.0 - redfac;
            double flmnrm = sqrt((t * fcnrm) * (t * fcnrm) + 2.0 * redfac * t * oftlm + (redfac * rsnrm) * (redfac * rsnrm));
            eta = fabs(fpnrm - flmnrm) / fcnrm;
        }
        if (ieta == 1) {
            etamin = eta * eta;
            eta = (fpnrm / fcnrm) * (fpnrm / fcnrm);
        }
        if (ieta == 2) {
            etamin = gamma * pow(eta, alpha);
            eta = gamma * pow(fpnrm / fcnrm, alpha);
        }
        if (ieta != 3) {
            if (etamin <= NIT_ETA_CUTOFF) etamin = 0.0;
            if (eta < etamin) eta = etamin;
            if (eta > NIT_ETA_MAX) eta = NIT_ETA_MAX;
            if (eta * fpnrm <= 2.0 * ftol) eta = (0.8 * ftol) / fpnrm;
        }
        if (ieta == 3) eta = NIT_ETA_FIXED;
        memcpy(x, xpls, sizeof(double) * n);
        memcpy(fcur, fpls, sizeof(double) * n);
        fcnrm = fpnrm;
        stpnrm = vnrm2(n, step);
        nni++;
    }
out:
    info[0] = nfe;
    info[1] = njve;
    info[2] = 0;
    info[3] = nli;
    info[4] = nni;
    info[5] = nbt;
    free(fcur); free(xpls); free(fpls); free(step); free(scr); free(vv); free(rr); free(svbig); free(svsml); free(w);
    return iterm;
}
