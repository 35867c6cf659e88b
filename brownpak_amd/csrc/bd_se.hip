// bd_se.hip -- the semi-implicit Euler step (integrate_se,
// bd_solver_m.f90:242-332): x - drift(x) - crd0 - w = 0 solved for the new
// coordinates x with the inexact Newton method of NITSOL v0.3 as
// integrate_se calls it (src/nitsol: nitdrv.f, nitgm.f, nitjv.f, nitbt.f):
// GMRES in its "simpler GMRES" form with restarts every se_kdmax vectors,
// finite-difference Jacobian-vector products (first order inside a cycle,
// second order for the residual at a restart, input(7) = 1), backtracking,
// and the Eisenstat-Walker forcing terms (choice 1) with NITSOL's
// safeguards.  drift(x) = dt (U(x) + D F(x)) uses the step's mobility D
// (held fixed over the solve, as calc_drift does) and the device force and
// product kernels.
//
// Every n-vector stays on the device; the Newton/GMRES scalars (the small
// triangular factor, the condition estimates, the forcing term) live on the
// host, which reads the few reductions it decides on after each Krylov
// vector.  Reductions run in one workgroup in a fixed order, so a solve
// repeats bit for bit.
#include <cfloat>
#include <cmath>
#include <vector>

#include "bd_internal.h"

namespace bd {

constexpr int SW = 1024;   // threads of the single-workgroup reductions

__device__ inline double wg_sum(double v, double *sh)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int h = SW / 2; h > 0; h >>= 1) {
        if (t < h) sh[t] += sh[t + h];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// out[slot] = a . b, or its square root (root: the norm of a = b)
__global__ __launch_bounds__(SW) void se_dot_kernel(const double *__restrict__ a, const double *__restrict__ b, int n,
                                                    double *out, int slot, int root)
{
    __shared__ double sh[SW];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += SW) s += a[i] * b[i];
    s = wg_sum(s, sh);
    if (threadIdx.x == 0) out[slot] = root ? sqrt(s) : s;
}

// v0 = -fcur / fcnrm (as nitgm.f: dcopy then dscal by -1/fcnrm);
// out[slot] = |v0|
__global__ __launch_bounds__(SW) void se_start_kernel(const double *__restrict__ fcur, double fcnrm, int n,
                                                      double *__restrict__ v0, double *out, int slot)
{
    __shared__ double sh[SW];
    const double t = -1.0 / fcnrm;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += SW) {
        const double v = fcur[i] * t;
        v0[i] = v;
        s += v * v;
    }
    s = wg_sum(s, sh);
    if (threadIdx.x == 0) out[slot] = sqrt(s);
}

// v *= alpha; out[slot] = |v| after scaling
__global__ __launch_bounds__(SW) void se_scal_norm_kernel(double *__restrict__ v, double alpha, int n, double *out,
                                                          int slot)
{
    __shared__ double sh[SW];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += SW) {
        const double x = v[i] * alpha;
        v[i] = x;
        s += x * x;
    }
    s = wg_sum(s, sh);
    if (threadIdx.x == 0) out[slot] = sqrt(s);
}

// One GMRES iteration after z = J v_kd landed in column kd (0-based): the
// modified Gram-Schmidt of z against columns 1..kd-1 (rr column out[0..kd-2]),
// rr(kd,kd) = |z| (out[kd-1]), z /= rr(kd,kd), w = v0 . z (out[kd]) and
// |z| after scaling (out[kd+1]) (nitgm.f, labels 200-210).
__global__ __launch_bounds__(SW) void se_gm_kernel(double *__restrict__ vv, long ld, int n, int kd, double *out)
{
    __shared__ double sh[SW];
    double *z = vv + (long)kd * ld;
    for (int i = 1; i < kd; i++) {
        const double *vi = vv + (long)i * ld;
        double s = 0.0;
        for (int k = threadIdx.x; k < n; k += SW) s += vi[k] * z[k];
        const double r = wg_sum(s, sh);
        for (int k = threadIdx.x; k < n; k += SW) z[k] += -r * vi[k];
        if (threadIdx.x == 0) out[i - 1] = r;
        __syncthreads();
    }
    double s = 0.0;
    for (int k = threadIdx.x; k < n; k += SW) s += z[k] * z[k];
    const double nrm = sqrt(wg_sum(s, sh));
    const double t = 1.0 / nrm;
    double sw = 0.0, sz = 0.0;
    for (int k = threadIdx.x; k < n; k += SW) {
        const double x = z[k] * t;
        z[k] = x;
        sw += vv[k] * x;
        sz += x * x;
    }
    const double w = wg_sum(sw, sh);
    const double zn = wg_sum(sz, sh);
    if (threadIdx.x == 0) {
        out[kd - 1] = nrm;
        out[kd] = w;
        out[kd + 1] = sqrt(zn);
    }
}

__global__ void se_axpy_kernel(double a, const double *__restrict__ x, double *__restrict__ y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

// y = a * x
__global__ void se_scale_copy_kernel(double a, const double *__restrict__ x, double *__restrict__ y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] * a;
}

// z = x + a * v
__global__ void se_shift_kernel(const double *__restrict__ x, double a, const double *__restrict__ v,
                                double *__restrict__ z, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = x[i] + a * v[i];
}

// se_fval (bd_solver_m.f90:312-332): f = x - drift - crd0 - w
__global__ void se_f_kernel(const double *__restrict__ x, const double *__restrict__ drift,
                            const double *__restrict__ crd0, const double *__restrict__ w, double *__restrict__ f, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = x[i] - drift[i] - crd0[i];
    if (w) v = v - w[i];
    f[i] = v;
}

// first-order difference: z = (z - fcur) / eps
__global__ void se_fd1_kernel(double *__restrict__ z, const double *__restrict__ fcur, double eps, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = (z[i] - fcur[i]) / eps;
}

// second-order difference: z = (z - fm) / (2 eps)
__global__ void se_fd2_kernel(double *__restrict__ z, const double *__restrict__ fm, double twoeps, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = (z[i] - fm[i]) / twoeps;
}

// the linear residual after a restart: v0 = -fcur - J step
__global__ void se_resid_kernel(double *__restrict__ v0, const double *__restrict__ fcur, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v0[i] = -fcur[i] - v0[i];
}

// pvcrd = sol and the bds_run tail's PBC wrap; counts the step
__global__ void se_finish_kernel(const double *__restrict__ x, double *__restrict__ crd, int nb, int imcon, double Lx,
                                 double Ly, double Lz, DevStatus *st)
{
    if (st->ierr) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    if (i == 0) st->steps_done += 1;
    const double L[3] = {Lx, Ly, Lz};
    for (int d = 0; d < 3; d++) {
        double r = x[3 * i + d];
        if (imcon) r = r - L[d] * floor(r / L[d]);  // smbx_wrap_all, simbox_m.f90:128-153
        crd[3 * i + d] = r;
    }
}

// dlaic1 (LAPACK incremental condition estimation), from its definition:
// the extreme singular value of M = [[sest, 0], [x.w, gamma]] and its right
// singular vector (s, c); job 1 largest, job 2 smallest.
static void laic1(int job, int j, const double *x, double sest, const double *w, double gamma, double &sestpr,
                  double &s, double &c)
{
    double alpha = 0.0;
    for (int k = 0; k < j; k++) alpha += x[k] * w[k];
    const double a = sest * sest + alpha * alpha, b = alpha * gamma, d = gamma * gamma;
    const double h = 0.5 * (a - d);
    const double lmax = 0.5 * (a + d) + std::sqrt(h * h + b * b);
    const double smax = std::sqrt(lmax);
    double v1, v2;
    if (a >= d) {
        v1 = lmax - d;
        v2 = b;
    } else {
        v1 = b;
        v2 = lmax - a;
    }
    const double nv = std::hypot(v1, v2);
    if (nv == 0.0) {
        v1 = 1.0;
        v2 = 0.0;
    } else {
        v1 /= nv;
        v2 /= nv;
    }
    if (job == 1) {
        sestpr = smax;
        s = v1;
        c = v2;
    } else {
        sestpr = smax > 0.0 ? std::fabs(sest * gamma) / smax : 0.0;
        s = -v2;
        c = v1;
    }
}

namespace {

// NITSOL parameters (nitdflts.h, nitbd.f) and the machine epsilon of its
// dlamch('e') (unit roundoff 2^-53) doubled.
constexpr double CHOICE1_EXP = 0.5 * (1.0 + 2.23606797749978981);
constexpr double ETA_CUTOFF = 0.1, ETA_MAX = 1.0 - 1.0e-4, THMIN = 0.1, THMAX = 0.5;
constexpr double EPSMACH = DBL_EPSILON;

struct Se {
    Ctx &c;
    const double *crd0, *w;
    int n;
    unsigned nbk;
    hipStream_t sm;
    double *vv, *xcur, *fcur, *xpls, *fpls, *step, *rwork, *scr, *sc;
    double *hsc;
    int nfe = 0, njve = 0, nli = 0, nni = 0, nbt = 0;
    bool ffail = false;

    double *col(int k) { return vv + (size_t)k * c.npad; }

    // F(x) -> f on the device
    int F(const double *x, double *f)
    {
        BD_HIP(hipMemcpyAsync(c.crd, x, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, sm));
        BD_TRY(compute_forces(c, c.stats_on, false));
        const double *v = c.frc;
        if (c.opt.lhdia) {
            BD_TRY(mob_product(c, c.frc, c.W, 1, c.npad, c.npad));
            v = c.W;
        }
        BD_TRY(launch_drift(c, v));
        hipLaunchKernelGGL(se_f_kernel, dim3(nbk), dim3(256), 0, sm, x, c.drift, crd0, w, f, n);
        return BD_OK;
    }
    // k scalars of sc to the host; notes a failed force evaluation
    int read(int k)
    {
        BD_HIP(hipMemcpyAsync(hsc, sc, (size_t)k * sizeof(double), hipMemcpyDeviceToHost, sm));
        BD_HIP(hipMemcpyAsync(c.st_host, c.st, sizeof(DevStatus), hipMemcpyDeviceToHost, sm));
        BD_HIP(hipStreamSynchronize(sm));
        if (c.st_host->ierr) ffail = true;
        return BD_OK;
    }
    void axpy(double a, const double *x, double *y)
    {
        hipLaunchKernelGGL(se_axpy_kernel, dim3(nbk), dim3(256), 0, sm, a, x, y, n);
    }
    void nrm_to(const double *a, int slot)
    {
        hipLaunchKernelGGL(se_dot_kernel, dim3(1), dim3(SW), 0, sm, a, a, n, sc, slot, 1);
    }
    void dot_to(const double *a, const double *b, int slot)
    {
        hipLaunchKernelGGL(se_dot_kernel, dim3(1), dim3(SW), 0, sm, a, b, n, sc, slot, 0);
    }
};

}  // namespace

// nitgm.f with iresup 1 (integrate_se's input(7)): itrmks as nitgm's; the
// final linear residual -(F + J step) in column 0, its norm in rsnrm.
// Returns a BD_* status.
static int se_gmres(Se &S, double fcnrm, double eta, double xnorm, int kdmax, int iksmax, double &rsnrm,
                    int &itrmks)
{
    Ctx &c = S.c;
    const int n = S.n;
    const double cndmax = 1.0 / (100.0 * EPSMACH);
    std::vector<double> rr((size_t)kdmax * kdmax, 0.0), svbig(kdmax), svsml(kdmax), w(kdmax);
    auto RR = [&](int i, int j) -> double & { return rr[(size_t)j * kdmax + i]; };
    int igm = 0, kd = 0, kdp1 = 0;
    itrmks = 0;
    double big = 0.0, small = 0.0, r0norm = 0.0;
    bool have_r0norm = false;   // r0norm is |v0| of the current column 0
    BD_HIP(hipMemsetAsync(S.step, 0, (size_t)n * sizeof(double), S.sm));
    double rsnrm0 = fcnrm;
    const double abstol = eta * rsnrm0;
    hipLaunchKernelGGL(se_start_kernel, dim3(1), dim3(SW), 0, S.sm, S.fcur, fcnrm, n, S.col(0), S.sc, 0);
    BD_TRY(S.read(1));
    double vnorm = S.hsc[0];   // |v_kd| for the next difference
    for (;;) {
        // label 100: a cycle
        kd = 0;
        rsnrm = 1.0;
        for (;;) {
            // label 200
            kd++;
            kdp1 = kd + 1;
            S.nli++;
            igm++;
            // nitfd, first order: z = (F(x + eps v) - F(x)) / eps
            if (vnorm == 0.0) { itrmks = 1; goto done; }
            const double eps = std::sqrt((1.0 + xnorm) * EPSMACH) / vnorm;
            hipLaunchKernelGGL(se_shift_kernel, dim3(S.nbk), dim3(256), 0, S.sm, S.xcur, eps, S.col(kd - 1), S.scr,
                               n);
            BD_TRY(S.F(S.scr, S.col(kd)));
            S.nfe++;
            hipLaunchKernelGGL(se_fd1_kernel, dim3(S.nbk), dim3(256), 0, S.sm, S.col(kd), S.fcur, eps, n);
            hipLaunchKernelGGL(se_gm_kernel, dim3(1), dim3(SW), 0, S.sm, S.vv, (long)c.npad, n, kd, S.sc);
            BD_TRY(S.read(kd + 2));
            if (S.ffail) { itrmks = 1; goto done; }
            S.njve++;
            for (int i = 0; i < kd - 1; i++) RR(i, kd - 1) = S.hsc[i];
            RR(kd - 1, kd - 1) = S.hsc[kd - 1];
            const double wk = S.hsc[kd];
            vnorm = S.hsc[kd + 1];
            have_r0norm = false;
            if (kd == 1) {
                big = RR(0, 0);
                small = big;
                svbig[0] = 1.0;
                svsml[0] = 1.0;
            } else {
                double sestpr, sn, cs;
                laic1(1, kd - 1, svbig.data(), big, &RR(0, kd - 1), RR(kd - 1, kd - 1), sestpr, sn, cs);
                big = sestpr;
                for (int i = 0; i < kd - 1; i++) svbig[i] *= sn;
                svbig[kd - 1] = cs;
                laic1(2, kd - 1, svsml.data(), small, &RR(0, kd - 1), RR(kd - 1, kd - 1), sestpr, sn, cs);
                small = sestpr;
                for (int i = 0; i < kd - 1; i++) svsml[i] *= sn;
                svsml[kd - 1] = cs;
            }
            if (big >= small * cndmax) {
                if (kd == 1) { itrmks = 5; goto done; }
                kdp1 = kd;
                kd = kd - 1;
                S.axpy(w[kd - 1], S.col(kdp1 - 1), S.col(0));
                break;
            }
            w[kd - 1] = wk;
            const double t = std::fmax(std::fmin(wk / rsnrm, 1.0), -1.0);
            rsnrm = rsnrm * std::sin(std::acos(t));
            if (rsnrm0 * rsnrm <= abstol || kd == kdmax || igm >= iksmax) break;
            S.axpy(-wk, S.col(kdp1 - 1), S.col(0));
        }
        // label 300: the step of this cycle from the small triangular system
        {
            std::vector<double> sv(w.begin(), w.begin() + kd);
            for (int i = kd; i >= 1; i--) {
                w[i - 1] = w[i - 1] / RR(i - 1, i - 1);
                if (i > 1)
                    for (int k = 0; k < i - 1; k++) w[k] += -w[i - 1] * RR(k, i - 1);
            }
            hipLaunchKernelGGL(se_scale_copy_kernel, dim3(S.nbk), dim3(256), 0, S.sm, w[0], S.col(0), S.rwork, n);
            if (kd > 1) {
                for (int k = 0; k < kd - 1; k++) w[k + 1] += w[0] * sv[k];
                for (int i = 2; i <= kd; i++) S.axpy(w[i - 1], S.col(i - 1), S.rwork);
            }
            for (int i = 0; i < kd; i++) svbig[i] = sv[i];
        }
        S.axpy(rsnrm0, S.rwork, S.step);
        // iresup 1: v0 = -F - J step, J step by second-order differences
        {
            S.nrm_to(S.step, 0);
            BD_TRY(S.read(1));
            const double snorm = S.hsc[0];
            if (snorm == 0.0) { itrmks = 1; goto done; }
            const double eps = std::pow((1.0 + xnorm) * EPSMACH, 1.0 / 3.0) / snorm;
            hipLaunchKernelGGL(se_shift_kernel, dim3(S.nbk), dim3(256), 0, S.sm, S.xcur, eps, S.step, S.scr, n);
            BD_TRY(S.F(S.scr, S.col(0)));
            S.nfe++;
            hipLaunchKernelGGL(se_shift_kernel, dim3(S.nbk), dim3(256), 0, S.sm, S.xcur, -eps, S.step, S.scr, n);
            BD_TRY(S.F(S.scr, S.rwork));
            S.nfe++;
            hipLaunchKernelGGL(se_fd2_kernel, dim3(S.nbk), dim3(256), 0, S.sm, S.col(0), S.rwork, 2.0 * eps, n);
            hipLaunchKernelGGL(se_resid_kernel, dim3(S.nbk), dim3(256), 0, S.sm, S.col(0), S.fcur, n);
            S.nrm_to(S.col(0), 0);
            BD_TRY(S.read(1));
            if (S.ffail) { itrmks = 1; goto done; }
            S.njve++;
            r0norm = S.hsc[0];
            have_r0norm = true;
        }
        if (rsnrm0 * rsnrm <= abstol) { itrmks = 0; goto done; }
        if (igm >= iksmax) { itrmks = 3; goto done; }
        if (big >= small * cndmax) { itrmks = 5; goto done; }
        {
            const double t = (double)kd * std::log(abstol / (rsnrm0 * rsnrm)) /
                             std::log(rsnrm / (1.0 + 10.0 * EPSMACH));
            if (t >= 1000.0 * (double)(iksmax - igm)) { itrmks = 4; goto done; }
        }
        // restart: v0 /= |v0|
        rsnrm0 = r0norm;
        hipLaunchKernelGGL(se_scal_norm_kernel, dim3(1), dim3(SW), 0, S.sm, S.col(0), 1.0 / rsnrm0, n, S.sc, 0);
        BD_TRY(S.read(1));
        vnorm = S.hsc[0];
        have_r0norm = false;
    }
done:
    if (itrmks != 1 && itrmks != 2) {
        if (!have_r0norm) {
            S.nrm_to(S.col(0), 0);
            BD_TRY(S.read(1));
            r0norm = S.hsc[0];
        }
        rsnrm = r0norm;
    }
    return BD_OK;
}

// integrate_se's nitsol call: iterm as nitsol's (0 converged).  Returns a
// BD_* status.
static int se_nitsol(Se &S, int &iterm)
{
    Ctx &c = S.c;
    const int n = S.n;
    const int nnimax = c.opt.se_nlmxitr > 0 ? c.opt.se_nlmxitr : 200;
    const int kdmax = c.opt.se_kdmax > 0 ? c.opt.se_kdmax : 20;
    const int iksmax = 1000, ibtmax = 10;
    const double ftol = c.opt.se_ftol, stptol = c.opt.se_stptol;
    double eta = 0.5, stpnrm = 0.0, rsnrm = 0.0;
    int itrmks = 0;
    // nitdrv.f: F at the initial guess
    BD_TRY(S.F(S.xcur, S.fcur));
    S.nfe++;
    S.nrm_to(S.fcur, 0);
    S.nrm_to(S.xcur, 1);
    BD_TRY(S.read(2));
    if (S.ffail) { iterm = 2; return BD_OK; }
    double fcnrm = S.hsc[0], xnorm = S.hsc[1];
    for (;;) {
        if (fcnrm <= ftol) { iterm = 0; return BD_OK; }
        if (S.nni > 0 && stpnrm <= stptol && itrmks == 0) { iterm = 0; return BD_OK; }
        if (S.nni >= nnimax) { iterm = 1; return BD_OK; }
        BD_TRY(se_gmres(S, fcnrm, eta, xnorm, kdmax, iksmax, rsnrm, itrmks));
        if (itrmks == 1 || itrmks == 2) { iterm = itrmks + 2; return BD_OK; }
        if (itrmks >= 3) {
            if (rsnrm / fcnrm >= 1.0) { iterm = 5; return BD_OK; }
            const double t = std::log(ftol / fcnrm) / std::log(rsnrm / ((1.0 + 10.0 * EPSMACH) * fcnrm));
            if (t > 1000.0 * (double)(nnimax - S.nni)) { iterm = 5; return BD_OK; }
        }
        S.dot_to(S.fcur, S.col(0), 0);
        BD_TRY(S.read(1));
        const double oftlm = -S.hsc[0];
        const double oftjs = oftlm - fcnrm * fcnrm;
        // nitbt.f: backtracking
        double fpnrm = 0.0, redfac = 1.0;
        int ibt = 0;
        for (;;) {
            hipLaunchKernelGGL(se_shift_kernel, dim3(S.nbk), dim3(256), 0, S.sm, S.xcur, 1.0, S.step, S.xpls, n);
            BD_TRY(S.F(S.xpls, S.fpls));
            S.nrm_to(S.fpls, 0);
            BD_TRY(S.read(1));
            if (S.ffail) { iterm = 2; return BD_OK; }
            S.nfe++;
            fpnrm = S.hsc[0];
            if (fpnrm <= (1.0 - 1.0e-4 * (1.0 - eta)) * fcnrm) break;
            ibt++;
            if (ibt > ibtmax) { iterm = 6; return BD_OK; }
            double theta = -(oftjs * redfac) / (fpnrm * fpnrm - fcnrm * fcnrm - 2.0 * oftjs * redfac);
            if (theta < THMIN) theta = THMIN;
            if (theta > THMAX) theta = THMAX;
            hipLaunchKernelGGL(se_scale_copy_kernel, dim3(S.nbk), dim3(256), 0, S.sm, theta, S.step, S.step, n);
            eta = 1.0 - theta * (1.0 - eta);
            redfac = theta * redfac;
        }
        S.nbt += ibt;
        // forcing term, choice 1 with safeguards (nitdrv.f)
        {
            double etamin = std::pow(eta, CHOICE1_EXP);
            const double t = 1.0 - redfac;
            const double flmnrm =
                std::sqrt((t * fcnrm) * (t * fcnrm) + 2.0 * redfac * t * oftlm + (redfac * rsnrm) * (redfac * rsnrm));
            eta = std::fabs(fpnrm - flmnrm) / fcnrm;
            if (etamin <= ETA_CUTOFF) etamin = 0.0;
            if (eta < etamin) eta = etamin;
            if (eta > ETA_MAX) eta = ETA_MAX;
            if (eta * fpnrm <= 2.0 * ftol) eta = (0.8 * ftol) / fpnrm;
        }
        std::swap(S.xcur, S.xpls);
        std::swap(S.fcur, S.fpls);
        fcnrm = fpnrm;
        S.nrm_to(S.step, 0);
        S.nrm_to(S.xcur, 1);
        BD_TRY(S.read(2));
        stpnrm = S.hsc[0];
        xnorm = S.hsc[1];
        S.nni++;
    }
}

int se_alloc(Ctx &c)
{
    const int kdmax = c.opt.se_kdmax > 0 ? c.opt.se_kdmax : 20;
    const size_t cols = (size_t)kdmax + 1 + 9;
    BD_HIP(hipMalloc(&c.se_buf, cols * c.npad * sizeof(double)));
    BD_HIP(hipMemset(c.se_buf, 0, cols * c.npad * sizeof(double)));
    BD_HIP(hipMalloc(&c.se_sc, 64 * sizeof(double)));
    BD_HIP(hipHostMalloc(&c.se_hsc, 64 * sizeof(double)));
    return BD_OK;
}

void se_free(Ctx &c)
{
    if (c.se_buf) (void)hipFree(c.se_buf);
    if (c.se_sc) (void)hipFree(c.se_sc);
    if (c.se_hsc) (void)hipHostFree(c.se_hsc);
    c.se_buf = nullptr;
    c.se_sc = nullptr;
    c.se_hsc = nullptr;
}

static const char *iterm_message(int iterm)
{
    switch (iterm) {
    case 1: return "max nonlinear iterations reached";
    case 2: return "failed to evaluate nonlinear function";
    case 3: return "failed to evaluate J*v";
    case 5: return "insufficient initial norm reduction";
    case 6: return "failed to reach an acceptable step by backtracking";
    default: return "nitsol failed";
    }
}

// integrate_se (bd_solver_m.f90:242-308) after calc_diffusion: the new
// coordinates in c.crd (wrapped when periodic); w is this step's Brownian
// column (null: none).
int se_step(Ctx &c, const double *w)
{
    const int kdmax = c.opt.se_kdmax > 0 ? c.opt.se_kdmax : 20;
    const size_t np = c.npad;
    double *b = c.se_buf + (size_t)(kdmax + 1) * np;
    Se S{c, b + 8 * np, w, c.n, (unsigned)((c.n + 255) / 256), c.stream, c.se_buf,
         b, b + np, b + 2 * np, b + 3 * np, b + 4 * np, b + 5 * np, b + 6 * np, c.se_sc, c.se_hsc};
    // crd0 = pvcrd; sol = pvcrd
    BD_HIP(hipMemcpyAsync(b + 8 * np, c.crd, (size_t)c.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    BD_HIP(hipMemcpyAsync(S.xcur, c.crd, (size_t)c.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    int iterm = 0;
    BD_TRY(se_nitsol(S, iterm));
    c.se_last[0] = S.nfe;
    c.se_last[1] = S.njve;
    c.se_last[2] = 0;
    c.se_last[3] = S.nli;
    c.se_last[4] = S.nni;
    c.se_last[5] = S.nbt;
    for (int k = 0; k < 6; k++) c.se_total[k] += c.se_last[k];
    if (iterm != 0) {
        set_error(std::string("integrate_se: ") + iterm_message(iterm) + " (nitsol iterm " + std::to_string(iterm) +
                  ")");
        return BD_ERR_SOLVE;
    }
    hipLaunchKernelGGL(se_finish_kernel, dim3((c.nb + 255) / 256), dim3(256), 0, c.stream, S.xcur, c.crd, c.nb,
                       c.imcon, c.box[0], c.box[1], c.box[2], c.st);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_se_info(bd_ctx *h, int *last6, long long *total6)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    for (int k = 0; k < 6; k++) {
        if (last6) last6[k] = h->c.se_last[k];
        if (total6) total6[k] = h->c.se_total[k];
    }
    return BD_OK;
}
