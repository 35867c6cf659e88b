// bd_adjoint.hip -- hand-written adjoint of the EM trajectory with the
// Brownian displacements held fixed.
//
// No reference counterpart: BROWNPAK has no gradients.  The forward map is
// the bds_run loop (bd_solver_m.f90:167-238) with the diffusion term
// replaced by a recorded displacement Delta_k:
//     r_{k+1} = P (r_k + dt D(r_b) F(r_k; theta) + Delta_k),
// P = centre-of-mass removal (simbox_m.f90:182-202, unbounded chain),
// b = the step at which the mobility was last assembled (every s steps).
// For a loss L(r_K) the adjoint sweep is
//     mu_k     = P lambda_{k+1}
//     lambda_k = mu_k + dt J_F(r_k) D(r_b) mu_k
//                + [k == b] sum_{k' in block b} dt grad_{r_b}(mu_k'^T D(r_b) F_k')
//     dL/dtheta += dt (dF(r_k)/dtheta)^T D(r_b) mu_k
// with J_F the (symmetric) force Jacobian.  States are recomputed from
// checkpoints taken every ckpt steps (a multiple of s), so memory is
// O(ckpt * n) instead of O(K * n).
#include <cmath>
#include <vector>

#include "bd_internal.h"
#include "bd_lanczos_dev.h"

namespace bd {

// grad_r [a^T M(r) b] for the RPY block M(r) = alpha(r) I + beta(r) r r^T
// (bead radius 1; overlap branch for r < 2, bd_solver_m.f90:484-504).
__device__ inline void rpy_phi_grad(double x, double y, double z, const double *a, const double *b, double &gx,
                                    double &gy, double &gz)
{
    double r2 = x * x + y * y + z * z, r = sqrt(r2);
    double al_p, be, be_p;
    if (r >= 2.0) {
        double i1 = 1.0 / r, i2 = i1 * i1, i3 = i2 * i1, i4 = i2 * i2, i5 = i4 * i1, i6 = i3 * i3;
        al_p = -0.75 * i2 - 1.5 * i4;
        be = 0.75 * i3 - 1.5 * i5;
        be_p = -2.25 * i4 + 7.5 * i6;
    } else {
        al_p = -9.0 / 32.0;
        be = 3.0 / (32.0 * r);
        be_p = -3.0 / (32.0 * r2);
    }
    double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    double ar = a[0] * x + a[1] * y + a[2] * z;
    double br = b[0] * x + b[1] * y + b[2] * z;
    double radial = (al_p * ab + be_p * ar * br) / r;
    gx = radial * x + be * (br * a[0] + ar * b[0]);
    gy = radial * y + be * (br * a[1] + ar * b[1]);
    gz = radial * z + be * (br * a[2] + ar * b[2]);
}

// out_j += scale * sum_{i != j} [grad phi(r_j - r_i; w_i, v_j) + grad phi(r_j - r_i; w_j, v_i)]
// = scale * grad_{r_j} (w^T D(r) v).
__global__ __launch_bounds__(256) void rpy_vjp_kernel(const double *__restrict__ crd, const double *__restrict__ w,
                                                      const double *__restrict__ v, int nb, double scale,
                                                      double *__restrict__ out)
{
    __shared__ double sr[256][3], sw[256][3], sv[256][3];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    double rj[3] = {0, 0, 0}, wj[3] = {0, 0, 0}, vj[3] = {0, 0, 0};
    if (j < nb)
        for (int d = 0; d < 3; d++) { rj[d] = crd[3 * j + d]; wj[d] = w[3 * j + d]; vj[d] = v[3 * j + d]; }
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int i0 = 0; i0 < nb; i0 += 256) {
        __syncthreads();
        int ii = i0 + threadIdx.x;
        for (int d = 0; d < 3; d++) {
            sr[threadIdx.x][d] = ii < nb ? crd[3 * ii + d] : 0.0;
            sw[threadIdx.x][d] = ii < nb ? w[3 * ii + d] : 0.0;
            sv[threadIdx.x][d] = ii < nb ? v[3 * ii + d] : 0.0;
        }
        __syncthreads();
        if (j < nb) {
            int lim = min(256, nb - i0);
            for (int k = 0; k < lim; k++) {
                if (i0 + k == j) continue;
                double x = rj[0] - sr[k][0], y = rj[1] - sr[k][1], z = rj[2] - sr[k][2];
                double ax, ay, az, bx, by, bz;
                rpy_phi_grad(x, y, z, sw[k], vj, ax, ay, az);
                rpy_phi_grad(x, y, z, wj, sv[k], bx, by, bz);
                gx += ax + bx;
                gy += ay + by;
                gz += az + bz;
            }
        }
    }
    if (j < nb) {
        out[3 * j] += scale * gx;
        out[3 * j + 1] += scale * gy;
        out[3 * j + 2] += scale * gz;
    }
}

// y = x - mean(x) per component (P, one block)
__global__ __launch_bounds__(1024) void project_kernel(const double *__restrict__ x, double *__restrict__ y, int nb)
{
    __shared__ double sh[3][16];
    __shared__ double m[3];
    double s[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) s[d] += x[3 * i + d];
    for (int d = 0; d < 3; d++) {
        double v = s[d];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[d][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++) v += sh[threadIdx.x][w];
        m[threadIdx.x] = v / nb;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) y[3 * i + d] = x[3 * i + d] - m[d];
}

// lambda = (2/N)(r - com): gradient of Rg^2 (stats_m.f90:441-446)
__global__ __launch_bounds__(1024) void rgsq_grad_kernel(const double *__restrict__ crd, int nb,
                                                         double *__restrict__ lam)
{
    __shared__ double sh[3][16];
    __shared__ double m[3];
    double s[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) s[d] += crd[3 * i + d];
    for (int d = 0; d < 3; d++) {
        double v = s[d];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[d][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++) v += sh[threadIdx.x][w];
        m[threadIdx.x] = v / nb;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) lam[3 * i + d] = (2.0 / nb) * (crd[3 * i + d] - m[d]);
}

// lam = mu + dt * h ; theta += dt * sum(tpart)
__global__ void adj_combine_kernel(const double *__restrict__ mu, const double *__restrict__ h, double dt, int n,
                                   double *__restrict__ lam, const double *__restrict__ tpart, int ntp,
                                   double *__restrict__ theta)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) lam[i] = mu[i] + dt * h[i];
    if (i == 0) {
        for (int m = 0; m < 3; m++) {
            theta[m] += dt * sum_strided(tpart + m, 3, ntp);
        }
    }
}

__global__ void add_kernel(double *__restrict__ a, double *__restrict__ b, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] += b[i];
        b[i] = 0.0;
    }
}

}  // namespace bd

using namespace bd;

namespace {

struct DevBuf {
    double *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count)
    {
        BD_HIP(hipMalloc(&p, count * sizeof(double)));
        BD_HIP(hipMemset(p, 0, count * sizeof(double)));
        return BD_OK;
    }
};

int replay_steps(Ctx &c, const double *disp_dev, int k0, int k1, double *states)
{
    // states (optional): positions before each step, (k1 - k0) x npad
    long long cons;
    for (int k = k0; k < k1; k++) {
        if (states)
            BD_HIP(hipMemcpyAsync(states + (size_t)(k - k0) * c.npad, c.crd, (size_t)c.npad * sizeof(double),
                                  hipMemcpyDeviceToDevice, c.stream));
        BD_TRY(em_step(c, nullptr, 0, &cons, disp_dev + (size_t)k * c.npad, nullptr));
    }
    return BD_OK;
}

}  // namespace

extern "C" {

int bd_set_theta(bd_ctx *h, double eps, double sigma, double kbond, int tie_rc)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    return set_theta(h->c, eps, sigma, kbond, tie_rc);
}

int bd_run_record(bd_ctx *h, int nsteps, const double *xi, long long nxi, long long *used, double *disp_out)
{
    if (!h || nsteps < 0 || !disp_out) { set_error("bad arguments"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    if (!c.have_topology) { set_error("topology not set"); return BD_ERR_STATE; }
    DevBuf rec;
    BD_TRY(rec.alloc((size_t)c.npad * std::max(nsteps, 1)));
    BD_HIP(hipMemsetAsync(&c.st->steps_done, 0, sizeof(int), c.stream));
    if (c.limud())
        BD_TRY(launch_to_origin(c, nullptr));
    long long pos = 0;
    int rc = BD_OK;
    for (int k = 0; k < nsteps && rc == BD_OK; k++) {
        long long cons = 0;
        rc = em_step(c, xi ? xi + pos : nullptr, xi ? nxi - pos : 0, &cons, nullptr, rec.p + (size_t)k * c.npad);
        pos += cons;
    }
    if (used) *used = pos;
    int rc2 = check_status(c, false);
    if (rc == BD_OK) rc = rc2;
    c.info.nts += c.st_host->steps_done;
    BD_HIP(hipMemcpy2D(disp_out, (size_t)c.n * sizeof(double), rec.p, (size_t)c.npad * sizeof(double),
                       (size_t)c.n * sizeof(double), nsteps, hipMemcpyDeviceToHost));
    return rc;
}

int bd_run_replay(bd_ctx *h, int nsteps, const double *disp)
{
    if (!h || nsteps < 0 || !disp) { set_error("bad arguments"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    if (!c.have_topology) { set_error("topology not set"); return BD_ERR_STATE; }
    DevBuf dd;
    BD_TRY(dd.alloc((size_t)c.npad * std::max(nsteps, 1)));
    BD_HIP(hipMemcpy2D(dd.p, (size_t)c.npad * sizeof(double), disp, (size_t)c.n * sizeof(double),
                       (size_t)c.n * sizeof(double), nsteps, hipMemcpyHostToDevice));
    BD_HIP(hipMemsetAsync(&c.st->steps_done, 0, sizeof(int), c.stream));
    if (c.limud())
        BD_TRY(launch_to_origin(c, nullptr));
    int rc = replay_steps(c, dd.p, 0, nsteps, nullptr);
    int rc2 = check_status(c, false);
    if (rc == BD_OK) rc = rc2;
    c.info.nts += c.st_host->steps_done;
    return rc;
}

int bd_adjoint_rgsq(bd_ctx *h, int nsteps, const double *disp, int ckpt, int tie_rc, double *loss, double *grad3)
{
    if (!h || nsteps <= 0 || !disp || !loss || !grad3) { set_error("bad arguments"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    if (!c.have_topology) { set_error("topology not set"); return BD_ERR_STATE; }
    if (c.imcon != 0) { set_error("the Rg^2 adjoint is for an isolated chain (imcon = 0)"); return BD_ERR_ARG; }
    if (c.opt.integrator != BD_EM) { set_error("the Rg^2 adjoint differentiates the EM step"); return BD_ERR_ARG; }
    if (c.vdw_general || c.nangles || c.ntethers || c.nexternals) {
        set_error("the Rg^2 adjoint differentiates bonds and LJ pairs only (no angles, tethers, external fields "
                  "or vdw styles 2-6)");
        return BD_ERR_ARG;
    }
    const int s = c.opt.lhdia ? c.s : 1;
    if (ckpt <= 0 || ckpt % s != 0) { set_error("ckpt must be a positive multiple of nts_mobsam"); return BD_ERR_ARG; }
    if (c.cntr_mobsam != 0) { set_error("the adjoint needs a run that starts a mobility block"); return BD_ERR_STATE; }
    const int np = c.npad, nb = c.nb, n = c.n;
    const double dt = c.opt.tim_stp;
    hipStream_t sm = c.stream;
    const int nseg = (nsteps + ckpt - 1) / ckpt;
    DevBuf dd, ck, states, lam, mu, dmu, hv, pend, tpart, theta;
    BD_TRY(dd.alloc((size_t)np * nsteps));
    BD_TRY(ck.alloc((size_t)np * (nseg + 1)));
    BD_TRY(states.alloc((size_t)np * ckpt));
    BD_TRY(lam.alloc(np));
    BD_TRY(mu.alloc(np));
    BD_TRY(dmu.alloc(np));
    BD_TRY(hv.alloc(np));
    BD_TRY(pend.alloc(np));
    const int ntp = std::max((nb + 255) / 256, 1);
    BD_TRY(tpart.alloc((size_t)ntp * 3));
    BD_TRY(theta.alloc(3));
    BD_HIP(hipMemcpy2D(dd.p, (size_t)np * sizeof(double), disp, (size_t)n * sizeof(double), (size_t)n * sizeof(double),
                       nsteps, hipMemcpyHostToDevice));
    BD_HIP(hipMemsetAsync(&c.st->steps_done, 0, sizeof(int), sm));
    BD_TRY(launch_to_origin(c, nullptr));
    // forward with checkpoints
    for (int g = 0; g < nseg; g++) {
        BD_HIP(hipMemcpyAsync(ck.p + (size_t)g * np, c.crd, (size_t)np * sizeof(double), hipMemcpyDeviceToDevice, sm));
        BD_TRY(replay_steps(c, dd.p, g * ckpt, std::min(nsteps, (g + 1) * ckpt), nullptr));
    }
    BD_HIP(hipMemcpyAsync(ck.p + (size_t)nseg * np, c.crd, (size_t)np * sizeof(double), hipMemcpyDeviceToDevice, sm));
    int rc = check_status(c, false);
    if (rc != BD_OK) return rc;
    BD_TRY(bd_rgsq(h, loss));
    hipLaunchKernelGGL(rgsq_grad_kernel, dim3(1), dim3(1024), 0, sm, c.crd, nb, lam.p);
    // backward sweep, segment by segment
    const int nbk = (n + 255) / 256;
    for (int g = nseg - 1; g >= 0; g--) {
        const int k0 = g * ckpt, k1 = std::min(nsteps, (g + 1) * ckpt);
        BD_HIP(hipMemcpyAsync(c.crd, ck.p + (size_t)g * np, (size_t)np * sizeof(double), hipMemcpyDeviceToDevice, sm));
        c.cntr_mobsam = 0;
        c.verlet_first = true;
        BD_TRY(replay_steps(c, dd.p, k0, k1, states.p));
        for (int k = k1 - 1; k >= k0; k--) {
            const int b = k - (k % s);
            const double *rk = states.p + (size_t)(k - k0) * np;
            const double *rb = states.p + (size_t)(b - k0) * np;
            hipLaunchKernelGGL(project_kernel, dim3(1), dim3(1024), 0, sm, lam.p, mu.p, nb);
            // forces and the force list at r_k
            BD_HIP(hipMemcpyAsync(c.crd, rk, (size_t)np * sizeof(double), hipMemcpyDeviceToDevice, sm));
            c.verlet_first = true;
            BD_TRY(compute_forces(c, false, false));
            const double *dm = mu.p;
            if (c.opt.lhdia) {
                BD_TRY(rpy_assemble_at(c, rb));
                BD_TRY(mob_product(c, mu.p, dmu.p, 1, np, np));
                dm = dmu.p;
            }
            BD_TRY(adj_force(c, dm, hv.p, tpart.p, tie_rc));
            if (c.opt.lhdia)
                hipLaunchKernelGGL(rpy_vjp_kernel, dim3((nb + 255) / 256), dim3(256), 0, sm, rb, mu.p, c.frc, nb, dt,
                                   pend.p);
            hipLaunchKernelGGL(adj_combine_kernel, dim3(nbk), dim3(256), 0, sm, mu.p, hv.p, dt, n, lam.p, tpart.p, ntp,
                               theta.p);
            if (c.opt.lhdia && k == b) hipLaunchKernelGGL(add_kernel, dim3(nbk), dim3(256), 0, sm, lam.p, pend.p, n);
            BD_HIP(hipGetLastError());
        }
    }
    // leave the final state in place, as a forward run would
    BD_HIP(hipMemcpyAsync(c.crd, ck.p + (size_t)nseg * np, (size_t)np * sizeof(double), hipMemcpyDeviceToDevice, sm));
    c.cntr_mobsam = 0;
    c.verlet_first = true;
    BD_HIP(hipMemcpyAsync(grad3, theta.p, 3 * sizeof(double), hipMemcpyDeviceToHost, sm));
    BD_HIP(hipStreamSynchronize(sm));
    return check_status(c, false);
}

}  // extern "C"
