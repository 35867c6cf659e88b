// bd_api.hip -- the extern "C" boundary of libbdstep and the EM step loop.
//
// The step loop restates bds_run / integrate_em / calc_diffusion /
// calc_drift (bd_solver_m.f90:128-238, 355-454) on device-resident data:
// coordinates never leave HBM between steps.
#include <cmath>
#include <cstring>
#include <mutex>

#include "bd_internal.h"

namespace bd {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// ------------------------------------------------------------ RNG
// Philox-4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11) counter-based
// generator; Box-Muller turns two 53-bit uniforms into two N(0,1) values.
__device__ inline void philox_round(uint32_t (&c)[4], uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

__device__ inline void philox4x32_10(uint32_t (&c)[4], uint32_t k0, uint32_t k1)
{
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    for (int r = 0; r < 10; r++) {
        philox_round(c, k0, k1);
        k0 += W0;
        k1 += W1;
    }
}

// out[col * ld + row] for row < rows, col < cols, stream position
// offset + (col * rows + row) (column-major order, as brn_calc_dw fills).
__global__ void gauss_kernel(double *out, int ld, int rows, int cols, uint64_t seed, uint64_t offset, double scale)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)rows * cols;
    long pair0 = 2 * g;
    if (pair0 >= total) return;
    uint64_t ctr = (offset + (uint64_t)pair0) >> 1;  // one Philox block per pair
    uint32_t c[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), 0x243F6A88u, 0x85A308D3u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    uint64_t a = ((uint64_t)c[0] << 32) | c[1], b = ((uint64_t)c[2] << 32) | c[3];
    double u1 = ((a >> 11) + 1) * 0x1.0p-53;  // (0, 1]
    double u2 = (b >> 11) * 0x1.0p-53;        // [0, 1)
    double r = sqrt(-2.0 * log(u1));
    double s, cth;
    sincospi(2.0 * u2, &s, &cth);
    double z[2] = {r * cth, r * s};
    for (int q = 0; q < 2; q++) {
        long e = pair0 + q;
        if (e < total) {
            int col = (int)(e / rows), row = (int)(e % rows);
            out[(long)col * ld + row] = scale * z[q];
        }
    }
}

// scale multiplies every value (the step's sqrt(2 dt) when the draws go
// straight into the Brownian displacement; 1 leaves them as drawn).
int launch_gauss(Ctx &c, double *out, long long count, uint64_t seed, uint64_t offset, int ld, int rows, int cols,
                 double scale)
{
    (void)count;
    long pairs = ((long)rows * cols + 1) / 2;
    if (pairs == 0) return BD_OK;
    hipLaunchKernelGGL(gauss_kernel, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0, c.stream, out, ld, rows,
                       cols, seed, offset, scale);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// ------------------------------------------------------------ step kernels
__global__ void scale_kernel(double *x, long n, double a, const DevStatus *st)
{
    if (st->ierr) return;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = a * x[i];
}

// drift = dt * (flow(r) + v), v = D F (HI) or F (free draining)
__global__ void drift_kernel(const double *__restrict__ v, const double *__restrict__ crd, int nb, int flow_style,
                             double rate, double dt, double *__restrict__ drift, const DevStatus *st)
{
    if (st->ierr) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    double f[3] = {0.0, 0.0, 0.0};
    switch (flow_style) {  // bd_solver_m.f90:420-441
    case 1: f[0] = rate * crd[3 * i + 1]; break;
    case 2: f[0] = rate * crd[3 * i]; f[1] = -rate * crd[3 * i + 1]; break;
    case 3:
        f[0] = rate * crd[3 * i];
        f[1] = -0.5 * rate * crd[3 * i + 1];
        f[2] = -0.5 * rate * crd[3 * i + 2];
        break;
    default: break;
    }
    for (int d = 0; d < 3; d++) drift[3 * i + d] = dt * (v[3 * i + d] + f[d]);
}

int launch_drift(Ctx &c, const double *v)
{
    hipLaunchKernelGGL(drift_kernel, dim3((c.nb + 255) / 256), dim3(256), 0, c.stream, v, c.crd, c.nb,
                       c.opt.flow_style, c.opt.flow_rate, c.opt.tim_stp, c.drift, c.st);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// r <- r + drift + diff (diff may be null: zero Brownian term); PBC wrap.
__global__ void update_kernel(double *__restrict__ crd, const double *__restrict__ drift, const double *__restrict__ diff,
                              int nb, int imcon, double Lx, double Ly, double Lz, DevStatus *st)
{
    if (st->ierr) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    if (i == 0) st->steps_done += 1;
    double L[3] = {Lx, Ly, Lz};
    for (int d = 0; d < 3; d++) {
        double r = crd[3 * i + d] + drift[3 * i + d];
        if (diff) r = r + diff[3 * i + d];
        if (imcon) r = r - L[d] * floor(r / L[d]);  // smbx_wrap_all, simbox_m.f90:128-153
        crd[3 * i + d] = r;
    }
}

// drift_kernel and update_kernel in one launch, bead by bead with the same
// operations (an EM step without the isolated-molecule tail): the bead's
// flow term reads its start-of-step coordinates before they are replaced.
__global__ void drift_update_kernel(const double *__restrict__ v, double *__restrict__ crd, int nb, int flow_style,
                                    double rate, double dt, double *__restrict__ drift,
                                    const double *__restrict__ diff, int imcon, double Lx, double Ly, double Lz,
                                    DevStatus *st)
{
    if (st->ierr) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    if (i == 0) st->steps_done += 1;
    double f[3] = {0.0, 0.0, 0.0};
    switch (flow_style) {  // bd_solver_m.f90:420-441
    case 1: f[0] = rate * crd[3 * i + 1]; break;
    case 2: f[0] = rate * crd[3 * i]; f[1] = -rate * crd[3 * i + 1]; break;
    case 3:
        f[0] = rate * crd[3 * i];
        f[1] = -0.5 * rate * crd[3 * i + 1];
        f[2] = -0.5 * rate * crd[3 * i + 2];
        break;
    default: break;
    }
    const double L[3] = {Lx, Ly, Lz};
    for (int d = 0; d < 3; d++) {
        const double dr = dt * (v[3 * i + d] + f[d]);
        drift[3 * i + d] = dr;
        double r = crd[3 * i + d] + dr;
        if (diff) r = r + diff[3 * i + d];
        if (imcon) r = r - L[d] * floor(r / L[d]);  // smbx_wrap_all, simbox_m.f90:128-153
        crd[3 * i + d] = r;
    }
}

// Predictor of the predictor-corrector step: r0 <- r, d0 <- d and
// r <- (r + d) + w, wrapped into the box when periodic (as after a step, so
// that the cell-list pair table sees coordinates in [0, L)).
__global__ void pc_predict_kernel(double *__restrict__ crd, const double *__restrict__ drift,
                                  const double *__restrict__ w, double *__restrict__ r0, double *__restrict__ d0, int n,
                                  int imcon, double Lx, double Ly, double Lz, const DevStatus *st)
{
    if (st->ierr) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double r = crd[i], d = drift[i];
    r0[i] = r;
    d0[i] = d;
    double p = r + d;
    if (w) p = p + w[i];
    if (imcon) {
        const double L = (i % 3 == 0) ? Lx : ((i % 3 == 1) ? Ly : Lz);
        p = p - L * floor(p / L);   // smbx_wrap_all, simbox_m.f90:128-153
    }
    crd[i] = p;
}

// Corrector: r <- r0 and drift <- (d0 + d(r*)) / 2; the EM update then adds
// the same w.
__global__ void pc_correct_kernel(double *__restrict__ crd, double *__restrict__ drift, const double *__restrict__ r0,
                                  const double *__restrict__ d0, int n, const DevStatus *st)
{
    if (st->ierr) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    crd[i] = r0[i];
    drift[i] = 0.5 * (d0[i] + drift[i]);
}

// smbx_to_origin (simbox_m.f90:182-202) in one block; com_acc += com.
__global__ __launch_bounds__(1024) void to_origin_kernel(double *crd, int nb, double *com_acc, const DevStatus *st)
{
    if (st->ierr) return;
    __shared__ double sh[3][16];
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) s[d] += crd[3 * i + d];
    for (int d = 0; d < 3; d++) {
        double v = s[d];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[d][threadIdx.x / 64] = v;
    }
    __syncthreads();
    __shared__ double com[3];
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++) v += sh[threadIdx.x][w];
        com[threadIdx.x] = v / nb;
        if (com_acc) com_acc[threadIdx.x] += v / nb;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) crd[3 * i + d] = crd[3 * i + d] - com[d];
}

// The tail of an EM step of an isolated molecule in one launch: drift
// (drift_kernel), r <- r + drift + w (update_kernel) and the centre of mass
// to the origin (to_origin_kernel), with the same arithmetic and the same
// summation order as the three kernels, so the coordinates are the same to
// the bit.  One block of 1024 threads.
__global__ __launch_bounds__(1024) void em_tail_kernel(double *__restrict__ crd, const double *__restrict__ v,
                                                       const double *__restrict__ diff, int nb, int flow_style,
                                                       double rate, double dt, double *__restrict__ drift,
                                                       double *com_acc, DevStatus *st)
{
    __shared__ double sh[3][16];
    __shared__ double com[3];
    double s[3] = {0.0, 0.0, 0.0};
    constexpr int RPT = 2;   // beads per thread kept in registers (the same order of sums)
    if (nb <= RPT * (int)blockDim.x) {
        // the beads' loads go out together with the status word's
        const int ierr = st->ierr;
        double c0[RPT][3], vv[RPT][3], dd[RPT][3];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int i = min((int)threadIdx.x + q * (int)blockDim.x, nb - 1);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                c0[q][d] = crd[3 * i + d];
                vv[q][d] = v[3 * i + d];
                dd[q][d] = diff ? diff[3 * i + d] : 0.0;
            }
        }
        if (ierr) return;
        if (threadIdx.x == 0) st->steps_done += 1;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int i = threadIdx.x + q * blockDim.x;
            if (i >= nb) continue;
            double f[3] = {0.0, 0.0, 0.0};
            switch (flow_style) {  // bd_solver_m.f90:420-441
            case 1: f[0] = rate * c0[q][1]; break;
            case 2: f[0] = rate * c0[q][0]; f[1] = -rate * c0[q][1]; break;
            case 3:
                f[0] = rate * c0[q][0];
                f[1] = -0.5 * rate * c0[q][1];
                f[2] = -0.5 * rate * c0[q][2];
                break;
            default: break;
            }
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const double dr = dt * (vv[q][d] + f[d]);
                drift[3 * i + d] = dr;
                // separately rounded adds, as update_kernel's (which reads dr from memory)
                double r = __dadd_rn(c0[q][d], dr);
                if (diff) r = __dadd_rn(r, dd[q][d]);
                c0[q][d] = r;
                s[d] += r;
            }
        }
        for (int d = 0; d < 3; d++) {
            double x = s[d];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            if (threadIdx.x % 64 == 0) sh[d][threadIdx.x / 64] = x;
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            double x = 0.0;
            for (int w = 0; w < (int)(blockDim.x / 64); w++) x += sh[threadIdx.x][w];
            com[threadIdx.x] = x / nb;
            if (com_acc) com_acc[threadIdx.x] += x / nb;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int i = threadIdx.x + q * blockDim.x;
            if (i < nb)
                for (int d = 0; d < 3; d++) crd[3 * i + d] = c0[q][d] - com[d];
        }
        return;
    }
    if (st->ierr) return;
    if (threadIdx.x == 0) st->steps_done += 1;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        double f[3] = {0.0, 0.0, 0.0};
        switch (flow_style) {  // bd_solver_m.f90:420-441
        case 1: f[0] = rate * crd[3 * i + 1]; break;
        case 2: f[0] = rate * crd[3 * i]; f[1] = -rate * crd[3 * i + 1]; break;
        case 3:
            f[0] = rate * crd[3 * i];
            f[1] = -0.5 * rate * crd[3 * i + 1];
            f[2] = -0.5 * rate * crd[3 * i + 2];
            break;
        default: break;
        }
        for (int d = 0; d < 3; d++) {
            const double dr = dt * (v[3 * i + d] + f[d]);
            drift[3 * i + d] = dr;
            double r = __dadd_rn(crd[3 * i + d], dr);
            if (diff) r = __dadd_rn(r, diff[3 * i + d]);
            crd[3 * i + d] = r;
            s[d] += r;
        }
    }
    for (int d = 0; d < 3; d++) {
        double x = s[d];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (threadIdx.x % 64 == 0) sh[d][threadIdx.x / 64] = x;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double x = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++) x += sh[threadIdx.x][w];
        com[threadIdx.x] = x / nb;
        if (com_acc) com_acc[threadIdx.x] += x / nb;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) crd[3 * i + d] = crd[3 * i + d] - com[d];
}

__global__ __launch_bounds__(1024) void rgsq_kernel(const double *crd, int nb, double *out)
{
    __shared__ double sh[4][16];
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) s[d] += crd[3 * i + d];
    for (int d = 0; d < 3; d++) {
        double v = s[d];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[d][threadIdx.x / 64] = v;
    }
    __syncthreads();
    __shared__ double com[3];
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++) v += sh[threadIdx.x][w];
        com[threadIdx.x] = v / nb;
    }
    __syncthreads();
    double a = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        double q = 0.0;
        for (int d = 0; d < 3; d++) {
            double t = crd[3 * i + d] - com[d];
            q += t * t;
        }
        a += q;
    }
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    __syncthreads();
    if (threadIdx.x % 64 == 0) sh[3][threadIdx.x / 64] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++) v += sh[3][w];
        *out = v / nb;
    }
}

int launch_to_origin(Ctx &c, double *com_acc)
{
    hipLaunchKernelGGL(to_origin_kernel, dim3(1), dim3(1024), 0, c.stream, c.crd, c.nb, com_acc, c.st);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// ------------------------------------------------------------ helpers
static int upload_xi(Ctx &c, const double *xi_host, int s)
{
    // host (n, s) column-major -> device (npad, s)
    BD_HIP(hipMemcpy2DAsync(c.xi_dev, (size_t)c.npad * sizeof(double), xi_host, (size_t)c.n * sizeof(double),
                            (size_t)c.n * sizeof(double), s, hipMemcpyHostToDevice, c.stream));
    return BD_OK;
}

static int download_block(Ctx &c, const double *dev, double *host, int s)
{
    BD_HIP(hipMemcpy2DAsync(host, (size_t)c.n * sizeof(double), dev, (size_t)c.npad * sizeof(double),
                            (size_t)c.n * sizeof(double), s, hipMemcpyDeviceToHost, c.stream));
    BD_HIP(hipStreamSynchronize(c.stream));
    return BD_OK;
}

int check_status(Ctx &c, bool clear)
{
    BD_HIP(hipMemcpyAsync(c.st_host, c.st, sizeof(DevStatus), hipMemcpyDeviceToHost, c.stream));
    BD_HIP(hipStreamSynchronize(c.stream));
    int e = c.st_host->ierr;
    if (e && clear) {
        int z = 0;
        BD_HIP(hipMemcpy(&c.st->ierr, &z, sizeof(int), hipMemcpyHostToDevice));
    }
    if (e == 1) { set_error("bond length too large (overstretched bond)"); return BD_ERR_BOND; }
    if (e == 2) { set_error("factorisation failed"); return BD_ERR_FACTOR; }
    if (e == 4) { set_error("negative coordinate in cell-list mode (cell_list_m.f90:193-196)"); return BD_ERR_CELL; }
    if (e == 5) { set_error("rigid tether violated (ia_tether_m.f90:110-128)"); return BD_ERR_BOND; }
    if (e == 6) { set_error("external wall crossed (ia_external_m.f90:93-101)"); return BD_ERR_BOND; }
    return BD_OK;
}

// B.xi for the block in c.xi_dev -> c.diff (scaled later by the caller).
static int brownian_dev(Ctx &c, int method, int s, int *lconv, int *iters, bool sync = true)
{
    if (method == BD_LANC) return brownian_lanczos(c, c.xi_dev, c.diff, s, lconv, iters);
    *lconv = 1;
    *iters = 0;
    return brownian_cholesky(c, c.xi_dev, c.diff, s, method, sync);
}

// One EM step (integrate_em + the bds_run loop tail).  xi_host: next block of
// the injected stream (or null for the device generator).  Returns the number
// of xi values consumed via *consumed.
// disp_dev (npad, may be null): replay a recorded Brownian displacement
// instead of computing B.xi; rec_dev (npad, may be null): record the
// displacement this step applied (zero on steps without noise).
int em_step(Ctx &c, const double *xi_host, long long avail, long long *consumed, const double *disp_dev,
            double *rec_dev)
{
    const int n = c.n, nb = c.nb;
    const bool hi = c.opt.lhdia != 0;
    const int s = hi ? c.s : 1;
    const double sqrt_two_dt = std::sqrt(2.0 * c.opt.tim_stp);
    hipStream_t sm = c.stream;
    *consumed = 0;
    const int nbk = (nb + 255) / 256;
    bool noise = true;
    // ---- calc_diffusion (bd_solver_m.f90:355-403)
    if (!hi) {
        if (disp_dev) {
        } else if (xi_host) {
            if (avail < n) { set_error("injected xi stream exhausted"); return BD_ERR_ARG; }
            BD_TRY(upload_xi(c, xi_host, 1));
            *consumed = n;
            BD_HIP(hipMemcpyAsync(c.diff, c.xi_dev, (size_t)c.npad * sizeof(double), hipMemcpyDeviceToDevice, sm));
            hipLaunchKernelGGL(scale_kernel, dim3((n + 255) / 256), dim3(256), 0, sm, c.diff, (long)n, sqrt_two_dt,
                               c.st);
        } else {
            // sqrt(2 dt) xi drawn straight into the displacement (the same
            // product as the scaling of an injected stream)
            BD_TRY(launch_gauss(c, c.diff, n, c.rng_seed, c.rng_offset, c.npad, n, 1, sqrt_two_dt));
            c.rng_offset += (uint64_t)n + (n & 1);
        }
    } else {
        if (c.cntr_mobsam == 0) {
            BD_TRY(rpy_assemble(c));
            if (!disp_dev) {
                long long need = (long long)n * s;
                if (xi_host) {
                    if (avail < need) { set_error("injected xi stream exhausted"); return BD_ERR_ARG; }
                    BD_TRY(upload_xi(c, xi_host, s));
                    *consumed = need;
                } else {
                    BD_TRY(launch_gauss(c, c.xi_dev, need, c.rng_seed, c.rng_offset, c.npad, n, s));
                    c.rng_offset += (uint64_t)need + (need & 1);
                }
                int lconv = 1, iters = 0;
                int method = c.opt.mob_fctr;
                // a Cholesky failure stops the step's later kernels (ierr) and
                // bd_run reports it; no host round trip per step
                int rc = brownian_dev(c, method, s, &lconv, &iters, false);
                if (rc != BD_OK) return rc;
                if (method == BD_LANC) {
                    c.info.last_lanczos_iters = iters;
                    c.info.total_lanczos_iters += iters;
                    if (!lconv) c.info.n_not_converged++;
                }
                hipLaunchKernelGGL(scale_kernel, dim3((unsigned)(((long)c.npad * s + 255) / 256)), dim3(256), 0, sm,
                                   c.diff, (long)c.npad * s, sqrt_two_dt, c.st);
            }
            c.info.n_mob_updates++;
        } else if (!c.opt.keep_reuse_noise) {
            noise = false;  // the reference zeroed the block (bd_solver_m.f90:364)
        }
        c.cntr_mobsam++;
    }
    // ---- integrate_se (bd_solver_m.f90:242-308): the new coordinates solve
    // x - drift(x) - crd0 - w = 0 with this step's D and w
    if (c.opt.integrator == BD_SE) {
        const double *w = nullptr;
        if (disp_dev) w = disp_dev;
        else if (noise) w = c.diff + (size_t)(hi ? c.cntr_mobsam - 1 : 0) * c.npad;
        if (rec_dev) {
            if (w) BD_HIP(hipMemcpyAsync(rec_dev, w, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, sm));
            else BD_HIP(hipMemsetAsync(rec_dev, 0, (size_t)n * sizeof(double), sm));
        }
        BD_TRY(se_step(c, w));
        // the bds_run tail (bd_solver_m.f90:176-186); se_step wrapped a periodic box
        if (c.limud())
            hipLaunchKernelGGL(to_origin_kernel, dim3(1), dim3(1024), 0, sm, c.crd, nb, c.scal + 16, c.st);
        if (c.stats_on) BD_TRY(stats_step(c));
        if (hi && c.cntr_mobsam == s) c.cntr_mobsam = 0;
        BD_HIP(hipGetLastError());
        return BD_OK;
    }
    // ---- calc_drift (bd_solver_m.f90:407-454): drift = dt (U(r) + D F(r))
    // (an EM step of an isolated molecule leaves the drift to its tail kernel)
    const bool fused_tail = c.opt.integrator != BD_PC && c.limud();
    // otherwise an EM step (no isolated-molecule tail then) updates in the
    // drift's launch
    const bool fused_update = c.opt.integrator == BD_EM && !fused_tail;
    const double *vdrift = nullptr;
    auto drift_at_crd = [&](bool launch, bool energies) -> int {
        BD_TRY(compute_forces(c, energies, false));
        const double *v = c.frc;
        if (hi) {
            BD_TRY(mob_product(c, c.frc, c.W, 1, c.npad, c.npad));
            v = c.W;
        }
        vdrift = v;
        if (launch)
            hipLaunchKernelGGL(drift_kernel, dim3(nbk), dim3(256), 0, sm, v, c.crd, nb, c.opt.flow_style,
                               c.opt.flow_rate, c.opt.tim_stp, c.drift, c.st);
        return BD_OK;
    };
    // the statistics take the energies of this evaluation (calc_drift at the
    // start-of-step coordinates, bd_solver_m.f90:407-454)
    BD_TRY(drift_at_crd(!fused_tail && !fused_update, c.stats_on));
    // Brownian displacement w of this step
    const double *dcol = nullptr;
    if (disp_dev) dcol = disp_dev;
    else if (noise) dcol = c.diff + (size_t)(hi ? c.cntr_mobsam - 1 : 0) * c.npad;
    // ---- predictor-corrector: drift <- (d(r) + d(r + d(r) + w)) / 2
    if (c.opt.integrator == BD_PC) {
        const int nbn = (n + 255) / 256;
        hipLaunchKernelGGL(pc_predict_kernel, dim3(nbn), dim3(256), 0, sm, c.crd, c.drift, dcol, c.pc_r0, c.pc_d0,
                           n, c.imcon, c.box[0], c.box[1], c.box[2], c.st);
        BD_TRY(drift_at_crd(true, false));
        hipLaunchKernelGGL(pc_correct_kernel, dim3(nbn), dim3(256), 0, sm, c.crd, c.drift, c.pc_r0, c.pc_d0, n,
                           c.st);
    }
    // ---- update (integrate_em, bd_solver_m.f90:231-236)
    if (rec_dev) {
        if (dcol) BD_HIP(hipMemcpyAsync(rec_dev, dcol, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, sm));
        else BD_HIP(hipMemsetAsync(rec_dev, 0, (size_t)n * sizeof(double), sm));
    }
    if (fused_tail) {
        // drift, update and the bds_run tail (bd_solver_m.f90:176-186) in one launch
        hipLaunchKernelGGL(em_tail_kernel, dim3(1), dim3(1024), 0, sm, c.crd, vdrift, dcol, nb, c.opt.flow_style,
                           c.opt.flow_rate, c.opt.tim_stp, c.drift, c.scal + 16, c.st);
    } else if (fused_update) {
        hipLaunchKernelGGL(drift_update_kernel, dim3(nbk), dim3(256), 0, sm, vdrift, c.crd, nb, c.opt.flow_style,
                           c.opt.flow_rate, c.opt.tim_stp, c.drift, dcol, c.imcon, c.box[0], c.box[1], c.box[2],
                           c.st);
    } else {
        hipLaunchKernelGGL(update_kernel, dim3(nbk), dim3(256), 0, sm, c.crd, c.drift, dcol, nb, c.imcon, c.box[0],
                           c.box[1], c.box[2], c.st);
        // ---- bds_run tail (bd_solver_m.f90:176-186)
        if (c.limud())
            hipLaunchKernelGGL(to_origin_kernel, dim3(1), dim3(1024), 0, sm, c.crd, nb, c.scal + 16, c.st);
    }
    // stats_accumulate (bd_solver_m.f90:189) on the end-of-step coordinates
    if (c.stats_on) BD_TRY(stats_step(c));
    if (hi && c.cntr_mobsam == s) c.cntr_mobsam = 0;
    BD_HIP(hipGetLastError());
    return BD_OK;
}

}  // namespace bd

using namespace bd;

extern "C" {

const char *bd_last_error(void) { return g_err.c_str(); }

bd_ctx *bd_ctx_create(int nbeads, const bd_opts *opts)
{
    if (nbeads <= 0 || !opts) { set_error("nbeads must be positive and opts non-null"); return nullptr; }
    if (opts->lhdia && (opts->nts_mobsam < 1 || opts->nts_mobsam > MAX_S)) {
        set_error("nts_mobsam must be in 1..8");
        return nullptr;
    }
    if (opts->mob_fctr < BD_CHOL || opts->mob_fctr > BD_CHOL_LT) { set_error("unknown mob_fctr"); return nullptr; }
    if (opts->precision != BD_FP64 && opts->precision != BD_FP32) { set_error("unknown precision"); return nullptr; }
    if (opts->integrator != BD_EM && opts->integrator != BD_PC && opts->integrator != BD_SE) {
        set_error("unknown integrator");
        return nullptr;
    }
    if (opts->integrator == BD_SE && (opts->se_nlmxitr < 0 || opts->se_kdmax < 0 || opts->se_kdmax > 60)) {
        set_error("se_nlmxitr must be >= 0 and se_kdmax in 0..60");
        return nullptr;
    }
    if (opts->mob_storage < BD_MOB_AUTO || opts->mob_storage > BD_MOB_MATFREE) {
        set_error("unknown mob_storage");
        return nullptr;
    }
    bd_ctx *h = new bd_ctx();
    Ctx &c = h->c;
    c.nb = nbeads;
    c.n = 3 * nbeads;
    c.opt = *opts;
    c.s = opts->lhdia ? opts->nts_mobsam : 1;
    int lmx = opts->lanc_mxitr;
    c.dimk_max = (lmx > 2 ? lmx : 2) + 1;  // brown_m.f90:96
    auto fail = [&](int) -> bd_ctx * {
        bd_ctx_destroy(h);
        return nullptr;
    };
    if (hipGetDevice(&c.device) != hipSuccess) { set_error("no HIP device"); return fail(0); }
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(0);
    }
    c.ntile = (c.n + TILE - 1) / TILE;
    c.npad = c.ntile * TILE;
    auto dalloc = [&](double **p, size_t count) -> bool {
        if (hipMalloc(p, count * sizeof(double)) != hipSuccess) return false;
        return hipMemset(*p, 0, count * sizeof(double)) == hipSuccess;
    };
    const int s = c.s;
    if (!dalloc(&c.crd, c.npad) || !dalloc(&c.frc, c.npad) || !dalloc(&c.drift, c.npad) ||
        !dalloc(&c.diff, (size_t)c.npad * s) || !dalloc(&c.xi_dev, (size_t)c.npad * s) || !dalloc(&c.scal, 32)) {
        set_error("device allocation failed");
        return fail(0);
    }
    if (opts->integrator == BD_PC && (!dalloc(&c.pc_r0, c.npad) || !dalloc(&c.pc_d0, c.npad))) {
        set_error("device allocation failed");
        return fail(0);
    }
    if (opts->integrator == BD_SE && se_alloc(c) != BD_OK) return fail(0);
    if (hipMalloc(&c.st, sizeof(DevStatus)) != hipSuccess || hipMemset(c.st, 0, sizeof(DevStatus)) != hipSuccess ||
        hipHostMalloc(&c.st_host, 2 * sizeof(DevStatus)) != hipSuccess) {
        set_error("status allocation failed");
        return fail(0);
    }
    if (opts->lhdia) {
        c.mf = mob_use_matfree(*opts, nbeads);
        if ((c.mf ? mf_alloc(c) : mob_alloc(c)) != BD_OK) return fail(0);
        if (opts->mob_fctr == BD_LANC && lanczos_alloc(c) != BD_OK) return fail(0);
        if (opts->mob_fctr != BD_LANC && !c.W) {
            if (!dalloc(&c.W, (size_t)c.npad * s)) { set_error("device allocation failed"); return fail(0); }
        }
    }
    if (!c.W && !dalloc(&c.W, (size_t)c.npad)) { set_error("device allocation failed"); return fail(0); }
    if (hipStreamSynchronize(c.stream) != hipSuccess) { set_error("stream sync failed"); return fail(0); }
    return h;
}

void bd_ctx_destroy(bd_ctx *h)
{
    if (!h) return;
    Ctx &c = h->c;
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (auto e : c.prof_ev) (void)hipEventDestroy(e);
    mob_free(c);
    mf_free(c);
    stats_free(c);
    se_free(c);
    lanczos_free(c);
    forces_free(c);
    double **dp[] = {&c.crd, &c.frc, &c.drift, &c.diff, &c.xi_dev, &c.scal, &c.pc_r0, &c.pc_d0};
    for (auto p : dp)
        if (*p) (void)hipFree(*p);
    if (c.W) (void)hipFree(c.W);
    if (c.st) (void)hipFree(c.st);
    if (c.st_host) (void)hipHostFree(c.st_host);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete h;
}

int bd_set_box(bd_ctx *h, int imcon, const double box[3])
{
    if (!h || (imcon != 0 && imcon != 1)) { set_error("bad arguments"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    c.imcon = imcon;
    for (int d = 0; d < 3; d++) {
        c.box[d] = box ? box[d] : 0.0;
        if (imcon && !(c.box[d] > 0.0)) { set_error("periodic box edges must be positive"); return BD_ERR_ARG; }
    }
    return BD_OK;
}

int bd_set_topology(bd_ctx *h, int nbonds, const int *bonds, int nbond_types, const int *bond_styles,
                    const double *bond_params, int natom_types, const int *atom_types, int nvdw_types,
                    const int *vdw_styles, const double *vdw_params, int lvdw, int excluded_atoms, int mth_ptgen,
                    double rcutoff, double tskin)
{
    if (!h || nbonds < 0 || (nbonds > 0 && (!bonds || !bond_styles || !bond_params)) || excluded_atoms < 0 ||
        excluded_atoms > 3 || mth_ptgen < 0 || mth_ptgen > 2 || (nvdw_types > 0 && (!vdw_styles || !vdw_params))) {
        set_error("bad topology arguments");
        return BD_ERR_ARG;
    }
    Ctx &c = h->c;
    c.lvdw = (lvdw && nvdw_types > 0) ? 1 : 0;  // ia_setup, interaction_m.f90:37-38
    c.excluded_atoms = excluded_atoms;
    c.mth_ptgen = mth_ptgen;
    c.rcutoff = rcutoff;
    c.tskin = tskin;
    if (c.lvdw && !(rcutoff > 0.0)) { set_error("rcutoff must be positive"); return BD_ERR_ARG; }
    return forces_setup(c, nbonds, bonds, nbond_types, bond_styles, bond_params, natom_types, atom_types,
                        nvdw_types, vdw_styles, vdw_params);
}

int bd_set_coordinates(bd_ctx *h, const double *crd)
{
    if (!h || !crd) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    BD_HIP(hipMemcpyAsync(c.crd, crd, (size_t)c.n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    BD_HIP(hipStreamSynchronize(c.stream));
    c.origin_pending = true;
    return BD_OK;
}

int bd_get_coordinates(bd_ctx *h, double *crd)
{
    if (!h || !crd) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    BD_HIP(hipMemcpyAsync(crd, c.crd, (size_t)c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    BD_HIP(hipStreamSynchronize(c.stream));
    return BD_OK;
}

int bd_rpy_assemble(bd_ctx *h)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    if (!c.opt.lhdia) { set_error("context created without hydrodynamic interaction"); return BD_ERR_STATE; }
    BD_TRY(rpy_assemble(c));
    BD_HIP(hipStreamSynchronize(c.stream));
    return BD_OK;
}

int bd_get_mobility(bd_ctx *h, double *mob)
{
    if (!h || !mob) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    if (!c.mob_valid) { set_error("mobility not assembled"); return BD_ERR_STATE; }
    double *dense = nullptr;
    BD_HIP(hipMalloc(&dense, (size_t)c.n * c.n * sizeof(double)));
    int rc = mob_unpack(c, dense, c.n);
    if (rc == BD_OK) {
        if (hipMemcpyAsync(mob, dense, (size_t)c.n * c.n * sizeof(double), hipMemcpyDeviceToHost, c.stream) !=
                hipSuccess ||
            hipStreamSynchronize(c.stream) != hipSuccess) {
            set_error("copy of the mobility failed");
            rc = BD_ERR_HIP;
        }
    }
    (void)hipFree(dense);
    return rc;
}

int bd_mob_symv(bd_ctx *h, const double *x, int s, double *y)
{
    if (!h || !x || !y || s < 1 || s > MAX_S) { set_error("bad arguments"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    double *dx = nullptr, *dy = nullptr;
    BD_HIP(hipMalloc(&dx, (size_t)c.npad * s * sizeof(double)));
    BD_HIP(hipMalloc(&dy, (size_t)c.npad * s * sizeof(double)));
    int rc = BD_OK;
    if (hipMemsetAsync(dx, 0, (size_t)c.npad * s * sizeof(double), c.stream) != hipSuccess ||
        hipMemcpy2DAsync(dx, (size_t)c.npad * sizeof(double), x, (size_t)c.n * sizeof(double),
                         (size_t)c.n * sizeof(double), s, hipMemcpyHostToDevice, c.stream) != hipSuccess) {
        set_error("upload failed");
        rc = BD_ERR_HIP;
    }
    if (rc == BD_OK) rc = mob_product(c, dx, dy, s, c.npad, c.npad);
    if (rc == BD_OK) rc = download_block(c, dy, y, s);
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

int bd_brownian(bd_ctx *h, const double *xi, int s, int method, double *out, int *lconv, int *iters)
{
    if (!h || !xi || !out || s < 1 || s > MAX_S || method < BD_CHOL || method > BD_CHOL_LT) {
        set_error("bad arguments");
        return BD_ERR_ARG;
    }
    Ctx &c = h->c;
    if (!c.opt.lhdia) { set_error("context created without hydrodynamic interaction"); return BD_ERR_STATE; }
    if (!c.mob_valid) { set_error("mobility not assembled"); return BD_ERR_STATE; }
    if (method == BD_LANC && (!c.V || s != c.s)) {
        set_error("Lanczos needs a context created with mob_fctr=LANC and nts_mobsam == s");
        return BD_ERR_ARG;
    }
    if (method != BD_LANC && s > c.s) { set_error("s exceeds the context's block size"); return BD_ERR_ARG; }
    BD_TRY(upload_xi(c, xi, s));
    int lc = 1, it = 0;
    int rc = brownian_dev(c, method, s, &lc, &it);
    if (lconv) *lconv = lc;
    if (iters) *iters = it;
    if (rc != BD_OK) return rc;
    BD_TRY(download_block(c, c.diff, out, s));
    return (method == BD_LANC && !lc) ? BD_WARN_NOTCONV : BD_OK;
}

int bd_forces(bd_ctx *h, double *frc, double *energies, double *stress9, double *bndlen3)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    BD_TRY(compute_forces(c, true, true));
    int rc = check_status(c, true);
    if (rc != BD_OK) return rc;
    double e[16];
    BD_HIP(hipMemcpy(e, c.scal, 15 * sizeof(double), hipMemcpyDeviceToHost));
    if (frc) BD_HIP(hipMemcpy(frc, c.frc, (size_t)c.n * sizeof(double), hipMemcpyDeviceToHost));
    if (energies) { energies[0] = e[0]; energies[1] = e[1]; energies[2] = e[2]; }
    if (stress9) for (int k = 0; k < 9; k++) stress9[k] = e[3 + k];
    if (bndlen3) { bndlen3[0] = e[12]; bndlen3[1] = e[13]; bndlen3[2] = e[14]; }
    return BD_OK;
}

int bd_energies(bd_ctx *h, double *energies7)
{
    if (!h || !energies7) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    double e[32];
    BD_HIP(hipStreamSynchronize(c.stream));
    BD_HIP(hipMemcpy(e, c.scal, sizeof(e), hipMemcpyDeviceToHost));
    const double out[7] = {e[0], e[24], e[25], e[1], e[26], e[27], e[2]};
    for (int k = 0; k < 7; k++) energies7[k] = out[k];
    return BD_OK;
}

int bd_set_angles(bd_ctx *h, int nangles, const int *angles, int nangle_types, const int *angle_styles,
                  const double *angle_params)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    BD_TRY(set_angles(h->c, nangles, angles, nangle_types, angle_styles, angle_params));
    return BD_OK;
}

int bd_set_tethers(bd_ctx *h, int ntethers, const int *tethers, const double *tether_points, int ntether_types,
                   const int *tether_styles, const double *tether_params)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    BD_TRY(set_tethers(h->c, ntethers, tethers, tether_points, ntether_types, tether_styles, tether_params));
    return BD_OK;
}

int bd_set_externals(bd_ctx *h, int nexternals, const int *external_styles, const double *external_params)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    BD_TRY(set_externals(h->c, nexternals, external_styles, external_params));
    return BD_OK;
}

int bd_set_charges(bd_ctx *h, const double *charge)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    BD_TRY(set_charges(h->c, charge));
    return BD_OK;
}

int bd_pair_table(bd_ctx *h, int *pairs2, int cap, int *npairs, int *rebuilt)
{
    if (!h || !npairs) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    std::vector<int> p;
    BD_TRY(pair_table_host(c, p));
    int np = (int)(p.size() / 2);
    *npairs = np;
    if (rebuilt) *rebuilt = c.last_rebuilt;
    if (np > cap || (!pairs2 && np > 0)) { set_error("pair buffer too small"); return BD_ERR_CAPACITY; }
    if (np) std::memcpy(pairs2, p.data(), p.size() * sizeof(int));
    return BD_OK;
}

int bd_get_rng(bd_ctx *h, uint64_t *seed, uint64_t *offset)
{
    if (!h || !seed || !offset) { set_error("null argument"); return BD_ERR_ARG; }
    *seed = h->c.rng_seed;
    *offset = h->c.rng_offset;
    return BD_OK;
}

int bd_resume_run(bd_ctx *h, const double *com)
{
    if (!h || !com) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    BD_HIP(hipMemcpyAsync(c.scal + 16, com, 3 * sizeof(double), hipMemcpyHostToDevice, c.stream));
    BD_HIP(hipStreamSynchronize(c.stream));
    for (int d = 0; d < 3; d++) c.info.molecule_com[d] = com[d];
    c.origin_pending = false;
    return BD_OK;
}

int bd_set_rng(bd_ctx *h, uint64_t seed, uint64_t offset)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    h->c.rng_seed = seed;
    h->c.rng_offset = offset;
    return BD_OK;
}

int bd_run(bd_ctx *h, int nsteps, const double *xi, long long nxi, long long *used)
{
    if (!h || nsteps < 0) { set_error("bad arguments"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    if (!c.have_topology) { set_error("topology not set"); return BD_ERR_STATE; }
    long long pos = 0;
    const long long nts0 = c.info.nts;
    BD_HIP(hipMemsetAsync(&c.st->steps_done, 0, sizeof(int), c.stream));
    if (c.limud() && c.origin_pending) {
        // isolated molecule: COM to the origin before the loop (bd_solver_m.f90:161),
        // once per run: later calls continue the same bds_run loop
        hipLaunchKernelGGL(to_origin_kernel, dim3(1), dim3(1024), 0, c.stream, c.crd, c.nb, (double *)nullptr, c.st);
    }
    c.origin_pending = false;
    int rc = BD_OK;
    for (int k = 0; k < nsteps && rc == BD_OK; k++) {
        long long cons = 0;
        rc = em_step(c, xi ? xi + pos : nullptr, xi ? nxi - pos : 0, &cons, nullptr, nullptr);
        pos += cons;
        if (rc == BD_OK && c.opt.lhdia == 0 && (k % 64) == 63) {
            rc = check_status(c, false);
            // an overflowed list is scanned in full by the forces until it is
            // rebuilt with room: grow it here rather than at the end of the run
            if (rc == BD_OK && c.st_host->overflow) rc = grow_pairs(c);
        }
    }
    if (used) *used = pos;
    int rc2 = check_status(c, false);
    if (rc == BD_OK) rc = rc2;
    if (c.st_host->overflow) BD_TRY(grow_pairs(c));  // next build gets more room
    c.info.n_pair_rebuilds = c.st_host->nrebuild;
    // nts counts completed steps only (bds_run returns before nts+1 on error)
    c.info.nts = nts0 + c.st_host->steps_done;
    return rc;
}

int bd_get_runinfo(bd_ctx *h, bd_runinfo *info)
{
    if (!h || !info) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    double com[3];
    BD_HIP(hipStreamSynchronize(c.stream));
    BD_HIP(hipMemcpy(com, c.scal + 16, 3 * sizeof(double), hipMemcpyDeviceToHost));
    *info = c.info;
    for (int d = 0; d < 3; d++) info->molecule_com[d] = com[d];
    return BD_OK;
}

int bd_reset_run(bd_ctx *h)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    c.cntr_mobsam = 0;
    c.info = bd_runinfo{};
    c.verlet_first = true;
    c.origin_pending = true;
    BD_HIP(hipMemsetAsync(c.scal + 16, 0, 3 * sizeof(double), c.stream));
    BD_HIP(hipMemsetAsync(c.st, 0, sizeof(DevStatus), c.stream));
    BD_HIP(hipStreamSynchronize(c.stream));
    return BD_OK;
}

int bd_rgsq(bd_ctx *h, double *rgsq)
{
    if (!h || !rgsq) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    hipLaunchKernelGGL(rgsq_kernel, dim3(1), dim3(1024), 0, c.stream, c.crd, c.nb, c.scal + 20);
    BD_HIP(hipGetLastError());
    BD_HIP(hipMemcpyAsync(rgsq, c.scal + 20, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    BD_HIP(hipStreamSynchronize(c.stream));
    return BD_OK;
}

int bd_prof_enable(bd_ctx *h, int on)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    prof_flush(c);
    if (on < 0 || on > 1 + MAX_S) { set_error("prof: on must be 0, 1 or 1 + block width"); return BD_ERR_ARG; }
    c.prof_on = on != 0;
    c.prof_width = on > 1 ? on - 1 : 0;
    c.prof_ms = 0.0;
    c.prof_count = 0;
    return BD_OK;
}

int bd_prof_get(bd_ctx *h, double *total_ms, long long *launches, double *tile_bytes)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    prof_flush(c);
    if (total_ms) *total_ms = c.prof_ms;
    if (launches) *launches = c.prof_count;
    if (tile_bytes)
        *tile_bytes = (double)c.ntiles_lower * TILE * TILE * (c.mob32 ? sizeof(float) : sizeof(double));
    return BD_OK;
}

int bd_sync(bd_ctx *h)
{
    if (!h) { set_error("null context"); return BD_ERR_ARG; }
    BD_HIP(hipStreamSynchronize(h->c.stream));
    return BD_OK;
}

}  // extern "C"
