// bd_lanczos.hip -- Brownian displacement B.xi on the device.
//
// Block Lanczos (brown_m.f90:292-512; s = 1 is the scalar Lanczos of
// brown_m.f90:185-288) and Cholesky (brown_m.f90:516-554).
//
// Per Krylov iteration j the device runs:
//   W = D V_j                                  (symmetric tile product)
//   W -= V_{j-1} B_{j-1}^T ; A_j = V_j^T W     (lanczos_a)
//   W -= V_j symL(A_j)                          (lanczos_b)
//   W = V_{j+1} B_j                             (CholeskyQR2, FP64)
//   y_j = V_{1:j} sqrt(H_j) E_1 R_0             (small problem, FP64)
//   err_j = ||y_j - y_{j-1}|| / ||y_{j-1}||     (2-norm for s = 1, max column
//                                                 1-norm for s > 1)
// The reference factorises with pivoted Householder QR (dgeqp3 + dorgqr) and
// diagonalises H_j with dstevd/dsbevd.  Here the QR is CholeskyQR2 (the same
// block Krylov space, B_j upper triangular, so H_j has bandwidth s) and
// sqrt(H_j) E_1 R_0 is evaluated as H_j * H_j^{-1/2} X with
//   H^{-1/2} = (2/pi) int_0^inf (H + t^2 I)^{-1} dt,
// discretised by a sinh-substituted trapezoid rule (t = exp(u), u = u0 +
// sinh(v)) whose node count is chosen from the spectral bounds [m, M] of H_j
// for 3e-15 relative accuracy.  M is the Gershgorin bound; m is the largest
// M 2^-k for which the banded Cholesky of H_j - m I succeeds.  Each node is
// one banded SPD solve of bandwidth s, run by its own wavefront, so a whole
// evaluation costs a few tens of microseconds instead of a serial O(js^3)
// eigensolve.  In exact arithmetic the result equals the reference's
// P sqrt(Lambda) P^T E_1 R.
#include "bd_internal.h"
#include "bd_lanczos_dev.h"

namespace bd {

// part[b] = A^T B over rows (npad rows), A, B npad x S.
template <int S>
__global__ __launch_bounds__(RB) void gram_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                                  int ld, int rows, double *__restrict__ part,
                                                  const DevStatus *st)
{
    if (lz_done(st)) return;
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += GBLK * RB) {
        double a[S], b[S];
#pragma unroll
        for (int k = 0; k < S; k++) { a[k] = A[(long)k * ld + r]; b[k] = B[(long)k * ld + r]; }
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += a[ci] * b[cj];
    }
    block_reduce_store<S>(acc, part + blockIdx.x * S * S);
}

// W -= V_prev B_prev^T (if V_prev); part = V_j^T W.
template <int S>
__global__ __launch_bounds__(RB) void lanczos_a_kernel(double *__restrict__ W, const double *__restrict__ Vj,
                                                       const double *__restrict__ Vp, const double *__restrict__ Bp,
                                                       int ld, int rows, double *__restrict__ part,
                                                       const DevStatus *st)
{
    if (lz_done(st)) return;
    double Bm[S][S];
    if (Vp) {
#pragma unroll
        for (int a = 0; a < S; a++)
#pragma unroll
            for (int b = 0; b < S; b++) Bm[a][b] = Bp[a + b * S];
    }
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += GBLK * RB) {
        double w[S], v[S];
#pragma unroll
        for (int k = 0; k < S; k++) { w[k] = W[(long)k * ld + r]; v[k] = Vj[(long)k * ld + r]; }
        if (Vp) {
            double p[S];
#pragma unroll
            for (int k = 0; k < S; k++) p[k] = Vp[(long)k * ld + r];
#pragma unroll
            for (int c = 0; c < S; c++) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < S; k++) t += p[k] * Bm[c][k];
                w[c] = w[c] - t;
            }
#pragma unroll
            for (int k = 0; k < S; k++) W[(long)k * ld + r] = w[k];
        }
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += v[ci] * w[cj];
    }
    block_reduce_store<S>(acc, part + blockIdx.x * S * S);
}

// A = sum(part) (stored); W -= V_j symL(A); gpart = W^T W.
template <int S>
__global__ __launch_bounds__(RB) void lanczos_b_kernel(double *__restrict__ W, const double *__restrict__ Vj,
                                                       int ld, int rows, const double *__restrict__ part,
                                                       double *__restrict__ Aout, int do_update,
                                                       double *__restrict__ gpart, const DevStatus *st)
{
    if (lz_done(st)) return;
    __shared__ double sA[S * S];
    sum_partials<S>(part, sA);
    if (blockIdx.x == 0 && threadIdx.x < S * S) Aout[threadIdx.x] = sA[threadIdx.x];
    if (!do_update) return;
    double As[S][S];
#pragma unroll
    for (int a = 0; a < S; a++)
#pragma unroll
        for (int b = 0; b < S; b++) As[a][b] = (a >= b) ? sA[a + b * S] : sA[b + a * S];
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += GBLK * RB) {
        double w[S], v[S];
#pragma unroll
        for (int k = 0; k < S; k++) { w[k] = W[(long)k * ld + r]; v[k] = Vj[(long)k * ld + r]; }
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < S; k++) t += v[k] * As[k][c];
            w[c] = w[c] - t;
        }
#pragma unroll
        for (int k = 0; k < S; k++) W[(long)k * ld + r] = w[k];
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += w[ci] * w[cj];
    }
    block_reduce_store<S>(acc, gpart + blockIdx.x * S * S);
}

// CholeskyQR pass: G = sum(gpart) = R^T R; Q = X R^{-1} (into dst);
// gpart_out = Q^T Q.  pass 2 also forms Rtot = R2 * R1 (R1 read from rprev).
template <int S>
__global__ __launch_bounds__(RB) void cholqr_kernel(const double *__restrict__ X, double *__restrict__ Q,
                                                    int ld, int rows, const double *__restrict__ gpart,
                                                    double *__restrict__ gpart_out, double *__restrict__ Rsave,
                                                    const double *__restrict__ rprev, DevStatus *st)
{
    if (lz_done(st)) return;
    __shared__ double sG[S * S];
    sum_partials<S>(gpart, sG);
    double R[S][S];
    bool ok = small_chol<S>(sG, R);
    if (!ok) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicExch(&st->ierr, 2);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (rprev) {
            // Rtot = R * Rprev (both upper triangular)
            for (int i = 0; i < S; i++)
                for (int j = 0; j < S; j++) {
                    double t = 0.0;
                    for (int k = 0; k < S; k++) t += R[i][k] * rprev[k + j * S];
                    Rsave[i + j * S] = t;
                }
        } else {
            for (int i = 0; i < S; i++)
                for (int j = 0; j < S; j++) Rsave[i + j * S] = R[i][j];
        }
    }
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += GBLK * RB) {
        double x[S], q[S];
#pragma unroll
        for (int k = 0; k < S; k++) x[k] = X[(long)k * ld + r];
        // solve q R = x
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = x[c];
#pragma unroll
            for (int k = 0; k < c; k++) t -= q[k] * R[k][c];
            q[c] = t / R[c][c];
        }
#pragma unroll
        for (int k = 0; k < S; k++) Q[(long)k * ld + r] = q[k];
        if (gpart_out) {
#pragma unroll
            for (int cj = 0; cj < S; cj++)
#pragma unroll
                for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += q[ci] * q[cj];
        }
    }
    if (gpart_out) block_reduce_store<S>(acc, gpart_out + blockIdx.x * S * S);
}

// ------------------------------------------------------------ reorthogonalisation
// part[b][m*S + k] = sum over the block's rows of V[r, m] W[r, k], m < ncol.
template <int S>
__global__ __launch_bounds__(RB) void reorth_coef_kernel(const double *__restrict__ V, const double *__restrict__ W,
                                                         int ld, int rows, int ncol, double *__restrict__ part,
                                                         const DevStatus *st)
{
    if (lz_done(st)) return;
    __shared__ double sW[RB][S];
    __shared__ double red[RB / 64][S];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    // each block owns a contiguous row range
    const int per = (rows + GBLK - 1) / GBLK;
    const int r0 = blockIdx.x * per, r1 = min(rows, r0 + per);
    for (int m = 0; m < ncol; m++) {
        double acc[S];
#pragma unroll
        for (int k = 0; k < S; k++) acc[k] = 0.0;
        for (int r = r0 + threadIdx.x; r < r1; r += RB) {
            double v = V[(long)m * ld + r];
#pragma unroll
            for (int k = 0; k < S; k++) acc[k] += v * W[(long)k * ld + r];
        }
#pragma unroll
        for (int k = 0; k < S; k++) {
            double x = acc[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            if (lane == 0) red[wv][k] = x;
        }
        __syncthreads();
        if (threadIdx.x < S) {
            double x = 0.0;
            for (int w = 0; w < RB / 64; w++) x += red[w][threadIdx.x];
            part[(long)blockIdx.x * ncol * S + m * S + threadIdx.x] = x;
        }
        __syncthreads();
    }
    (void)sW;
}

// coef[e] = sum_b part[b][e], e < ncol*S
__global__ void reorth_sum_kernel(const double *__restrict__ part, int ne, double *__restrict__ coef,
                                  const DevStatus *st)
{
    if (st->converged_j != 0 || st->ierr != 0) return;
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    double v = 0.0;
    for (int b = 0; b < GBLK; b++) v += part[(long)b * ne + e];
    coef[e] = v;
}

// W -= V(:, 0:ncol) coef; optional Gram partial of the result.
template <int S>
__global__ __launch_bounds__(RB) void reorth_apply_kernel(const double *__restrict__ V, double *__restrict__ W, int ld,
                                                          int rows, int ncol, const double *__restrict__ coef,
                                                          double *__restrict__ gpart, const DevStatus *st)
{
    if (lz_done(st)) return;
    extern __shared__ double sc[];
    for (int e = threadIdx.x; e < ncol * S; e += blockDim.x) sc[e] = coef[e];
    __syncthreads();
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += GBLK * RB) {
        double w[S];
#pragma unroll
        for (int k = 0; k < S; k++) w[k] = W[(long)k * ld + r];
        for (int m = 0; m < ncol; m++) {
            double v = V[(long)m * ld + r];
#pragma unroll
            for (int k = 0; k < S; k++) w[k] -= v * sc[m * S + k];
        }
#pragma unroll
        for (int k = 0; k < S; k++) W[(long)k * ld + r] = w[k];
        if (gpart) {
#pragma unroll
            for (int cj = 0; cj < S; cj++)
#pragma unroll
                for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += w[ci] * w[cj];
        }
    }
    if (gpart) block_reduce_store<S>(acc, gpart + blockIdx.x * S * S);
}

// ------------------------------------------------------------ small problem
// One block (64 threads) per probe: Cholesky of H - mu_b I, mu_b = M 2^-(b+1).
template <int S>
__global__ __launch_bounds__(64) void bounds_kernel(const double *__restrict__ Ablk, const double *__restrict__ Bblk,
                                                    int js, int *__restrict__ ok, DevStatus *st)
{
    if (lz_done(st)) return;
    extern __shared__ double band[];
    const int ldb = js;
    double M = gersh_hi<S>(Ablk, Bblk, js);
    double mu = ldexp(M, -(int)(blockIdx.x + 1));
    load_band<S>(Ablk, Bblk, js, mu, band, ldb);
    bool good = band_chol<S>(band, ldb, js);
    if (threadIdx.x == 0) ok[blockIdx.x] = good ? 1 : 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->lam_hi = M;
}


// Node k of the quadrature: solves (H + t_k^2 I) Z = X, X = E1 R0, and
// stores w_k Z into part[k] (js x S col-major, ld js).
template <int S>
__global__ __launch_bounds__(64) void quad_kernel(const double *__restrict__ Ablk, const double *__restrict__ Bblk,
                                                  const double *__restrict__ R0, int js, const int *__restrict__ ok,
                                                  double *__restrict__ part, DevStatus *st)
{
    if (lz_done(st)) return;
    extern __shared__ double smem[];
    double *band = smem;               // (S+1) x js
    double *z = smem + (S + 1) * js;   // js x S
    const double M = st->lam_hi;
    int kb = -1;
    for (int b = 0; b < NBOUND; b++)
        if (ok[b]) { kb = b; break; }
    if (kb < 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicExch(&st->ierr, 2);
        return;
    }
    const double m = ldexp(M, -(kb + 1));
    double kappa = M / m;
    int dec = (int)ceil(log10(kappa));
    if (dec < 1) dec = 1;
    if (dec > 13) dec = 13;
    const int N = c_qN[dec];
    const double h = c_qh[dec];
    const int k = (int)blockIdx.x - N;
    double *out = part + (long)blockIdx.x * js * S;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->lam_lo = m;
        st->nq = 2 * N + 1;
    }
    if ((int)blockIdx.x > 2 * N) return;
    const double u0 = 0.25 * (log(m) + log(M));
    const double v = k * h;
    const double u = u0 + sinh(v);
    const double w = (2.0 / M_PI) * h * cosh(v) * exp(u);
    const double t2 = exp(2.0 * u);
    load_band<S>(Ablk, Bblk, js, -t2, band, js);
    if (!band_chol<S>(band, js, js)) {
        if (threadIdx.x == 0) atomicExch(&st->ierr, 2);
        return;
    }
    // right-hand sides: rows 0..S-1 = R0, others 0
    for (int e = threadIdx.x; e < js * S; e += blockDim.x) {
        int i = e % js, c = e / js;
        z[e] = (i < S) ? R0[i + c * S] : 0.0;
    }
    __syncthreads();
    if (threadIdx.x < S) {
        double *zc = z + threadIdx.x * js;
        for (int i = 0; i < js; i++) {        // L y = x
            double t = zc[i];
            for (int d = 1; d <= S && d <= i; d++) t -= band[d * js + i - d] * zc[i - d];
            zc[i] = t / band[i];
        }
        for (int i = js - 1; i >= 0; i--) {   // L^T z = y
            double t = zc[i];
            for (int d = 1; d <= S && i + d < js; d++) t -= band[d * js + i] * zc[i + d];
            zc[i] = t / band[i];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < js * S; e += blockDim.x) out[e] = w * z[e];
}

// ysum = sum_k part[k] over the st->nq nodes, fixed order.  Block of 256
// threads handles 64 entries x 4 node stripes.
template <int S>
__global__ __launch_bounds__(256) void nodesum_kernel(const double *__restrict__ part, int js,
                                                      double *__restrict__ ysum, const DevStatus *st)
{
    if (lz_done(st)) return;
    __shared__ double sh[4][64];
    const int e = blockIdx.x * 64 + threadIdx.x % 64, q = threadIdx.x / 64;
    const int nq = st->nq;
    double v = 0.0;
    if (e < js * S)
        for (int k = q; k < nq; k += 4) v += part[(long)k * js * S + e];
    sh[q][threadIdx.x % 64] = v;
    __syncthreads();
    if (q == 0 && e < js * S) ysum[e] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

// y = V(:, 1:js) C ; partial norms of (y - yold) and yold per column.
template <int S>
__global__ __launch_bounds__(RB) void ycomb_kernel(const double *__restrict__ V, int ld, int rows, int js,
                                                   const double *__restrict__ Ablk, const double *__restrict__ Bblk,
                                                   const double *__restrict__ ysum, double *__restrict__ y,
                                                   const double *__restrict__ yold, int want_err, int two_norm,
                                                   double *__restrict__ part, const DevStatus *st)
{
    if (lz_done(st)) return;
    extern __shared__ double sC[];
    // C = H ysum (banded, symmetric)
    for (int e = threadIdx.x; e < js * S; e += blockDim.x) {
        int i = e % js, c = e / js;
        double v = 0.0;
        for (int l = max(0, i - S); l <= min(js - 1, i + S); l++) {
            double hv = (i >= l) ? h_entry<S>(Ablk, Bblk, js, i, l) : h_entry<S>(Ablk, Bblk, js, l, i);
            v += hv * ysum[l + c * js];
        }
        sC[e] = v;
    }
    __syncthreads();
    double dn[S], on[S];
#pragma unroll
    for (int k = 0; k < S; k++) { dn[k] = 0.0; on[k] = 0.0; }
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += GBLK * RB) {
        double acc[S];
#pragma unroll
        for (int k = 0; k < S; k++) acc[k] = 0.0;
        for (int m = 0; m < js; m++) {
            double vm = V[(long)m * ld + r];
#pragma unroll
            for (int k = 0; k < S; k++) acc[k] += vm * sC[m + k * js];
        }
#pragma unroll
        for (int k = 0; k < S; k++) {
            y[(long)k * ld + r] = acc[k];
            if (want_err) {
                double o = yold[(long)k * ld + r];
                double d = acc[k] - o;
                if (two_norm) { dn[k] += d * d; on[k] += o * o; }
                else { dn[k] += fabs(d); on[k] += fabs(o); }
            }
        }
    }
    if (!want_err) return;
    __shared__ double sh[RB / 64][2 * S];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < 2 * S; k++) {
        double v = k < S ? dn[k] : on[k - S];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sh[wv][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < 2 * S) {
        double v = 0.0;
        for (int w = 0; w < RB / 64; w++) v += sh[w][threadIdx.x];
        part[blockIdx.x * 2 * S + threadIdx.x] = v;
    }
}

template <int S>
__global__ __launch_bounds__(64) void finalize_kernel(const double *__restrict__ part, int j, int want_err,
                                                      int two_norm, double tol, DevStatus *st)
{
    if (lz_done(st)) return;
    const int t = threadIdx.x;
    __shared__ double tot[2 * S];
    if (want_err) {
        // lane t: entry t % (2S), block stripe t / (2S)
        constexpr int E = 2 * S, STR = 64 / E;
        double v = 0.0;
        if (t < E * STR)
            for (int b = t / E; b < GBLK; b += STR) v += part[b * E + t % E];
        __shared__ double sh[64];
        sh[t] = v;
        __syncthreads();
        if (t < E) {
            double a = 0.0;
            for (int q = 0; q < STR; q++) a += sh[q * E + t];
            tot[t] = a;
        }
        __syncthreads();
    }
    if (t != 0) return;
    st->last_j = j;
    if (!want_err) return;
    double num = 0.0, den = 0.0;
    if (two_norm) {
        for (int k = 0; k < S; k++) { num += tot[k]; den += tot[S + k]; }
        num = sqrt(num); den = sqrt(den);
    } else {
        for (int k = 0; k < S; k++) { num = fmax(num, tot[k]); den = fmax(den, tot[S + k]); }
    }
    double err = num / den;
    st->err = err;
    if (err <= tol) st->converged_j = j;
}

__global__ void status_reset_kernel(DevStatus *st)
{
    st->converged_j = 0;
    st->last_j = 0;
    st->err = 0.0;
}

// ---------------------------------------------------------------- host
int lanczos_alloc(Ctx &c)
{
    const int s = c.s;
    const int ncol = s * c.dimk_max;
    BD_HIP(hipMalloc(&c.V, (size_t)c.npad * ncol * sizeof(double)));
    BD_HIP(hipMemsetAsync(c.V, 0, (size_t)c.npad * ncol * sizeof(double), c.stream));
    BD_HIP(hipMalloc(&c.W, (size_t)c.npad * s * sizeof(double)));
    BD_HIP(hipMemsetAsync(c.W, 0, (size_t)c.npad * s * sizeof(double), c.stream));
    BD_HIP(hipMalloc(&c.Y, (size_t)2 * c.npad * s * sizeof(double)));
    BD_HIP(hipMemsetAsync(c.Y, 0, (size_t)2 * c.npad * s * sizeof(double), c.stream));
    BD_HIP(hipMalloc(&c.Ablk, (size_t)s * s * c.dimk_max * sizeof(double)));
    BD_HIP(hipMalloc(&c.Bblk, (size_t)s * s * c.dimk_max * sizeof(double)));
    BD_HIP(hipMalloc(&c.R0, (size_t)s * s * 2 * sizeof(double)));
    BD_HIP(hipMalloc(&c.gpart, (size_t)2 * GBLK * MAX_S * MAX_S * sizeof(double)));
    BD_HIP(hipMalloc(&c.Cq, (size_t)(2 * 125 + 1) * ncol * s * sizeof(double)));
    BD_HIP(hipMalloc(&c.Ccoef, (size_t)ncol * s * sizeof(double)));
    BD_HIP(hipMalloc(&c.bound_ok, NBOUND * sizeof(int)));
    BD_HIP(hipMalloc(&c.npart, (size_t)GBLK * 2 * MAX_S * sizeof(double)));
    if (c.opt.lanc_reorth) {
        BD_HIP(hipMalloc(&c.Rq, (size_t)GBLK * ncol * s * sizeof(double)));
        BD_HIP(hipMalloc(&c.Rcoef, (size_t)ncol * s * sizeof(double)));
    }
    BD_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    for (int k = 0; k < LZ_NEV; k++) BD_HIP(hipEventCreateWithFlags(&c.lz_ev[k], hipEventDisableTiming));
    for (int k = 0; k < 2; k++) BD_HIP(hipEventCreateWithFlags(&c.lz_chunk_ev[k], hipEventDisableTiming));
    return BD_OK;
}

void lanczos_free(Ctx &c)
{
    if (c.side) (void)hipStreamSynchronize(c.side);
    for (int k = 0; k < LZ_NEV; k++)
        if (c.lz_ev[k]) (void)hipEventDestroy(c.lz_ev[k]);
    for (int k = 0; k < 2; k++)
        if (c.lz_chunk_ev[k]) (void)hipEventDestroy(c.lz_chunk_ev[k]);
    if (c.side) (void)hipStreamDestroy(c.side);
    c.side = nullptr;
    double **ps[] = {&c.V, &c.W, &c.Y, &c.Ablk, &c.Bblk, &c.R0, &c.gpart, &c.Cq, &c.Ccoef, &c.npart,
                     &c.Rq, &c.Rcoef};
    for (auto p : ps) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (c.bound_ok) (void)hipFree(c.bound_ok);
    c.bound_ok = nullptr;
    if (c.chol) (void)hipFree(c.chol);
    c.chol = nullptr;
}

// Host driver of one block-Lanczos solve.  The Krylov recurrence (product,
// three-term update, CholeskyQR2) runs on the context stream; the small
// problem and the convergence test of iteration j run on a second stream as
// soon as iteration j's blocks exist, overlapping the next products.  The
// host looks at the convergence flag once per chunk of iterations, one chunk
// behind, so it never stalls the device; iterations enqueued past the first
// converged j find the flag set and return at once.
template <int S>
static int lanczos_t(Ctx &c, const double *xi, double *out, int *lconv, int *iters)
{
    const int ld = c.npad, rows = c.npad;
    const int dmax = c.dimk_max;
    const bool two_norm = (S == 1);
    double *g0 = c.gpart, *g1 = c.gpart + GBLK * MAX_S * MAX_S;
    hipStream_t sm = c.stream, ss = c.side;

    hipLaunchKernelGGL(status_reset_kernel, dim3(1), dim3(1), 0, sm, c.st);
    // xi = V_1 R0 (CholeskyQR2)
    hipLaunchKernelGGL((gram_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, xi, xi, ld, rows, g0, c.st);
    hipLaunchKernelGGL((cholqr_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, xi, c.V, ld, rows, g0, g1,
                       c.R0 + S * S, (const double *)nullptr, c.st);
    hipLaunchKernelGGL((cholqr_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, c.V, c.V, ld, rows, g1,
                       (double *)nullptr, c.R0, (const double *)(c.R0 + S * S), c.st);
    BD_HIP(hipGetLastError());

    const int chunk = 4;
    int pending = -1;   // slot of the status copy of the previous chunk
    int slot = 0;
    bool stop = false;
    for (int j = 1; j <= dmax && !stop; j++) {
        double *Vj = c.V + (size_t)(j - 1) * S * ld;
        const double *Vp = (j > 1) ? c.V + (size_t)(j - 2) * S * ld : nullptr;
        const double *Bp = (j > 1) ? c.Bblk + (size_t)(j - 2) * S * S : nullptr;
        BD_TRY(mob_product_lz(c, Vj, c.W, S, ld, ld));
        hipLaunchKernelGGL((lanczos_a_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, c.W, Vj, Vp, Bp, ld, rows, g0, c.st);
        const int upd = (j < dmax) ? 1 : 0;
        hipLaunchKernelGGL((lanczos_b_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, c.W, Vj, ld, rows, g0,
                           c.Ablk + (size_t)(j - 1) * S * S, upd, g1, c.st);
        // A_j and V_1..V_j are final here: hand iteration j to the side stream
        hipEvent_t ev = c.lz_ev[j % LZ_NEV];
        BD_HIP(hipEventRecord(ev, sm));
        if (upd && c.opt.lanc_reorth) {
            // two classical Gram-Schmidt passes against V_1..V_j (FP64)
            const int ncol = j * S;
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL((reorth_coef_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, c.V, c.W, ld, rows, ncol,
                                   c.Rq, c.st);
                hipLaunchKernelGGL(reorth_sum_kernel, dim3((ncol * S + 255) / 256), dim3(256), 0, sm, c.Rq,
                                   ncol * S, c.Rcoef, c.st);
                hipLaunchKernelGGL((reorth_apply_kernel<S>), dim3(GBLK), dim3(RB), (size_t)ncol * S * sizeof(double),
                                   sm, c.V, c.W, ld, rows, ncol, c.Rcoef, pass == 1 ? g1 : (double *)nullptr, c.st);
            }
        }
        if (upd) {
            double *Vn = c.V + (size_t)j * S * ld;
            double *Bj = c.Bblk + (size_t)(j - 1) * S * S;
            hipLaunchKernelGGL((cholqr_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, c.W, Vn, ld, rows, g1, g0,
                               c.R0 + S * S /* R1 scratch */, (const double *)nullptr, c.st);
            hipLaunchKernelGGL((cholqr_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, Vn, Vn, ld, rows, g0,
                               (double *)nullptr, Bj, (const double *)(c.R0 + S * S), c.st);
        }
        BD_HIP(hipGetLastError());
        if (j >= 2) {
            BD_HIP(hipStreamWaitEvent(ss, ev, 0));
            const int js = j * S;
            size_t band_bytes = (size_t)(S + 1) * js * sizeof(double);
            hipLaunchKernelGGL((bounds_kernel<S>), dim3(NBOUND), dim3(64), band_bytes, ss, c.Ablk, c.Bblk, js,
                               c.bound_ok, c.st);
            const int nq = 2 * 125 + 1;
            hipLaunchKernelGGL((quad_kernel<S>), dim3(nq), dim3(64), band_bytes + (size_t)js * S * sizeof(double), ss,
                               c.Ablk, c.Bblk, c.R0, js, c.bound_ok, c.Cq, c.st);
            hipLaunchKernelGGL((nodesum_kernel<S>), dim3((js * S + 63) / 64), dim3(256), 0, ss, c.Cq, js, c.Ccoef,
                               c.st);
            double *ycur = c.Y + (size_t)(j % 2) * S * ld;
            const double *yold = c.Y + (size_t)((j + 1) % 2) * S * ld;
            const int want = (j > 2) ? 1 : 0;
            hipLaunchKernelGGL((ycomb_kernel<S>), dim3(GBLK), dim3(RB), (size_t)js * S * sizeof(double), ss, c.V, ld,
                               rows, js, c.Ablk, c.Bblk, c.Ccoef, ycur, yold, want, two_norm ? 1 : 0, c.npart, c.st);
            hipLaunchKernelGGL((finalize_kernel<S>), dim3(1), dim3(64), 0, ss, c.npart, j, want, two_norm ? 1 : 0,
                               c.opt.lanc_tol, c.st);
            BD_HIP(hipGetLastError());
        }
        if (j % chunk == 0 || j == dmax) {
            // snapshot the status after this chunk's tests; read the previous one
            BD_HIP(hipMemcpyAsync(&c.st_host[slot], c.st, sizeof(DevStatus), hipMemcpyDeviceToHost, ss));
            BD_HIP(hipEventRecord(c.lz_chunk_ev[slot], ss));
            if (pending >= 0) {
                BD_HIP(hipEventSynchronize(c.lz_chunk_ev[pending]));
                if (c.st_host[pending].ierr || c.st_host[pending].converged_j) stop = true;
            }
            pending = slot;
            slot ^= 1;
        }
    }
    BD_HIP(hipStreamSynchronize(ss));
    BD_HIP(hipStreamSynchronize(sm));
    BD_HIP(hipMemcpy(c.st_host, c.st, sizeof(DevStatus), hipMemcpyDeviceToHost));
    if (c.st_host->ierr == 1) {
        set_error("bond length too large (overstretched bond)");
        return BD_ERR_BOND;
    }
    if (c.st_host->ierr) {
        set_error("Lanczos small problem or CholeskyQR failed (matrix not SPD?)");
        return BD_ERR_FACTOR;
    }
    int jf = c.st_host->converged_j ? c.st_host->converged_j : c.st_host->last_j;
    *lconv = c.st_host->converged_j ? 1 : 0;
    *iters = c.st_host->converged_j ? c.st_host->converged_j : dmax;
    const double *yres = c.Y + (size_t)(jf % 2) * S * ld;
    BD_HIP(hipMemcpyAsync(out, yres, (size_t)S * ld * sizeof(double), hipMemcpyDeviceToDevice, sm));
    return BD_OK;
}

int brownian_lanczos(Ctx &c, const double *xi, double *out, int s, int *lconv, int *iters)
{
    if (s != c.s) { set_error("block size differs from the context's nts_mobsam"); return BD_ERR_ARG; }
    switch (s) {
    case 1: return lanczos_t<1>(c, xi, out, lconv, iters);
    case 2: return lanczos_t<2>(c, xi, out, lconv, iters);
    case 3: return lanczos_t<3>(c, xi, out, lconv, iters);
    case 4: return lanczos_t<4>(c, xi, out, lconv, iters);
    case 5: return lanczos_t<5>(c, xi, out, lconv, iters);
    case 6: return lanczos_t<6>(c, xi, out, lconv, iters);
    case 7: return lanczos_t<7>(c, xi, out, lconv, iters);
    case 8: return lanczos_t<8>(c, xi, out, lconv, iters);
    default: set_error("s must be in 1..8"); return BD_ERR_ARG;
    }
}

// ------------------------------------------------------------ Cholesky
constexpr int CNB = 32;

// Factor the diagonal block A[k0:k0+nb, k0:k0+nb] (lower) in LDS.
__global__ __launch_bounds__(256) void potrf_diag_kernel(double *A, int ld, int k0, int nb, DevStatus *st)
{
    if (st->ierr) return;
    __shared__ double a[CNB][CNB + 1];
    __shared__ int fail;
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        int i = e % nb, j = e / nb;
        a[i][j] = A[(long)(k0 + j) * ld + k0 + i];
    }
    if (threadIdx.x == 0) fail = 0;
    __syncthreads();
    for (int j = 0; j < nb; j++) {
        if (threadIdx.x == 0) {
            double d = a[j][j];
            if (!(d > 0.0)) fail = 1;
            a[j][j] = sqrt(fmax(d, 0.0));
        }
        __syncthreads();
        if (fail) {
            if (threadIdx.x == 0) st->ierr = 2;
            return;
        }
        int i = j + 1 + threadIdx.x;
        if (i < nb) a[i][j] /= a[j][j];
        __syncthreads();
        for (int e = threadIdx.x; e < (nb - j - 1) * (nb - j - 1); e += blockDim.x) {
            int ii = j + 1 + e % (nb - j - 1), jj = j + 1 + e / (nb - j - 1);
            if (ii >= jj) a[ii][jj] -= a[ii][j] * a[jj][j];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        int i = e % nb, j = e / nb;
        A[(long)(k0 + j) * ld + k0 + i] = (i >= j) ? a[i][j] : 0.0;
    }
}

// Panel: rows k0+nb..n-1 of block column k0 := A21 L11^{-T}.
__global__ __launch_bounds__(256) void potrf_panel_kernel(double *A, int ld, int n, int k0, int nb, DevStatus *st)
{
    if (st->ierr) return;
    __shared__ double L[CNB][CNB + 1];
    for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
        int i = e % nb, j = e / nb;
        L[i][j] = A[(long)(k0 + j) * ld + k0 + i];
    }
    __syncthreads();
    int r = k0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double x[CNB];
    for (int j = 0; j < nb; j++) x[j] = A[(long)(k0 + j) * ld + r];
    for (int j = 0; j < nb; j++) {
        double t = x[j];
        for (int k = 0; k < j; k++) t -= x[k] * L[j][k];
        x[j] = t / L[j][j];
    }
    for (int j = 0; j < nb; j++) A[(long)(k0 + j) * ld + r] = x[j];
}

// Trailing update A22 -= L21 L21^T (lower part), 16x16 output tiles.
__global__ __launch_bounds__(256) void potrf_update_kernel(double *A, int ld, int n, int k0, int nb, DevStatus *st)
{
    if (st->ierr) return;
    const int base = k0 + nb;
    const int ti = base + blockIdx.x * 16, tj = base + blockIdx.y * 16;
    if (tj > ti + 15) return;
    __shared__ double Li[16][CNB + 1], Lj[16][CNB + 1];
    for (int e = threadIdx.x; e < 16 * nb; e += blockDim.x) {
        int r = e % 16, k = e / 16;
        Li[r][k] = (ti + r < n) ? A[(long)(k0 + k) * ld + ti + r] : 0.0;
        Lj[r][k] = (tj + r < n) ? A[(long)(k0 + k) * ld + tj + r] : 0.0;
    }
    __syncthreads();
    int r = threadIdx.x % 16, cc = threadIdx.x / 16;
    int gi = ti + r, gj = tj + cc;
    if (gi >= n || gj >= n || gi < gj) return;
    double t = 0.0;
    for (int k = 0; k < nb; k++) t += Li[r][k] * Lj[cc][k];
    A[(long)gj * ld + gi] -= t;
}

// out = L^T x (reference U.xi) when trans, else L x.
__global__ void trmv_kernel(const double *L, int ld, int n, const double *x, double *out, int s, int ldx, int trans)
{
    int row = blockIdx.x;
    int c = blockIdx.y;
    if (row >= n) return;
    const double *xc = x + (long)c * ldx;
    double t = 0.0;
    if (trans)
        for (int k = row + threadIdx.x; k < n; k += 64) t += L[(long)row * ld + k] * xc[k];
    else
        for (int k = threadIdx.x; k <= row; k += 64) t += L[(long)k * ld + row] * xc[k];
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if (threadIdx.x == 0) out[(long)c * ldx + row] = t;
}

int brownian_cholesky(Ctx &c, const double *xi, double *out, int s, int method)
{
    const int n = c.n;
    if (!c.chol) {
        c.chol_ld = n;
        BD_HIP(hipMalloc(&c.chol, (size_t)n * n * sizeof(double)));
    }
    BD_TRY(mob_unpack(c, c.chol, c.chol_ld));
    hipStream_t sm = c.stream;
    const int ld = c.chol_ld;
    for (int k0 = 0; k0 < n; k0 += CNB) {
        int nb = std::min(CNB, n - k0);
        hipLaunchKernelGGL(potrf_diag_kernel, dim3(1), dim3(256), 0, sm, c.chol, ld, k0, nb, c.st);
        int rest = n - k0 - nb;
        if (rest > 0) {
            hipLaunchKernelGGL(potrf_panel_kernel, dim3((rest + 255) / 256), dim3(256), 0, sm, c.chol, ld, n, k0, nb,
                               c.st);
            int nt = (rest + 15) / 16;
            hipLaunchKernelGGL(potrf_update_kernel, dim3(nt, nt), dim3(256), 0, sm, c.chol, ld, n, k0, nb, c.st);
        }
    }
    BD_HIP(hipGetLastError());
    hipLaunchKernelGGL(trmv_kernel, dim3(n, s), dim3(64), 0, sm, c.chol, ld, n, xi, out, s, c.npad,
                       method == BD_CHOL ? 1 : 0);
    BD_HIP(hipGetLastError());
    BD_HIP(hipMemcpyAsync(c.st_host, c.st, sizeof(DevStatus), hipMemcpyDeviceToHost, sm));
    BD_HIP(hipStreamSynchronize(sm));
    if (c.st_host->ierr == 1) {
        set_error("bond length too large (overstretched bond)");
        return BD_ERR_BOND;
    }
    if (c.st_host->ierr) {
        set_error("Cholesky factorisation failed: mobility not positive definite");
        return BD_ERR_FACTOR;
    }
    return BD_OK;
}

}  // namespace bd
