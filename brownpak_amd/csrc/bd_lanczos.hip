// bd_lanczos.hip -- Brownian displacement B.xi on the device.
//
// Block Lanczos (brown_m.f90:292-512; s = 1 is the scalar Lanczos of
// brown_m.f90:185-288) and Cholesky (brown_m.f90:516-554).
//
// Per Krylov iteration j the device runs:
//   W = D V_j                                  (symmetric tile product)
//   W -= V_{j-1} B_{j-1}^T ; A_j = V_j^T W     (lanczos_a)
//   W -= V_j symL(A_j)                          (lanczos_b)
//   W = V_{j+1} B_j                             (CholeskyQR2, FP64; the second
//                                                 pass inside the next product)
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
// M 2^-k for which the Cholesky of H_j - m I succeeds.  Each node (and each
// probe of m) is one block-tridiagonal SPD solve with s x s blocks held in
// registers, run by its own thread, so a whole evaluation is one block of
// one kernel instead of a serial O(js^3) eigensolve.  In exact arithmetic
// the result equals the reference's P sqrt(Lambda) P^T E_1 R.
#include "bd_internal.h"

#include <thread>
#include "bd_lanczos_dev.h"
#include "bd_pcr_dev.h"

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
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += (int)gridDim.x * RB) {
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
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += (int)gridDim.x * RB) {
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
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += (int)gridDim.x * RB) {
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
                                                    const double *__restrict__ rprev, DevStatus *st, int blk)
{
    __shared__ double sG[S * S];
    __shared__ int s_done;
    // this thread's first row, loaded before the Gram sum so that the two
    // memory round trips overlap, and the status with them
    if (threadIdx.x == 0) s_done = lz_done(st) ? 1 : 0;
    const int r0 = blockIdx.x * RB + threadIdx.x;
    double x0[S];
#pragma unroll
    for (int k = 0; k < S; k++) x0[k] = X[(long)k * ld + min(r0, rows - 1)];
    sum_partials<S>(gpart, sG);   // ends with a barrier
    if (s_done) return;
    double R[S][S];
    bool ok = small_chol<S>(sG, R);
    if (!ok) {
        if (blockIdx.x == 0 && threadIdx.x == 0) lz_factor_failed(st, blk);
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
    for (int r = r0; r < rows; r += (int)gridDim.x * RB) {
        double x[S], q[S];
#pragma unroll
        for (int k = 0; k < S; k++) x[k] = (r == r0) ? x0[k] : X[(long)k * ld + r];
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

// ------------------------------------------------------------ fused step
// The Krylov step of one iteration in one workgroup, for n <= KT * KR rows:
//   W -= V_{j-1} B_{j-1}^T ; A_j = V_j^T W ; W -= V_j symL(A_j) ;
//   CholeskyQR2: W = Q R1, Q = V_{j+1} R2, B_j = R2 R1.
// The rows stay in registers across the three block reductions, so the
// four multi-block kernels (and their launch gaps) become one launch.  A
// thread owns KR consecutive rows, so one base address per column serves
// all of them (immediate offsets) and the registers go to the rows.
constexpr int KT = 1024;   // threads of the fused step
constexpr int KR = 6;      // rows per thread (n <= 6144)

template <int S>
__device__ inline void kblock_sum(double (&acc)[S * S], double (*sh)[S * S], double *out)
{
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
    for (int e = 0; e < S * S; e++) {
        double v = acc[e];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sh[wv][e] = v;
    }
    __syncthreads();
    if (threadIdx.x < S * S) {
        double v = 0.0;
        for (int w = 0; w < KT / 64; w++) v += sh[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
    __syncthreads();
}

template <int S>
__global__ __launch_bounds__(KT) void krylov_fused_kernel(const double *__restrict__ W, const double *__restrict__ Vj,
                                                          const double *__restrict__ Vp, const double *__restrict__ Bp,
                                                          double *__restrict__ Vn, int ld, int rows,
                                                          double *__restrict__ Aout, double *__restrict__ Bout,
                                                          int upd, DevStatus *st, int blk)
{
    // the S x S matrices live in LDS and are read as broadcasts, which keeps
    // the registers for the rows
    __shared__ int s_done, s_fail;
    __shared__ double sh[KT / 64][S * S];
    __shared__ double sM[S * S], sB[S * S], sR1[S * S], sR2[S * S], sI1[S], sI2[S];
    if (threadIdx.x == 0) {
        s_done = lz_done(st) ? 1 : 0;
        s_fail = 0;
    }
    if (threadIdx.x < S * S) sB[threadIdx.x] = Vp ? Bp[threadIdx.x] : 0.0;
    __syncthreads();
    if (s_done) return;
    double w[KR][S];
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    // W -= V_{j-1} B^T ; A = V_j^T W.  Loads come from clamped in-range rows,
    // so they carry no branch; rows past the end are zeroed afterwards.
#pragma unroll
    for (int m = 0; m < KR; m++) {
        const int r = min((int)threadIdx.x * KR + m, rows - 1);
        const bool in = (int)threadIdx.x * KR + m < rows;
        double v[S], p[S];
#pragma unroll
        for (int k = 0; k < S; k++) {
            w[m][k] = W[(long)k * ld + r];
            v[k] = Vj[(long)k * ld + r];
            p[k] = Vp ? Vp[(long)k * ld + r] : 0.0;
        }
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < S; k++) t += p[k] * sB[c + k * S];
            w[m][c] = in ? w[m][c] - t : 0.0;
        }
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += (in ? v[ci] : 0.0) * w[m][cj];
        // two rows of loads in flight per thread; hoisting all would spill
        if (m % 2 == 1) __builtin_amdgcn_sched_barrier(0);
    }
    kblock_sum<S>(acc, sh, sM);
    if (threadIdx.x < S * S) Aout[threadIdx.x] = sM[threadIdx.x];
    if (!upd) return;
    // W -= V_j symL(A) ; G1 = W^T W
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
#pragma unroll
    for (int m = 0; m < KR; m++) {
        const int r = min((int)threadIdx.x * KR + m, rows - 1);
        const bool in = (int)threadIdx.x * KR + m < rows;
        double v[S];
#pragma unroll
        for (int k = 0; k < S; k++) v[k] = Vj[(long)k * ld + r];
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < S; k++) t += v[k] * ((k >= c) ? sM[k + c * S] : sM[c + k * S]);
            w[m][c] = in ? w[m][c] - t : 0.0;
        }
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += w[m][ci] * w[m][cj];
    }
    // CholeskyQR2 with the factors in LDS
    kblock_sum<S>(acc, sh, sM);
    if (threadIdx.x == 0) {
        double R[S][S];
        if (!small_chol<S>(sM, R)) s_fail = 1;
        for (int i = 0; i < S; i++) {
            sI1[i] = 1.0 / R[i][i];
            for (int j = 0; j < S; j++) sR1[i + j * S] = R[i][j];
        }
    }
    __syncthreads();
    if (s_fail) {
        if (threadIdx.x == 0) lz_factor_failed(st, blk);
        return;
    }
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
#pragma unroll
    for (int m = 0; m < KR; m++) {
        double q[S];
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = w[m][c];
#pragma unroll
            for (int k = 0; k < c; k++) t -= q[k] * sR1[k + c * S];
            q[c] = t * sI1[c];
        }
#pragma unroll
        for (int k = 0; k < S; k++) w[m][k] = q[k];
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += q[ci] * q[cj];
    }
    kblock_sum<S>(acc, sh, sM);
    if (threadIdx.x == 0) {
        double R[S][S];
        if (!small_chol<S>(sM, R)) s_fail = 1;
        for (int i = 0; i < S; i++) {
            sI2[i] = 1.0 / R[i][i];
            for (int j = 0; j < S; j++) sR2[i + j * S] = R[i][j];
        }
        // B_j = R2 R1
        for (int i = 0; i < S; i++)
            for (int j = 0; j < S; j++) {
                double t = 0.0;
                for (int k = 0; k < S; k++) t += R[i][k] * sR1[k + j * S];
                Bout[i + j * S] = t;
            }
    }
    __syncthreads();
    if (s_fail) {
        if (threadIdx.x == 0) lz_factor_failed(st, blk);
        return;
    }
#pragma unroll
    for (int m = 0; m < KR; m++) {
        const int r = (int)threadIdx.x * KR + m;
        double q[S];
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = w[m][c];
#pragma unroll
            for (int k = 0; k < c; k++) t -= q[k] * sR2[k + c * S];
            q[c] = t * sI2[c];
        }
        if (r < rows) {
#pragma unroll
            for (int k = 0; k < S; k++) Vn[(long)k * ld + r] = q[k];
        }
    }
}

// ------------------------------------------------------------ two-pass step
// After the product's epilogue (W1 = W - V_{j-1} B^T and the partials of
// A = V_j^T W1, P = W1^T W1), one kernel finishes A, forms the Gram matrix of
// W2 = W1 - V_j symL(A) without another pass over the rows,
//   W2^T W2 = P - A^T Ah - Ah^T A + Ah^T Ah     (V_j^T V_j = I, Ah = symL(A)),
// factors it (R1) and writes Q = W2 R1^{-1} with the partial Gram matrix of
// Q.  The second CholeskyQR pass (V_{j+1} = Q R2^{-1}, B_j = R2 R1) is done
// by the next iteration's product as it stages Q (XPre), which saves a
// launch per iteration; cholqr_kernel does it where no product follows.
template <int S>
__global__ __launch_bounds__(RB) void krylov_k2_kernel(const double *__restrict__ W, const double *__restrict__ Vj,
                                                       double *__restrict__ Q, int ld, int rows,
                                                       const double *__restrict__ lzpart, int nblk,
                                                       double *__restrict__ Aout, double *__restrict__ R1save,
                                                       int upd, double *__restrict__ gpart, DevStatus *st,
                                                       float *__restrict__ qrec32, double *__restrict__ qrec64, int blk)
{
    __shared__ int s_done, s_fail;
    __shared__ double sAP[2 * S * S], sAh[S * S], sG[S * S], sR[S * S], sI[S], s_scr[RB];
    if (threadIdx.x == 0) {
        s_done = lz_done(st) ? 1 : 0;
        s_fail = 0;
    }
    // this thread's first row, loaded before the partial sums so that the
    // two memory round trips overlap
    const int r0 = blockIdx.x * RB + threadIdx.x, rr0 = min(r0, rows - 1);
    double w0[S], v0[S];
#pragma unroll
    for (int k = 0; k < S; k++) {
        w0[k] = W[(long)k * ld + rr0];
        v0[k] = Vj[(long)k * ld + rr0];
    }
    // fixed-order sums of the per-block partials (ends with a barrier)
    block_sum_parts(lzpart, 2 * S * S, nblk, s_scr, sAP);
    if (s_done) return;
    if (blockIdx.x == 0 && threadIdx.x < S * S) Aout[threadIdx.x] = sAP[threadIdx.x];
    if (!upd) return;
    const double *A = sAP, *P = sAP + S * S;
    if (threadIdx.x < S * S) {
        const int a = threadIdx.x % S, b = threadIdx.x / S;
        sAh[a + b * S] = (a >= b) ? A[a + b * S] : A[b + a * S];
    }
    __syncthreads();
    if (threadIdx.x < S * S) {
        // G = W2^T W2 entry (a, b), symmetrised (the three products are
        // symmetric only up to rounding), one entry per thread
        const int a = threadIdx.x % S, b = threadIdx.x / S;
        auto gab = [&](int x, int y) {
            double t = P[x + y * S];
            for (int k = 0; k < S; k++) {
                t -= A[k + x * S] * sAh[k + y * S];     // (A^T Ah)(x,y)
                t -= sAh[k + x * S] * A[k + y * S];     // (Ah^T A)(x,y)
                t += sAh[k + x * S] * sAh[k + y * S];   // (Ah^T Ah)(x,y)
            }
            return t;
        };
        if (a == b) sG[a + b * S] = gab(a, a);
        else if (a > b) sG[a + b * S] = sG[b + a * S] = 0.5 * (gab(a, b) + gab(b, a));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double R[S][S];
        if (!small_chol<S>(sG, R)) s_fail = 1;
        for (int i = 0; i < S; i++) {
            sI[i] = 1.0 / R[i][i];
            for (int j = 0; j < S; j++) sR[i + j * S] = R[i][j];
        }
    }
    __syncthreads();
    if (s_fail) {
        if (blockIdx.x == 0 && threadIdx.x == 0) lz_factor_failed(st, blk);
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x < S * S) R1save[threadIdx.x] = sR[threadIdx.x];
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    for (int r = r0; r < rows; r += (int)gridDim.x * RB) {
        double w[S], v[S], q[S];
#pragma unroll
        for (int k = 0; k < S; k++) {
            w[k] = (r == r0) ? w0[k] : W[(long)k * ld + r];
            v[k] = (r == r0) ? v0[k] : Vj[(long)k * ld + r];
        }
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < S; k++) t += v[k] * sAh[k + c * S];
            w[c] = w[c] - t;
        }
#pragma unroll
        for (int c = 0; c < S; c++) {
            double t = w[c];
#pragma unroll
            for (int k = 0; k < c; k++) t -= q[k] * sR[k + c * S];
            q[c] = t * sI[c];
        }
#pragma unroll
        for (int k = 0; k < S; k++) Q[(long)k * ld + r] = q[k];
        // row-major copy of Q for a matrix-free product (bd_matfree.hip)
        if (qrec32) {
#pragma unroll
            for (int k = 0; k < S; k++) qrec32[(long)r * S + k] = (float)q[k];
        } else if (qrec64) {
#pragma unroll
            for (int k = 0; k < S; k++) qrec64[(long)r * S + k] = q[k];
        }
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += q[ci] * q[cj];
    }
    block_reduce_store<S>(acc, gpart + blockIdx.x * S * S);
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
    if (lz_done(st)) return;
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
    for (int r = blockIdx.x * RB + threadIdx.x; r < rows; r += (int)gridDim.x * RB) {
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
// The small problem of iteration j in one block (sp::small_problem):
// spectral bounds, quadrature, and the coefficients C = sqrt(H_j) E_1 R_0
// of y_j = V_{1:j} C.
template <int S>
__global__ __launch_bounds__(256) void small_kernel(const double *__restrict__ Ablk, const double *__restrict__ Bblk,
                                                    const double *__restrict__ R0, int nb, double *__restrict__ scr,
                                                    double *__restrict__ Cout, DevStatus *st)
{
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = (lz_tests_done(st) || !lz_valid(st, nb)) ? 1 : 0;
    __syncthreads();
    if (s_done) return;
    const sp::SmallOut o = sp::small_problem<S>(Ablk, Bblk, R0, nb, NBOUND, 13, scr, 256, Cout);
    if (threadIdx.x == 0) {
        st->lam_hi = o.lam_hi;
        st->lam_lo = o.lam_lo;
        st->nq = o.nq;
        if (o.err) atomicExch(&st->ierr, 2);
    }
}

// The small problem split over workgroups (sp::small_problem_split).
template <int S>
__global__ __launch_bounds__(sp::WG_T) void small_split_kernel(const double *__restrict__ Ablk,
                                                               const double *__restrict__ Bblk,
                                                               const double *__restrict__ R0, int nb,
                                                               double *__restrict__ part, double *__restrict__ Cout,
                                                               sp::SplitState *ss, DevStatus *st)
{
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = (lz_tests_done(st) || !lz_valid(st, nb)) ? 1 : 0;
    __syncthreads();
    if (s_done) return;
    const bool fin = sp::small_problem_split<S>(Ablk, Bblk, R0, nb, NBOUND, 13, part, Cout, ss, ss, 1, nb & 1);
    if (fin && threadIdx.x == 0) {
        st->lam_hi = ss->lam_hi;
        st->lam_lo = ss->lam_lo;
        st->nq = ss->nq;
        if (ss->err) atomicExch(&st->ierr, 2);
    }
}

constexpr int YR = 32, YG = 8, YC = 32;   // combine: rows per block, column groups, columns per group
constexpr int YJS = 256;                   // largest j s of the grouped combination (64 blocks of 4)

// ------------------------------------------------------------ grouped tests
// The small problems of LG consecutive iterations j0 .. j0+LG-1 run as one
// launch (blockIdx.y = member, each split over workgroups), and one combine
// launch forms their y_j and applies the reference's test to each in order.
// Both run on the side stream behind one event per group, so the Krylov
// recurrence never waits for a test; the host learns the verdict through a
// flag in mapped host memory (report_tests) and stops issuing iterations
// once it has read a converged one.  Groups hold up to LG consecutive
// iterations; parity (the group's count mod 2) selects which of the two
// node-placement hint slots a group reads and which it writes.
constexpr int LG = 4;

template <int S>
__global__ __launch_bounds__(sp::WG_T) void small_group_kernel(const double *__restrict__ Ablk,
                                                               const double *__restrict__ Bblk,
                                                               const double *__restrict__ R0, int j0, int jend,
                                                               double *__restrict__ part, long part_stride,
                                                               double *__restrict__ Cg, long c_stride,
                                                               sp::SplitState *ssv, int parity, DevStatus *st)
{
    const int j = j0 + (int)blockIdx.y;
    __shared__ int s_done;
    if (threadIdx.x == 0) s_done = lz_tests_done(st) ? 1 : 0;
    __syncthreads();
    if (s_done || j > jend || j < 2) return;
    sp::SplitState *ss = ssv + blockIdx.y;
    const bool fin = sp::small_problem_split<S>(Ablk, Bblk, R0, j, NBOUND, 13, part + blockIdx.y * part_stride,
                                                Cg + blockIdx.y * c_stride, ss, ssv, LG, parity);
    if (fin && threadIdx.x == 0) {
        st->lam_hi = ss->lam_hi;
        st->lam_lo = ss->lam_lo;
        st->nq = ss->nq;
        if (ss->err && lz_valid(st, j)) atomicExch(&st->ierr, 2);
    }
}

// The small problems of iterations j0..jend (blockIdx.y = j - j0) by
// parallel cyclic reduction (bd_pcr_dev.h), for nb = j <= 64 blocks.  Every
// workgroup runs one probe wave (H - m I positive definite?) and three
// quadrature-node waves, one wave per SIMD.  Bound and nodes come from the
// hint, the largest probe index the previous launch accepted: the smallest
// Ritz value only decreases with j, and a node range that reaches below the
// spectrum is as accurate as a tight one.  If the hinted bound fails, all
// four waves probe the next halvings together and the nodes are solved again
// with the first bound that passes; every workgroup reaches the same
// decision.  Each workgroup writes the weighted sum of its nodes' Z in
// part[wg]; the last one sums the parts in workgroup order and forms C = H Z.
// Grid: x = member, y = node batch, so that the batches that have nodes
// (the low ones) of every member are dispatched before the idle ones; the
// host sizes y from the node count the previous group reported.
template <int S>
__global__ __launch_bounds__(pcr::THREADS) void small_pcr_kernel(const double *__restrict__ Ablk,
                                                                 const double *__restrict__ Bblk,
                                                                 const double *__restrict__ R0, int j0, int jend,
                                                                 double *__restrict__ part, long part_stride,
                                                                 double *__restrict__ Cg, long c_stride,
                                                                 sp::SplitState *ssv, int parity, DevStatus *st)
{
    using namespace pcr;
    constexpr int SS = S * S;
    constexpr int BUF = (PROBES * 2 + NODES * 3) * SS * ROWS;   // task slots (doubles)
    __shared__ double s_buf[BUF];
    __shared__ double sA[ROWS * SS], sB[ROWS * SS], sR[SS];
    __shared__ double s_red[WAVES];
    __shared__ int s_done, s_ok[WAVES], s_last;
    const int j = j0 + (int)blockIdx.x;
    // the blocks of H_j are loaded together with the status
    if (threadIdx.x == 0) s_done = lz_tests_done(st) ? 1 : 0;
    const int nbl = (j >= 2 && j <= jend) ? j : 0;
    for (int e = threadIdx.x; e < nbl * SS; e += THREADS) {
        sA[e] = Ablk[e];
        sB[e] = Bblk[e];
    }
    if (threadIdx.x < SS) sR[threadIdx.x] = R0[threadIdx.x];
    __syncthreads();
    if (s_done || j > jend || j < 2) return;
    const int nb = j, js = nb * S, t = threadIdx.x, w = t / 64, lane = t % 64;
    sp::SplitState *ss = ssv + blockIdx.x;
    double *P = part + blockIdx.x * part_stride;
    const double M = sp::gersh_block<S>(sA, sB, js, s_red);
    int hint = -1;
    for (int g = 0; g < LG; g++) hint = max(hint, ssv[g].kb[parity ^ 1]);
    const int kb_guess = hint >= 0 ? min(hint, NBOUND - 1) : 0;
    constexpr int DEC_MAX = 13;
    auto place = [&](int kb, double &m, int &N, double &h) -> int {
        m = ldexp(M, -(kb + 1));
        int dec = (int)ceil(log10(M / m));
        if (dec < 1) dec = 1;
        if (dec > DEC_MAX) dec = DEC_MAX + 1;
        N = c_qN[min(dec, 13)];
        h = c_qh[min(dec, 13)];
        return dec;
    };
    // node batches (NODES nodes each) of the hinted placement; workgroups
    // without one leave at once
    int nact_g = 1;
    {
        double mg, hg;
        int Ng;
        if (place(kb_guess, mg, Ng, hg) <= DEC_MAX) nact_g = (2 * Ng + 1 + NODES - 1) / NODES;
    }
    // the grid may hold fewer batches than the placement (sized from the
    // previous group's node count): then every workgroup takes several
    nact_g = min(nact_g, (int)gridDim.y);
    if ((int)blockIdx.y >= nact_g) return;
    // this wave's slots
    double *buf = s_buf + (w < PROBES ? w * 2 * SS * ROWS : (PROBES * 2 + (w - PROBES) * 3) * SS * ROWS);
    double Z[S][S];
    double wt = 0.0;
    bool node_on = false;
    // one round: probes at kb_p (probe waves) and the nodes of batch `batch`
    // of placement kb_n
    auto round = [&](int kbp0, int kbp_step, bool nodes, int kb_n, int batch) __attribute__((always_inline)) {
        double shift = 1.0;
        bool rhs = false, active = false;
        if (w < PROBES || !nodes) {
            const int kbp = kbp0 + w * kbp_step;
            active = kbp >= 0 && kbp < NBOUND;
            shift = -ldexp(M, -(kbp + 1));
        } else {
            double m, h;
            int N;
            const int dec = place(kb_n, m, N, h);
            const int node = batch * NODES + (w - PROBES);
            if (dec <= DEC_MAX && node < 2 * N + 1) {
                const double u0 = 0.25 * (log(m) + log(M));
                const double v = (node - N) * h;
                const double u = u0 + sinh(v);
                wt = (2.0 / M_PI) * h * cosh(v) * exp(u);
                shift = exp(2.0 * u);
                rhs = active = true;
            }
            node_on = active;
        }
        const bool ok = reduce<S>(sA, sB, sR, nb, shift, rhs, buf, Z);
        if (lane == 0) s_ok[w] = (active && ok) ? 1 : 0;
        __syncthreads();
    };
    round(kb_guess, 1, true, kb_guess, (int)blockIdx.y);
    int use = kb_guess, kb_found = -1;
    if (s_ok[0]) {
        kb_found = kb_guess;
    } else {
        // the hinted bound is above the spectrum: four halvings per round
        for (int kb0 = kb_guess + 1; kb0 < NBOUND && kb_found < 0; kb0 += WAVES) {
            round(kb0, 1, false, 0, 0);
            for (int q = 0; q < WAVES; q++)
                if (s_ok[q]) { kb_found = kb0 + q; break; }
        }
        use = kb_found;
    }
    int err = 0;
    double m = 0.0, h = 0.0;
    int N = 0;
    if (kb_found < 0) err = 2;
    else if (place(use, m, N, h) > DEC_MAX) err = 3;
    const int nq = err ? 1 : 2 * N + 1;
    const int nact = err ? 1 : (nq + NODES - 1) / NODES;
    // node batches blockIdx.y, blockIdx.y + nact_g, ... of the placement in
    // use (one, unless the fallback placed more nodes than the hint did)
    if (t == 0) s_last = 0;
    __syncthreads();
    for (int b = (int)blockIdx.y; b < nact; b += nact_g) {
        if (!err && !(b == (int)blockIdx.y && use == kb_guess)) round(-1, 0, true, use, b);   // probe wave idle
        // weighted node solutions summed in wave order into part[b]
        __syncthreads();
        if (!err && w >= PROBES && lane < nb) {
            double *zw = s_buf + (size_t)(w - PROBES) * js * S;
#pragma unroll
            for (int a = 0; a < S; a++)
#pragma unroll
                for (int c = 0; c < S; c++) zw[(lane * S + a) + c * js] = node_on ? wt * Z[a][c] : 0.0;
        }
        __syncthreads();
        double *pw = P + (long)b * js * S;
        for (int e = t; e < js * S; e += THREADS) {
            double acc = 0.0;
            for (int q = 0; q < NODES; q++) acc += err ? 0.0 : s_buf[(size_t)q * js * S + e];
            store_dev(pw + e, acc);
        }
        stores_complete();
        __syncthreads();
        // the batch completing the count is this workgroup's last one
        if (t == 0 && atomicAdd(&ss->cnt, 1) == nact - 1) s_last = 1;
        __syncthreads();
    }
    if (!s_last) return;
    if (t == 0) {
        ss->cnt = 0;
        ss->kb[parity] = kb_found;
        ss->nq = nq;
        ss->err = err;
        ss->lam_lo = m;
        ss->lam_hi = M;
        st->lam_hi = M;
        st->lam_lo = m;
        st->nq = nq;
        if (err && lz_valid(st, j)) atomicExch(&st->ierr, 2);
    }
    if (err) return;
    double *sZ = s_buf;
    for (int e = t; e < js * S; e += THREADS) sZ[e] = sum_strided<true>(P + e, (long)js * S, nact);
    __syncthreads();
    double *Cout = Cg + blockIdx.x * c_stride;
    for (int e = t; e < js * S; e += THREADS) {
        const int i = e % js, c = e / js;
        Cout[e] = sp::block_row_dot<S>(sA, sB, nb, i, sZ + (long)c * js);
    }
}

// The host's view of a group of tests: converged_j, ierr and the small
// problem's node count, then the last iteration the group covered; the host
// waits for the latter before it reads the former.  The host reads nothing else the device wrote, so the
// three system-scope stores with a wait for completion before the last one
// are the whole protocol (no cache write-back).
__device__ inline void report_tests(int *flag, int jend, int conv, int ierr, int nq)
{
    __hip_atomic_store(flag + 1, conv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(flag + 2, ierr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(flag + 3, nq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    stores_complete();
    __hip_atomic_store(flag, jend, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// y_j = V(:, 1:jS) C_j for the group's members (Y is a ring of LG + 1
// vectors indexed by j mod (LG + 1)); partial norms of y_j - y_{j-1} and
// y_{j-1}; the last block tests the members in order and records the first
// j that passes.
template <int S>
__global__ __launch_bounds__(256) void ycomb_group_kernel(const double *__restrict__ V, int ld, int rows, int j0,
                                                          int jend, const double *__restrict__ Cg, long c_stride,
                                                          double *__restrict__ Y, int two_norm,
                                                          double *__restrict__ part, double tol, DevStatus *st,
                                                          int *flag)
{
    constexpr int NR = LG + 1;
    __shared__ int s_done, s_last;
    __shared__ double s_acc[YG][LG][S][YR];   // row index fastest: conflict-free
    __shared__ double s_nrm[LG][2 * S];
    __shared__ double s_C[LG][YJS * S];
    __shared__ double s_scr[256];
    __shared__ double s_np[LG * 2 * S][YR + 1];
    __shared__ double s_np2[LG * 2 * S][256 / (LG * 2 * S)];
    const int ng = jend - j0 + 1;               // members (<= LG)
    const int jsmax = jend * S;
    if (threadIdx.x == 0) s_done = lz_tests_done(st) ? 1 : 0;
    // this thread's entries of V (one column range of at most YC columns:
    // jsmax <= YJS <= YG * YC) and of y_{j0-1}, loaded before the
    // coefficients so that the memory round trips overlap
    static_assert(YG * YC >= YJS, "one pass over the column range");
    const int lr = threadIdx.x % YR, gq = threadIdx.x / YR;
    const int r = blockIdx.x * YR + lr;
    const int rr = min(r, rows - 1);
    const int per = (jsmax + YG - 1) / YG, m0 = gq * per, m1 = min(jsmax, m0 + per);
    double vm[YC];
#pragma unroll
    for (int q = 0; q < YC; q++) vm[q] = (m0 + q < m1) ? V[(long)(m0 + q) * ld + rr] : 0.0;
    double prev[S];
#pragma unroll
    for (int k = 0; k < S; k++)
        prev[k] = (threadIdx.x < YR && r < rows && j0 >= 3) ? Y[((long)((j0 - 1) % NR) * S + k) * ld + r] : 0.0;
    {
        // the members' coefficients (row-major into s_C), all loads of a
        // thread issued before the first store
        int off[LG + 1];
        off[0] = 0;
#pragma unroll
        for (int g = 0; g < LG; g++) {
            const int jg = j0 + g;
            off[g + 1] = off[g] + ((g < ng && jg >= 2) ? jg * S * S : 0);
        }
        const int tot = off[LG];
        constexpr int U = 16;
        auto member = [&](int x) __attribute__((always_inline)) {
            int g = 0;
#pragma unroll
            for (int k = 1; k < LG; k++) g += (x >= off[k]) ? 1 : 0;
            return g;
        };
        for (int b0 = threadIdx.x; b0 < tot; b0 += U * (int)blockDim.x) {
            double v[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int x = min(b0 + q * (int)blockDim.x, tot - 1);
                const int g = member(x);
                v[q] = Cg[g * c_stride + (x - off[g])];
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int x = b0 + q * (int)blockDim.x;
                if (x < tot) {
                    const int g = member(x), e = x - off[g], js = (j0 + g) * S;
                    s_C[g][(e % js) * S + e / js] = v[q];
                }
            }
        }
    }
    __syncthreads();
    if (s_done) {
        // nothing to test: report the state reached, as the last block would
        if (blockIdx.x == 0 && threadIdx.x == 0) report_tests(flag, jend, st->converged_j, st->ierr, st->nq);
        return;
    }
    double acc[LG][S];
#pragma unroll
    for (int g = 0; g < LG; g++)
#pragma unroll
        for (int k = 0; k < S; k++) acc[g][k] = 0.0;
#pragma unroll
    for (int q = 0; q < YC; q++) {
        const int m = m0 + q;
        if (m < m1) {
#pragma unroll
            for (int g = 0; g < LG; g++)
                if (g < ng && j0 + g >= 2 && m < (j0 + g) * S) {
#pragma unroll
                    for (int k = 0; k < S; k++) acc[g][k] += vm[q] * s_C[g][m * S + k];
                }
        }
    }
#pragma unroll
    for (int g = 0; g < LG; g++)
#pragma unroll
        for (int k = 0; k < S; k++) s_acc[gq][g][k][lr] = acc[g][k];
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        double dn[LG][S], on[LG][S];
#pragma unroll
        for (int g = 0; g < LG; g++) {
            const int j = j0 + g;
#pragma unroll
            for (int k = 0; k < S; k++) {
                dn[g][k] = 0.0;
                on[g][k] = 0.0;
            }
            if (g < ng && j >= 2 && lane < YR && r < rows) {
#pragma unroll
                for (int k = 0; k < S; k++) {
                    double v = 0.0;
#pragma unroll
                    for (int q = 0; q < YG; q++) v += s_acc[q][g][k][lane];
                    Y[((long)(j % NR) * S + k) * ld + r] = v;
                    if (j > 2) {
                        const double o = prev[k], d = v - o;
                        if (two_norm) { dn[g][k] = d * d; on[g][k] = o * o; }
                        else { dn[g][k] = fabs(d); on[g][k] = fabs(o); }
                    }
                    prev[k] = v;
                }
            }
        }
        if (lane < YR) {
#pragma unroll
            for (int g = 0; g < LG; g++)
#pragma unroll
                for (int e = 0; e < 2 * S; e++) s_np[g * 2 * S + e][lane] = e < S ? dn[g][e] : on[g][e - S];
        }
    }
    __syncthreads();
    {
        // the block's partial norms: the YR rows summed in CHK chunks, then
        // the chunks (fixed order)
        constexpr int NV = LG * 2 * S, CHK = 256 / NV, RPC = (YR + CHK - 1) / CHK;
        const int t = threadIdx.x;
        if (t < NV * CHK) {
            const int v = t / CHK, c = t % CHK;
            double a = 0.0;
#pragma unroll
            for (int q = 0; q < RPC; q++)
                if (c * RPC + q < YR) a += s_np[v][c * RPC + q];
            s_np2[v][c] = a;
        }
        __syncthreads();
        if (t < NV) {
            double a = 0.0;
#pragma unroll
            for (int c = 0; c < CHK; c++) a += s_np2[t][c];
            store_dev(part + (long)blockIdx.x * NV + t, a);
        }
    }
    stores_complete();
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(&st->ycnt, 1) == (int)gridDim.x - 1) ? 1 : 0;
    __syncthreads();
    if (!s_last) return;
    // part[b][g][e] summed over the blocks b (ends with a barrier)
    block_sum_parts<true>(part, LG * 2 * S, gridDim.x, s_scr, &s_nrm[0][0]);
    if (threadIdx.x != 0) return;
    st->ycnt = 0;
    st->last_j = jend;
    for (int g = 0; g < ng; g++) {
        const int j = j0 + g;
        if (!lz_valid(st, j)) {
            // the recurrence could not form V_j: no converged iteration before it
            st->ierr = 2;
            break;
        }
        if (j <= 2) continue;
        double num = 0.0, den = 0.0;
        if (two_norm) {
            for (int k = 0; k < S; k++) { num += s_nrm[g][k]; den += s_nrm[g][S + k]; }
            num = sqrt(num);
            den = sqrt(den);
        } else {
            for (int k = 0; k < S; k++) { num = fmax(num, s_nrm[g][k]); den = fmax(den, s_nrm[g][S + k]); }
        }
        const double err = num / den;
        st->err = err;
        if (err <= tol) {
            st->converged_j = j;
            break;
        }
    }
    report_tests(flag, jend, st->converged_j, st->ierr, st->nq);
}

// y = V(:, 1:js) C, 32 rows per block and the js columns split over eight
// column groups (up to 32 columns each, all loads issued together); then
// the partial norms of (y - yold) and yold; the last block to finish applies
// the reference's test (brown_m.f90:497-506) and records j.
template <int S>
__global__ __launch_bounds__(256) void ycomb_kernel(const double *__restrict__ V, int ld, int rows, int js,
                                                    const double *__restrict__ C, double *__restrict__ y,
                                                    const double *__restrict__ yold, int want_err, int two_norm,
                                                    double *__restrict__ part, int j, double tol, DevStatus *st)
{
    __shared__ int s_done, s_last;
    __shared__ double s_acc[YG][S][YR];   // row index fastest: conflict-free
    __shared__ double s_nrm[2 * S];
    __shared__ double s_C[512 * S];
    __shared__ double s_scr[256];
    if (threadIdx.x == 0) s_done = lz_tests_done(st) ? 1 : 0;
    for (int b0 = threadIdx.x; b0 < js * S; b0 += 8 * (int)blockDim.x) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = C[min(b0 + q * (int)blockDim.x, js * S - 1)];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int e = b0 + q * (int)blockDim.x;
            if (e < js * S) s_C[(e % js) * S + e / js] = v[q];
        }
    }
    __syncthreads();
    if (s_done) return;
    const int lr = threadIdx.x % YR, g = threadIdx.x / YR;
    const int r = blockIdx.x * YR + lr;
    const int rr = min(r, rows - 1);   // clamped: loads stay unconditional
    const int per = (js + YG - 1) / YG, m0 = g * per, m1 = min(js, m0 + per);
    double acc[S];
#pragma unroll
    for (int k = 0; k < S; k++) acc[k] = 0.0;
    for (int c0 = m0; c0 < m1; c0 += YC) {   // one pass while js <= YG * YC
        double vm[YC];
#pragma unroll
        for (int q = 0; q < YC; q++) vm[q] = V[(long)min(c0 + q, js - 1) * ld + rr];
#pragma unroll
        for (int q = 0; q < YC; q++) {
            if (c0 + q < m1) {
#pragma unroll
                for (int k = 0; k < S; k++) acc[k] += vm[q] * s_C[(c0 + q) * S + k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < S; k++) s_acc[g][k][lr] = acc[k];
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        double dn[S], on[S];
#pragma unroll
        for (int k = 0; k < S; k++) { dn[k] = 0.0; on[k] = 0.0; }
        if (lane < YR && r < rows) {
#pragma unroll
            for (int k = 0; k < S; k++) {
                double v = 0.0;
#pragma unroll
                for (int gg = 0; gg < YG; gg++) v += s_acc[gg][k][lane];
                y[(long)k * ld + r] = v;
                if (want_err) {
                    const double o = yold[(long)k * ld + r], d = v - o;
                    if (two_norm) { dn[k] = d * d; on[k] = o * o; }
                    else { dn[k] = fabs(d); on[k] = fabs(o); }
                }
            }
        }
#pragma unroll
        for (int e = 0; e < 2 * S; e++) {
            double v = e < S ? dn[e] : on[e - S];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) store_dev(part + blockIdx.x * 2 * S + e, v);
        }
    }
    // last block: sum the partials in block order and test
    stores_complete();
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(&st->ycnt, 1) == (int)gridDim.x - 1) ? 1 : 0;
    __syncthreads();
    if (!s_last) return;
    block_sum_parts<true>(part, 2 * S, gridDim.x, s_scr, s_nrm);   // ends with a barrier
    if (threadIdx.x != 0) return;
    st->ycnt = 0;
    st->last_j = j;
    if (!lz_valid(st, j)) {
        st->ierr = 2;
        return;
    }
    if (!want_err) return;
    double num = 0.0, den = 0.0;
    if (two_norm) {
        for (int k = 0; k < S; k++) { num += s_nrm[k]; den += s_nrm[S + k]; }
        num = sqrt(num);
        den = sqrt(den);
    } else {
        for (int k = 0; k < S; k++) { num = fmax(num, s_nrm[k]); den = fmax(den, s_nrm[S + k]); }
    }
    const double err = num / den;
    st->err = err;
    if (err <= tol) st->converged_j = j;
}

__global__ void status_reset_kernel(DevStatus *st, sp::SplitState *ss)
{
    st->fail_j = 0;
    st->converged_j = 0;
    st->last_j = 0;
    st->err = 0.0;
    st->ycnt = 0;
    if (ss)
        for (int g = 0; g < LG; g++) ss[g].cnt = 0;
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
    BD_HIP(hipMalloc(&c.Y, (size_t)(LG + 1) * c.npad * s * sizeof(double)));
    BD_HIP(hipMemsetAsync(c.Y, 0, (size_t)(LG + 1) * c.npad * s * sizeof(double), c.stream));
    BD_HIP(hipMalloc(&c.Ablk, (size_t)s * s * c.dimk_max * sizeof(double)));
    BD_HIP(hipMalloc(&c.Bblk, (size_t)s * s * c.dimk_max * sizeof(double)));
    BD_HIP(hipMalloc(&c.R0, (size_t)s * s * 2 * sizeof(double)));
    BD_HIP(hipMalloc(&c.gpart, (size_t)2 * GBLK * MAX_S * MAX_S * sizeof(double)));
    if (ncol > 512) {
        set_error("lanc_mxitr * nts_mobsam too large (Krylov order <= 512)");
        return BD_ERR_ARG;
    }
    // small-problem scratch: per block of H, factors of 256 node shifts
    BD_HIP(hipMalloc(&c.Cq, (size_t)sp::scratch_doubles(s, c.dimk_max, 256) * sizeof(double)));
    BD_HIP(hipMalloc(&c.Ccoef, (size_t)LG * ncol * s * sizeof(double)));
    BD_HIP(hipMalloc(&c.bound_ok, NBOUND * sizeof(int)));
    BD_HIP(hipMalloc(&c.lzpart, (size_t)(c.npad / TILE + 1) * 2 * MAX_S * MAX_S * sizeof(double)));
    BD_HIP(hipMalloc(&c.Qbuf, (size_t)c.npad * MAX_S * sizeof(double)));
    BD_HIP(hipMalloc(&c.split, LG * sizeof(sp::SplitState)));
    {
        sp::SplitState z[LG];
        for (int g = 0; g < LG; g++) z[g] = sp::SplitState{0, {-1, -1}, 0, 0, 0.0, 0.0};
        BD_HIP(hipMemcpy(c.split, z, sizeof(z), hipMemcpyHostToDevice));
    }
    // combine partials: one set of 2 s norms per block, per group member
    BD_HIP(hipMalloc(&c.npart, (size_t)(c.npad / YR + 1) * LG * 2 * MAX_S * sizeof(double)));
    if (c.opt.lanc_reorth) {
        BD_HIP(hipMalloc(&c.Rq, (size_t)GBLK * ncol * s * sizeof(double)));
        BD_HIP(hipMalloc(&c.Rcoef, (size_t)ncol * s * sizeof(double)));
    }
    BD_HIP(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    for (int k = 0; k < LZ_NEV; k++) BD_HIP(hipEventCreateWithFlags(&c.lz_ev[k], kStreamEvent));
    for (int k = 0; k < 2; k++) BD_HIP(hipEventCreateWithFlags(&c.lz_chunk_ev[k], hipEventDisableTiming));
    void *fl = nullptr;
    BD_HIP(hipHostMalloc(&fl, 64, hipHostMallocMapped | hipHostMallocCoherent));
    c.lz_flag = static_cast<volatile int *>(fl);
    BD_HIP(hipHostGetDevicePointer((void **)&c.lz_flag_dev, fl, 0));
    return BD_OK;
}

void lanczos_free(Ctx &c)
{
    if (c.side) (void)hipStreamSynchronize(c.side);
    for (int k = 0; k < LZ_NEV; k++)
        if (c.lz_ev[k]) (void)hipEventDestroy(c.lz_ev[k]);
    for (int k = 0; k < 2; k++)
        if (c.lz_chunk_ev[k]) (void)hipEventDestroy(c.lz_chunk_ev[k]);
    if (c.lz_flag) (void)hipHostFree((void *)c.lz_flag);
    c.lz_flag = nullptr;
    c.lz_flag_dev = nullptr;
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
    if (c.split) (void)hipFree(c.split);
    c.split = nullptr;
    if (c.lzpart) (void)hipFree(c.lzpart);
    c.lzpart = nullptr;
    if (c.Qbuf) (void)hipFree(c.Qbuf);
    c.Qbuf = nullptr;
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
// Waits until the grouped tests through iteration jend have reported (see
// report_tests).  Polls the flag; after a while it waits for the tests'
// stream instead, after which the flag must be there.
static int wait_tests(Ctx &c, hipStream_t sm, int jend)
{
    for (long it = 0; it < 2000000; it++) {
        if (__atomic_load_n(&c.lz_flag[0], __ATOMIC_ACQUIRE) >= jend) return BD_OK;
        if (it > 1000) std::this_thread::yield();
    }
    BD_HIP(hipStreamSynchronize(sm));
    if (__atomic_load_n(&c.lz_flag[0], __ATOMIC_ACQUIRE) >= jend) return BD_OK;
    set_error("internal: Lanczos tests did not report");
    return BD_ERR_STATE;
}

template <int S>
static int lanczos_t(Ctx &c, const double *xi, double *out, int *lconv, int *iters)
{
    const int ld = c.npad, rows = c.npad;
    const int dmax = c.dimk_max;
    const bool two_norm = (S == 1);
    // one-workgroup Krylov step when the rows fit its registers
    const bool fused = S <= 2 && rows <= KT * KR && !c.opt.lanc_reorth;
    // otherwise the Krylov step rides on the product's reduction (no FP64
    // reorthogonalisation requested)
    const bool twopass = !fused && !c.opt.lanc_reorth;
    double *g0 = c.gpart, *g1 = c.gpart + GBLK * MAX_S * MAX_S;
    hipStream_t sm = c.stream, ss = c.side;
    // grid of every kernel that leaves or sums Gram partials (one partial
    // per block): a block per RB rows, at most GBLK
    const int gb = std::min(GBLK, (rows + RB - 1) / RB);

    hipLaunchKernelGGL(status_reset_kernel, dim3(1), dim3(1), 0, sm, c.st, c.split);
    // small problem split over workgroups when its blocks fit their LDS
    const bool split = S <= 4 && dmax <= sp::WG_NBM<S>;
    // by parallel cyclic reduction when every H_j has at most 64 blocks
    const bool use_pcr = S <= 4 && dmax <= pcr::ROWS;
    const int nwg = use_pcr ? (2 * 125 + 1 + pcr::NODES - 1) / pcr::NODES
                            : (2 * 125 + 1 + sp::WG_NPB<S> - 1) / sp::WG_NPB<S>;
    // grouped tests: LG iterations' small problems per launch, on the side
    // stream (needs the split or the reduction small problem)
    const bool grouped = split || use_pcr;
    // their stream: the side stream, except while the tile products are
    // being timed (events on the solve's stream would then also count the
    // wait for CUs held by the tests)
    const hipStream_t sg = c.prof_on ? sm : ss;
    {
        // host-side capacity checks of the scratch every launch below indexes
        const size_t ncol = (size_t)S * dmax;
        const size_t cq = (size_t)sp::scratch_doubles(S, dmax, 256);
        const size_t need_part = grouped ? (size_t)LG * nwg * dmax * S * S : cq;
        const size_t nblk_y = (size_t)(rows + YR - 1) / YR;
        if (need_part > cq || nblk_y * LG * 2 * S > (size_t)(c.npad / YR + 1) * LG * 2 * MAX_S ||
            (size_t)LG * dmax * S * S > (size_t)LG * ncol * S || (size_t)mob_rowblocks(c) > (size_t)(c.npad / TILE + 1) ||
            (grouped && (size_t)dmax * S > (size_t)YJS)) {
            set_error("internal: Lanczos scratch smaller than its launches need");
            return BD_ERR_STATE;
        }
    }
    // xi = V_1 R0 (CholeskyQR2)
    hipLaunchKernelGGL((gram_kernel<S>), dim3(gb), dim3(RB), 0, sm, xi, xi, ld, rows, g0, c.st);
    hipLaunchKernelGGL((cholqr_kernel<S>), dim3(gb), dim3(RB), 0, sm, xi, c.V, ld, rows, g0, g1,
                       c.R0 + S * S, (const double *)nullptr, c.st, 1);
    hipLaunchKernelGGL((cholqr_kernel<S>), dim3(gb), dim3(RB), 0, sm, c.V, c.V, ld, rows, g1,
                       (double *)nullptr, c.R0, (const double *)(c.R0 + S * S), c.st, 1);
    BD_HIP(hipGetLastError());

    const int chunk = LG;
    int pending = -1;   // slot of the status copy of the previous chunk
    int slot = 0;
    bool stop = false;
    int pcr_batches = 0;   // node batches of the last reported placement (0: unknown)
    // grouped tests: a group ends after LG iterations, and every iteration
    // from hint - 1 to hint + 1 is a group of its own (hint: the iteration at
    // which this context's previous solve converged), so that a solve that
    // converges there is seen without waiting for the rest of a group.  The
    // host takes a group's verdict once it is LG iterations old (waiting if
    // needed: the device then still has LG iterations queued) and reads newer
    // ones whenever they have already arrived.
    const int hint = c.lz_hint;
    auto in_window = [&](int jj) { return hint > 0 && jj >= hint - 1 && jj <= hint + 1; };
    int jg_start = 1;      // first iteration of the open group
    int ngroups = 0;       // groups launched (their parity alternates)
    int pend[64];          // last iterations of the groups whose verdict is not read yet
    int npend = 0;
    if (grouped) {
        c.lz_flag[0] = 0;
        c.lz_flag[1] = 0;
        c.lz_flag[2] = 0;
        c.lz_flag[3] = 0;
    }
    for (int j = 1; j <= dmax && !stop; j++) {
        double *Vj = c.V + (size_t)(j - 1) * S * ld;
        const double *Vp = (j > 1) ? c.V + (size_t)(j - 2) * S * ld : nullptr;
        const double *Bp = (j > 1) ? c.Bblk + (size_t)(j - 2) * S * S : nullptr;
        const int upd = (j < dmax) ? 1 : 0;
        if (twopass) {
            // product (which, from j = 2 on, first completes V_j from Q by
            // the second CholeskyQR pass and writes B_{j-1}) + epilogue, then
            // one kernel to A_j, R1 and Q of the next block
            LzEpi epi{Vj, Vp, Bp, c.lzpart};
            XPre pre;
            if (j > 1) {
                pre.blk = j;
                pre.gpart = g0;
                pre.ng = gb;
                pre.R1 = c.R0 + S * S;
                pre.Bout = c.Bblk + (size_t)(j - 2) * S * S;
                pre.V = Vj;
                pre.st = c.st;
            }
            BD_TRY(mob_product_lz(c, j > 1 ? c.Qbuf : Vj, c.W, S, ld, ld, &epi, j > 1 ? &pre : nullptr));
            hipLaunchKernelGGL((krylov_k2_kernel<S>), dim3(gb), dim3(RB), 0, sm, c.W, Vj, c.Qbuf, ld, rows,
                               c.lzpart, mob_rowblocks(c), c.Ablk + (size_t)(j - 1) * S * S, c.R0 + S * S, upd, g0,
                               c.st, (c.mf && c.opt.precision == BD_FP32) ? (float *)c.mf_x : nullptr,
                               (c.mf && c.opt.precision != BD_FP32) ? (double *)c.mf_x : nullptr, j + 1);
            BD_HIP(hipGetLastError());
        } else {
            BD_TRY(mob_product_lz(c, Vj, c.W, S, ld, ld));
        }
        if (twopass) {
        } else if (fused) {
            hipLaunchKernelGGL((krylov_fused_kernel<S>), dim3(1), dim3(KT), 0, sm, c.W, Vj, Vp, Bp,
                               c.V + (size_t)j * S * ld, ld, rows, c.Ablk + (size_t)(j - 1) * S * S,
                               c.Bblk + (size_t)(j - 1) * S * S, upd, c.st, j + 1);
            BD_HIP(hipGetLastError());
        } else {
            hipLaunchKernelGGL((lanczos_a_kernel<S>), dim3(gb), dim3(RB), 0, sm, c.W, Vj, Vp, Bp, ld, rows, g0,
                               c.st);
            hipLaunchKernelGGL((lanczos_b_kernel<S>), dim3(gb), dim3(RB), 0, sm, c.W, Vj, ld, rows, g0,
                               c.Ablk + (size_t)(j - 1) * S * S, upd, g1, c.st);
        }
        // A_j and V_1..V_j are final here: hand iteration j to the side stream
        hipEvent_t ev = c.lz_ev[j % LZ_NEV];
        if (!grouped) BD_HIP(hipEventRecord(ev, sm));
        if (!twopass && !fused && upd && c.opt.lanc_reorth) {
            // two classical Gram-Schmidt passes against V_1..V_j (FP64)
            const int ncol = j * S;
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL((reorth_coef_kernel<S>), dim3(GBLK), dim3(RB), 0, sm, c.V, c.W, ld, rows, ncol,
                                   c.Rq, c.st);
                hipLaunchKernelGGL(reorth_sum_kernel, dim3((ncol * S + 255) / 256), dim3(256), 0, sm, c.Rq,
                                   ncol * S, c.Rcoef, c.st);
                hipLaunchKernelGGL((reorth_apply_kernel<S>), dim3(gb), dim3(RB), (size_t)ncol * S * sizeof(double),
                                   sm, c.V, c.W, ld, rows, ncol, c.Rcoef, pass == 1 ? g1 : (double *)nullptr, c.st);
            }
        }
        if (!twopass && !fused && upd) {
            double *Vn = c.V + (size_t)j * S * ld;
            double *Bj = c.Bblk + (size_t)(j - 1) * S * S;
            hipLaunchKernelGGL((cholqr_kernel<S>), dim3(gb), dim3(RB), 0, sm, c.W, Vn, ld, rows, g1, g0,
                               c.R0 + S * S /* R1 scratch */, (const double *)nullptr, c.st, j + 1);
            hipLaunchKernelGGL((cholqr_kernel<S>), dim3(gb), dim3(RB), 0, sm, Vn, Vn, ld, rows, g0,
                               (double *)nullptr, Bj, (const double *)(c.R0 + S * S), c.st, j + 1);
        }
        BD_HIP(hipGetLastError());
        const bool group_end = grouped && (j == dmax || j - jg_start + 1 >= LG || in_window(j) || in_window(j + 1));
        if (group_end) {
            // the tests of iterations jg0..j on the side stream, behind the
            // Krylov kernels of iteration j: the recurrence goes on meanwhile
            // (it writes only later blocks of V, A and B than the tests read)
            const int jg0 = jg_start;
            const int parity = ngroups & 1;
            jg_start = j + 1;
            ngroups++;
            if (sg != sm) {
                BD_HIP(hipEventRecord(ev, sm));
                BD_HIP(hipStreamWaitEvent(sg, ev, 0));
            }
            if constexpr (S <= 4) {
                const long part_stride = (long)nwg * dmax * S * S;
                const long c_stride = (long)c.dimk_max * S * S;
                // node batches: those of the previous group's placement and one
                // more (a placement of more nodes is covered by the kernel's loop)
                const int nby = pcr_batches > 0 ? std::min(nwg, pcr_batches + 1) : nwg;
                if (use_pcr)
                    hipLaunchKernelGGL((small_pcr_kernel<S>), dim3(j - jg0 + 1, nby), dim3(pcr::THREADS), 0, sg,
                                       c.Ablk, c.Bblk, c.R0, jg0, j, c.Cq, part_stride, c.Ccoef, c_stride, c.split,
                                       parity, c.st);
                else
                    hipLaunchKernelGGL((small_group_kernel<S>), dim3(nwg, j - jg0 + 1), dim3(sp::WG_T), 0, sg,
                                       c.Ablk, c.Bblk, c.R0, jg0, j, c.Cq, part_stride, c.Ccoef, c_stride, c.split,
                                       parity, c.st);
                hipLaunchKernelGGL((ycomb_group_kernel<S>), dim3((rows + YR - 1) / YR), dim3(256), 0, sg, c.V, ld,
                                   rows, jg0, j, c.Ccoef, c_stride, c.Y, two_norm ? 1 : 0, c.npart, c.opt.lanc_tol,
                                   c.st, c.lz_flag_dev);
            }
            BD_HIP(hipGetLastError());
        }
        if (!grouped && j >= 2) {
            BD_HIP(hipStreamWaitEvent(ss, ev, 0));
            const int js = j * S;
            bool launched = false;
            if constexpr (S <= 4) {
                if (split) {
                    hipLaunchKernelGGL((small_split_kernel<S>), dim3(nwg), dim3(sp::WG_T), 0, ss, c.Ablk, c.Bblk, c.R0,
                                       j, c.Cq, c.Ccoef, c.split, c.st);
                    launched = true;
                }
            }
            if (!launched)
                hipLaunchKernelGGL((small_kernel<S>), dim3(1), dim3(256), 0, ss, c.Ablk, c.Bblk, c.R0, j, c.Cq,
                                   c.Ccoef, c.st);
            double *ycur = c.Y + (size_t)(j % 2) * S * ld;
            const double *yold = c.Y + (size_t)((j + 1) % 2) * S * ld;
            const int want = (j > 2) ? 1 : 0;
            hipLaunchKernelGGL((ycomb_kernel<S>), dim3((rows + YR - 1) / YR), dim3(256), 0, ss, c.V, ld, rows, js, c.Ccoef,
                               ycur, yold, want, two_norm ? 1 : 0, c.npart, j, c.opt.lanc_tol, c.st);
            BD_HIP(hipGetLastError());
        }
        if (grouped) {
            if (group_end) {
                if (npend == 64) {   // cannot happen: verdicts older than LG iterations are taken each iteration
                    set_error("internal: too many Lanczos test groups pending");
                    return BD_ERR_STATE;
                }
                pend[npend++] = j;
            }
            // verdicts in group order: those LG iterations old (3 for the
            // single iterations around the hint, whose verdict is the likely
            // stop; the device then still has two iterations queued), waiting
            // for them, then any newer one that has already arrived
            while (npend > 0 && !stop) {
                const int e = pend[0];
                if (j - e >= (in_window(e) ? 3 : LG)) BD_TRY(wait_tests(c, sg, e));
                else if (__atomic_load_n(&c.lz_flag[0], __ATOMIC_ACQUIRE) < e) break;
                for (int q = 1; q < npend; q++) pend[q - 1] = pend[q];
                npend--;
                if (c.lz_flag[2] || c.lz_flag[1]) stop = true;
                if (c.lz_flag[3] > 0) pcr_batches = (c.lz_flag[3] + pcr::NODES - 1) / pcr::NODES;
            }
        } else if (j % chunk == 0 || j == dmax) {
            // snapshot the status after this chunk's tests; read the previous one
            hipStream_t ts = grouped ? sm : ss;
            BD_HIP(hipMemcpyAsync(&c.st_host[slot], c.st, sizeof(DevStatus), hipMemcpyDeviceToHost, ts));
            BD_HIP(hipEventRecord(c.lz_chunk_ev[slot], ts));
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
    // a CholeskyQR failure beyond the converged iteration belongs to an
    // iteration the reference never runs (the tests stop at convergence and
    // report failures only before it)
    if (c.st_host->ierr || (c.st_host->converged_j == 0 && c.st_host->fail_j != 0)) {
        set_error("Lanczos small problem or CholeskyQR failed (matrix not SPD?)");
        return BD_ERR_FACTOR;
    }
    int jf = c.st_host->converged_j ? c.st_host->converged_j : c.st_host->last_j;
    c.lz_hint = c.st_host->converged_j;
    *lconv = c.st_host->converged_j ? 1 : 0;
    *iters = c.st_host->converged_j ? c.st_host->converged_j : dmax;
    const double *yres = c.Y + (size_t)(jf % (grouped ? LG + 1 : 2)) * S * ld;
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

// Panel: rows k0+nb..n-1 of block column k0 := A21 L11^{-T}.  The row is
// solved in registers: the loops run over the full CNB with compile-time
// indices (columns j >= nb are skipped uniformly).
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
#pragma unroll
    for (int j = 0; j < CNB; j++) x[j] = j < nb ? A[(long)(k0 + j) * ld + r] : 0.0;
#pragma unroll
    for (int j = 0; j < CNB; j++) {
        if (j < nb) {
            double t = x[j];
#pragma unroll
            for (int k = 0; k < j; k++) t -= x[k] * L[j][k];
            x[j] = t / L[j][j];
        }
    }
#pragma unroll
    for (int j = 0; j < CNB; j++)
        if (j < nb) A[(long)(k0 + j) * ld + r] = x[j];
}

// The whole factorisation and B.xi in one workgroup of 8 wavefronts, for
// n <= CS_MAX, right-looking over 32-column panels with a one-panel
// lookahead.  For panel p (columns k0..k0+nb-1):
//  A. one thread per row below the diagonal block solves that row against
//     the block (L21 = A21 L11^-T) and stores it transposed in LDS,
//     PT[j][t] = L(k0+nb+t, k0+j), and in A;
//  B. wavefront 0 applies the panel to the next diagonal block itself and
//     factors it, while wavefronts 1-7 form the panel's share of out and
//     update the rest of the trailing lower triangle.
// The diagonal block's factor (a chain of 32 dependent pivots on one
// wavefront) is thus off the critical path except for the first panel.  It
// runs in registers, lane i holding row i; each finished column of L goes to
// DS[j][q] = L(k0+j+q, k0+j) and returns to every lane from there, and a
// short last block is padded with the identity, which leaves its factor
// unchanged.  out = L^T xi (the reference's U.xi, trans) takes column k0+j of
// L whole, so rows k0..k0+nb-1 of out are final after panel p; out = L xi
// adds the panel's columns into every row below k0.  The trailing update
// works in 32 x 32 tiles, one per wavefront, each lane a 4 x 4 block whose two
// row pairs and two column pairs come from PT as 16-byte reads (8 lanes cover
// 128 contiguous bytes, the rest broadcast), accumulated onto A's values
// loaded up front.  Only the last panel can be short (nb < 32), and it has no
// rows below, so A and the trailing update always run with nb = 32.  Two
// workgroup barriers per panel.  A non-positive pivot sets ierr = 2.
constexpr int CS_MAX = 480;
constexpr int CS_T = 512;
constexpr int CS_PT = CS_MAX - CNB;   // trailing rows per panel, at most
constexpr int CS_XS = 2;              // columns of xi staged in LDS

// orders a wavefront's LDS stores before its later loads by other lanes
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lane l's value of v (l wave-uniform), through the scalar unit
__device__ inline double read_lane(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One pivot j of the diagonal block.  Lane i holds row i from column j on,
// x[q] = A(i, j + q); entries from LIVE on are zero.  The next pivot is lane
// j+1's x[1] - L(j+1, j)^2, all its own values, so it is formed (dn) before
// the column makes its round trip through LDS; x[0] after the update is the
// same number.
template <int LIVE>
__device__ inline void diag_pivot(double (&x)[CNB], double &dn, int j, int lane, int nb, double *Acol,
                                  double (*DS)[64], bool &bad, double &inv)
{
    const double djj = read_lane(dn, j);
    if (!(djj > 0.0)) bad = true;
    const double rp = rsqrt(fmax(djj, 1e-300));
    const double cj = lane == j ? djj * rp : x[0] * rp;   // column j of L on lanes >= j
    dn = x[1] - cj * cj;
    inv = lane == j ? rp : inv;
    if (lane >= j) DS[j][lane - j] = cj;
    if (lane >= j && lane < nb) Acol[lane] = cj;
    wave_sync();
    double2 cv[LIVE / 2];
#pragma unroll
    for (int m = 0; m < LIVE / 2; m++) cv[m] = *reinterpret_cast<const double2 *>(&DS[j][2 * m]);
#pragma unroll
    for (int q = 1; q < LIVE; q++) x[q - 1] = x[q] - cj * ((q & 1) ? cv[q / 2].y : cv[q / 2].x);
    x[LIVE - 1] = 0.0;
}

// Factor the diagonal block at (kb, kb) of order nbk on one wavefront, after
// subtracting the panel in PT from it when UPD.  L goes to A's lower part and
// to DS, the pivots' reciprocals to Dinv; a non-positive pivot sets *fail.
template <bool UPD>
__device__ inline void diag_block(double *A, int ld, int kb, int nbk, int lane, double (*DS)[64], double *Dinv,
                                  const double (*PT)[CS_PT], int *fail)
{
    double x[CNB];
#pragma unroll
    for (int j = 0; j < CNB; j++)
        x[j] = (j <= lane && lane < nbk) ? A[(long)(kb + j) * ld + kb + lane] : (j == lane ? 1.0 : 0.0);
    if (UPD) {
#pragma unroll 2
        for (int k = 0; k < CNB; k++) {
            const double pi = lane < nbk ? PT[k][lane] : 0.0;
            double2 pv[CNB / 2];
#pragma unroll
            for (int m = 0; m < CNB / 2; m++) pv[m] = *reinterpret_cast<const double2 *>(&PT[k][2 * m]);
#pragma unroll
            for (int q = 0; q < CNB; q++) x[q] -= pi * ((q & 1) ? pv[q / 2].y : pv[q / 2].x);
        }
    }
    bool bad = false;
    double inv = 0.0, dn = x[0];
    double *Ak = A + (long)kb * ld + kb;
#pragma unroll 1
    for (int j = 0; j < 8; j++) diag_pivot<32>(x, dn, j, lane, nbk, Ak + (long)j * ld, DS, bad, inv);
#pragma unroll 1
    for (int j = 8; j < 16; j++) diag_pivot<24>(x, dn, j, lane, nbk, Ak + (long)j * ld, DS, bad, inv);
#pragma unroll 1
    for (int j = 16; j < 24; j++) diag_pivot<16>(x, dn, j, lane, nbk, Ak + (long)j * ld, DS, bad, inv);
#pragma unroll 1
    for (int j = 24; j < 32; j++) diag_pivot<8>(x, dn, j, lane, nbk, Ak + (long)j * ld, DS, bad, inv);
    if (lane < CNB) Dinv[lane] = inv;
    if (bad && lane == 0) *fail = 1;
}

// One column j of a panel row's solve x L11^T = a: x[q] = the row's entry in
// column j + q, entries from LIVE on are zero.
template <int LIVE>
__device__ inline void panel_col(double (&x)[CNB], int j, int t, const double (*DS)[64], const double *Dinv,
                                 double (*PT)[CS_PT], double *Aj)
{
    double2 dv[LIVE / 2];
#pragma unroll
    for (int m = 0; m < LIVE / 2; m++) dv[m] = *reinterpret_cast<const double2 *>(&DS[j][2 * m]);
    const double xj = x[0] * Dinv[j];
    PT[j][t] = xj;
    *Aj = xj;
#pragma unroll
    for (int q = 1; q < LIVE; q++) x[q - 1] = x[q] - xj * ((q & 1) ? dv[q / 2].y : dv[q / 2].x);
    x[LIVE - 1] = 0.0;
}

__global__ __launch_bounds__(CS_T) void chol_small_kernel(double *A, int ld, int n, const double *__restrict__ xi,
                                                          double *__restrict__ out, int s, int ldx, int trans,
                                                          DevStatus *st)
{
    __shared__ __align__(16) double PT[CNB][CS_PT];
    __shared__ __align__(16) double DS[2][CNB][64];   // DS[p & 1][j][q] = L(k0+j+q, k0+j), zero for j+q >= nb
    __shared__ double Dinv[CNB];
    __shared__ double XS[CS_XS][CS_MAX];
    __shared__ int fail;
    constexpr int W = CS_T / 64, W1 = W - 1, JW = (CNB + W1 - 1) / W1;
    const int t = threadIdx.x, lane = t % 64, w = t / 64;
    if (st->ierr) return;
    if (t == 0) fail = 0;
    const int sx = min(s, CS_XS);
    for (int e = t; e < n * sx; e += CS_T) XS[e / n][e % n] = xi[(long)(e / n) * ldx + e % n];
    auto xv = [&](int c, int i) { return c < CS_XS ? XS[c][i] : xi[(long)c * ldx + i]; };
    if (w == 0) diag_block<false>(A, ld, 0, min(CNB, n), lane, DS[0], Dinv, PT, &fail);
    __syncthreads();
    for (int k0 = 0, p = 0; k0 < n; k0 += CNB, p++) {
        if (fail) {
            if (t == 0) atomicCAS(&st->ierr, 0, 2);
            return;
        }
        const int nb = min(CNB, n - k0), m2 = n - k0 - nb, base = k0 + nb;
        const double(*D)[64] = DS[p & 1];
        if (t < m2) {   // A.  nb == CNB; x[q] = the row's entry in column j + q
            const int r = base + t;
            double x[CNB];
#pragma unroll
            for (int j = 0; j < CNB; j++) x[j] = A[(long)(k0 + j) * ld + r];
            double *Ar = A + (long)k0 * ld + r;
#pragma unroll 2
            for (int j = 0; j < 8; j++) panel_col<32>(x, j, t, D, Dinv, PT, Ar + (long)j * ld);
#pragma unroll 2
            for (int j = 8; j < 16; j++) panel_col<24>(x, j, t, D, Dinv, PT, Ar + (long)j * ld);
#pragma unroll 2
            for (int j = 16; j < 24; j++) panel_col<16>(x, j, t, D, Dinv, PT, Ar + (long)j * ld);
#pragma unroll 2
            for (int j = 24; j < 32; j++) panel_col<8>(x, j, t, D, Dinv, PT, Ar + (long)j * ld);
        }
        __syncthreads();
        if (w == 0) {   // B on wavefront 0: the next diagonal block
            if (m2 > 0) diag_block<true>(A, ld, base, min(CNB, m2), lane, DS[(p + 1) & 1], Dinv, PT, &fail);
        } else {
            const int w1 = w - 1, t1 = t - 64;
            if (trans) {   // columns w1, w1 + W1, ... of L, whole, against xi
                for (int c = 0; c < s; c++) {
                    double acc[JW];
#pragma unroll
                    for (int u = 0; u < JW; u++) {
                        const int j = w1 + W1 * u;
                        acc[u] = (j < nb && lane >= j && lane < nb) ? D[j][lane - j] * xv(c, k0 + lane) : 0.0;
                    }
                    for (int i = lane; i < m2; i += 64) {   // m2 > 0 only when nb == CNB
                        const double xr = xv(c, base + i);
#pragma unroll
                        for (int u = 0; u < JW; u++)
                            if (w1 + W1 * u < CNB) acc[u] += PT[w1 + W1 * u][i] * xr;
                    }
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
                        for (int u = 0; u < JW; u++) acc[u] += __shfl_xor(acc[u], off, 64);
                    if (lane == 0)
#pragma unroll
                        for (int u = 0; u < JW; u++)
                            if (w1 + W1 * u < nb) out[(long)c * ldx + k0 + w1 + W1 * u] = acc[u];
                }
            } else {   // the panel's columns into rows k0.. of L xi
                for (int r = k0 + t1; r < n; r += CS_T - 64)
                    for (int c = 0; c < s; c++) {
                        double acc = 0.0;
                        if (r < base)
                            for (int j = 0; j <= r - k0; j++) acc += D[j][r - k0 - j] * xv(c, k0 + j);
                        else
                            for (int j = 0; j < nb; j++) acc += PT[j][r - base] * xv(c, k0 + j);
                        double *o = out + (long)c * ldx + r;
                        *o = k0 == 0 ? acc : *o + acc;
                    }
            }
            // trailing update A22 -= L21 L21^T but its first diagonal tile
            const int qt = (m2 + 31) / 32, ntile = qt * (qt + 1) / 2;
            const int lr = 2 * (lane % 8), lc = 2 * (lane / 8);
            for (int e = 1 + w1; e < ntile; e += W1) {
                int tr = (int)((sqrt(8.0 * e + 1.0) - 1.0) * 0.5);
                while ((tr + 1) * (tr + 2) / 2 <= e) tr++;
                while (tr * (tr + 1) / 2 > e) tr--;
                const int tc = e - tr * (tr + 1) / 2;
                const int R0 = 32 * tr + lr, C0 = 32 * tc + lc;
                double acc[4][4];
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int r = R0 + (a & 1) + 16 * (a >> 1), cc = C0 + (b & 1) + 16 * (b >> 1);
                        acc[a][b] = (r < m2 && cc < m2 && r >= cc) ? A[(long)(base + cc) * ld + base + r] : 0.0;
                    }
#pragma unroll 8
                for (int k = 0; k < CNB; k++) {
                    const double2 u0 = *reinterpret_cast<const double2 *>(&PT[k][R0]);
                    const double2 u1 = *reinterpret_cast<const double2 *>(&PT[k][R0 + 16]);
                    const double2 v0 = *reinterpret_cast<const double2 *>(&PT[k][C0]);
                    const double2 v1 = *reinterpret_cast<const double2 *>(&PT[k][C0 + 16]);
                    const double u[4] = {u0.x, u0.y, u1.x, u1.y}, v[4] = {v0.x, v0.y, v1.x, v1.y};
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int b = 0; b < 4; b++) acc[a][b] -= u[a] * v[b];
                }
#pragma unroll
                for (int a = 0; a < 4; a++)
#pragma unroll
                    for (int b = 0; b < 4; b++) {
                        const int r = R0 + (a & 1) + 16 * (a >> 1), cc = C0 + (b & 1) + 16 * (b >> 1);
                        if (r < m2 && cc < m2 && r >= cc) A[(long)(base + cc) * ld + base + r] = acc[a][b];
                    }
            }
        }
        __syncthreads();
    }
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

int brownian_cholesky(Ctx &c, const double *xi, double *out, int s, int method, bool sync)
{
    const int n = c.n;
    if (!c.chol) {
        c.chol_ld = n;
        BD_HIP(hipMalloc(&c.chol, (size_t)n * n * sizeof(double)));
    }
    BD_TRY(mob_unpack(c, c.chol, c.chol_ld));
    hipStream_t sm = c.stream;
    const int ld = c.chol_ld;
    if (n <= CS_MAX) {
        hipLaunchKernelGGL(chol_small_kernel, dim3(1), dim3(CS_T), 0, sm, c.chol, ld, n, xi, out, s, c.npad,
                           method == BD_CHOL ? 1 : 0, c.st);
        BD_HIP(hipGetLastError());
        if (!sync) return BD_OK;   // a failure stops the step's later kernels (ierr)
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
    if (!sync) return BD_OK;
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
