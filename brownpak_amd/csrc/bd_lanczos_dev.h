// bd_lanczos_dev.h -- device helpers shared by the single-chain and the
// batched block-Lanczos kernels: block reductions, S x S Cholesky, the band
// of H_j, banded Cholesky, and the quadrature node table.
#pragma once

#include "bd_internal.h"

namespace bd {

constexpr int GBLK = 120;     // blocks of the Gram/partial-norm reductions
constexpr int RB = 256;       // threads per reduction block

// Quadrature table: for kappa <= 10^(d), (N, h) with 2N+1 nodes (d = 1..13).
static __constant__ int c_qN[14] = {20, 20, 27, 34, 42, 50, 57, 66, 79, 87, 97, 109, 110, 125};
static __constant__ double c_qh[14] = {0.205, 0.205, 0.155, 0.125, 0.1, 0.085, 0.075,
                                0.065, 0.055, 0.05, 0.045, 0.04, 0.04, 0.035};

static __device__ inline bool lz_done(const DevStatus *st) { return st->converged_j != 0 || st->ierr != 0; }

// ------------------------------------------------------------ reductions
template <int S>
__device__ void block_reduce_store(double (&acc)[S * S], double *out)
{
    __shared__ double sh[RB / 64][S * S];
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
        for (int w = 0; w < RB / 64; w++) v += sh[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
}

// Sum of GBLK partials of S*S values, fixed order, by the whole block
// (blockDim.x == RB).  Result in g (shared), valid after the call.
template <int S>
__device__ void sum_partials(const double *part, double *g)
{
    constexpr int E = S * S;
    constexpr int STR = RB / E;          // stripes over the partial index
    __shared__ double sh[RB];
    const int t = threadIdx.x;
    double v = 0.0;
    if (t < E * STR) {
        const int e = t % E, st = t / E;
        for (int b = st; b < GBLK; b += STR) v += part[b * E + e];
    }
    sh[t] = v;
    __syncthreads();
    if (t < E) {
        double a = 0.0;
        for (int st = 0; st < STR; st++) a += sh[st * E + t];
        g[t] = a;
    }
    __syncthreads();
}

// Cholesky of an S x S SPD matrix g (col-major): g = R^T R, R upper.
template <int S>
__device__ bool small_chol(const double *g, double (&R)[S][S])
{
    for (int i = 0; i < S; i++)
        for (int j = 0; j < S; j++) R[i][j] = 0.0;
    for (int j = 0; j < S; j++) {
        double d = g[j + j * S];
        for (int k = 0; k < j; k++) d -= R[k][j] * R[k][j];
        if (!(d > 0.0)) return false;
        d = sqrt(d);
        R[j][j] = d;
        for (int i = j + 1; i < S; i++) {
            double t = g[j + i * S];
            for (int k = 0; k < j; k++) t -= R[k][j] * R[k][i];
            R[j][i] = t / d;
        }
    }
    return true;
}

// Band of H_j (lower, bandwidth S): Hb[d * ldb + col] = H(col + d, col).
template <int S>
__device__ double h_entry(const double *Ablk, const double *Bblk, int js, int row, int col)
{
    // row >= col assumed, row - col <= S
    int bi = row / S, bj = col / S, r = row % S, c = col % S;
    if (bi == bj) return Ablk[bi * S * S + r + c * S];       // lower of A (r >= c)
    if (bi == bj + 1) return Bblk[bj * S * S + r + c * S];   // B_{bj} = H_{bj+1, bj}
    return 0.0;
}

// Gershgorin upper bound of H_j, computed by one wavefront (64 lanes).
template <int S>
__device__ double gersh_hi(const double *Ablk, const double *Bblk, int js)
{
    double mx = 0.0;
    for (int i = threadIdx.x; i < js; i += 64) {
        double rs = 0.0;
        for (int l = max(0, i - S); l <= min(js - 1, i + S); l++) {
            double h = (i >= l) ? h_entry<S>(Ablk, Bblk, js, i, l) : h_entry<S>(Ablk, Bblk, js, l, i);
            rs += fabs(h);
        }
        mx = fmax(mx, rs);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    return mx;
}

// Banded Cholesky in LDS by one wavefront.  band[d][col] (d = 0..S),
// stride ldb.  Returns false if not positive definite.
template <int S>
__device__ bool band_chol(double *band, int ldb, int js)
{
    const int lane = threadIdx.x;
    __shared__ int s_fail;
    if (lane == 0) s_fail = 0;
    __syncthreads();
    // lane -> (d1, d2) pair for the trailing update, 1 <= d2 <= d1 <= S
    int d1 = 0, d2 = 0;
    {
        int idx = lane, k = 0;
        for (int a = 1; a <= S && !d1; a++)
            for (int b = 1; b <= a; b++) {
                if (k == idx) { d1 = a; d2 = b; }
                k++;
            }
    }
    for (int c = 0; c < js; c++) {
        if (lane == 0) {
            double a = band[c];
            if (!(a > 0.0)) s_fail = 1;
            band[c] = sqrt(fmax(a, 0.0));
        }
        __syncthreads();
        if (s_fail) return false;
        if (lane >= 1 && lane <= S && c + lane < js) band[lane * ldb + c] /= band[c];
        __syncthreads();
        if (d1 && c + d1 < js) {
            // A(c+d1, c+d2) -= L(c+d1, c) L(c+d2, c)
            band[(d1 - d2) * ldb + c + d2] -= band[d1 * ldb + c] * band[d2 * ldb + c];
        }
        __syncthreads();
    }
    return true;
}

template <int S>
__device__ void load_band(const double *Ablk, const double *Bblk, int js, double shift, double *band, int ldb)
{
    for (int e = threadIdx.x; e < (S + 1) * ldb; e += blockDim.x) {
        int d = e / ldb, col = e % ldb;
        double v = 0.0;
        if (col < js && col + d < js) v = h_entry<S>(Ablk, Bblk, js, col + d, col);
        if (d == 0) v -= shift;
        band[e] = v;
    }
    __syncthreads();
}

}  // namespace bd
