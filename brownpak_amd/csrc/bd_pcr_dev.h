// bd_pcr_dev.h -- the Lanczos small problem by parallel cyclic reduction.
//
// H_j (order js = nb S) is block tridiagonal: diagonal blocks A_b (symmetric)
// and sub-diagonal blocks B_{b-1} = H(b, b-1).  Every quadrature node needs
// (H + sigma I)^{-1} X with X = E_1 R_0, and every spectral probe the
// verdict "H - m I positive definite".  A block Cholesky sweep does either in
// nb dependent block steps.  Parallel cyclic reduction does it in
// ceil(log2 nb) levels with one lane per block row: at the level of stride h
// row i eliminates its couplings to rows i - h and i + h,
//   a_i = -L_i D_{i-h}^{-1},   g_i = -L_{i+h}^T D_{i+h}^{-1},
//   D_i <- D_i + a_i L_i^T + g_i L_{i+h},   L_i <- a_i L_{i-h},
//   r_i <- r_i + a_i r_{i-h} + g_i r_{i+h},
// where L_i couples row i to row i - h and the coupling to row i + h is
// L_{i+h}^T (every reduced system stays symmetric).  Each D is a Schur
// complement of H, so all are positive definite exactly when H is (the odd-
// even reduction's pivots are among them); after the last level the rows are
// decoupled: Z_i = D_i^{-1} r_i.  A wave holds one task (node or probe), a
// lane one block row, so nb <= 64.
#pragma once

#include "bd_lanczos_dev.h"

namespace bd {
namespace pcr {

constexpr int ROWS = 64;        // block rows per task (lanes of a wave)
constexpr int PROBES = 1;       // probe waves per workgroup
constexpr int NODES = 3;        // node waves per workgroup (one wave per SIMD in all)
constexpr int WAVES = PROBES + NODES;
constexpr int THREADS = 64 * WAVES;

template <int S>
__device__ inline void zero(double (&a)[S][S])
{
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c < S; c++) a[r][c] = 0.0;
}

// 1/d from the hardware estimate refined by two Newton-Raphson steps
// (x <- x + x (1 - d x)), each of which squares the relative error.
__device__ inline double rcp_nr(double d)
{
    double x = __builtin_amdgcn_rcp(d);
    double e = fma(-d, x, 1.0);
    x = fma(x, e, x);
    e = fma(-d, x, 1.0);
    return fma(x, e, x);
}

// D = L diag(d) L^T of the symmetric D (lower triangle read; L unit lower),
// then inv = D^{-1} = M^T diag(1/d) M with M = L^{-1}.  No square roots: the
// pivots d_j are positive exactly when D is positive definite, which is the
// verdict returned (inv is finite garbage otherwise).
template <int S>
__device__ inline bool chol_inv(const double (&D)[S][S], double (&inv)[S][S])
{
    double L[S][S], W[S][S], rd[S];   // W[i][k] = L[i][k] d_k
    bool ok = true;
#pragma unroll
    for (int j = 0; j < S; j++) {
#pragma unroll
        for (int i = 0; i < j; i++) {
            double t = D[j][i];
#pragma unroll
            for (int k = 0; k < i; k++) t -= W[j][k] * L[i][k];
            W[j][i] = t;
            L[j][i] = t * rd[i];
        }
        double dj = D[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) dj -= W[j][k] * L[j][k];
        ok = ok && (dj > 0.0);
        rd[j] = rcp_nr(dj);
    }
    // M = L^{-1}, unit lower triangular (M[i][k], k < i)
    double M[S][S];
#pragma unroll
    for (int i = 0; i < S; i++)
#pragma unroll
        for (int k = 0; k < i; k++) {
            double t = -L[i][k];
#pragma unroll
            for (int m = k + 1; m < i; m++) t -= L[i][m] * M[m][k];
            M[i][k] = t;
        }
    // inv(a, b) = sum_{k >= a} M(k, a) rd_k M(k, b), b <= a, with M(k, k) = 1
#pragma unroll
    for (int a = 0; a < S; a++)
#pragma unroll
        for (int b = 0; b <= a; b++) {
            double t = rd[a] * (a == b ? 1.0 : M[a][b]);
#pragma unroll
            for (int k = a + 1; k < S; k++) t += M[k][a] * rd[k] * M[k][b];
            inv[a][b] = t;
            inv[b][a] = t;
        }
    return ok;
}

// Shared-memory slots of one task: element e of row i at buf[e * ROWS + i]
// (lanes read consecutive addresses).  Node tasks publish inv, L, r (3 S^2
// values), probes inv, L (2 S^2).
template <int S>
__device__ inline void put_blk(double *buf, int e0, int i, const double (&a)[S][S])
{
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c < S; c++) buf[(e0 + r * S + c) * ROWS + i] = a[r][c];
}
template <int S>
__device__ inline void get_blk(const double *buf, int e0, int i, double (&a)[S][S])
{
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c < S; c++) a[r][c] = buf[(e0 + r * S + c) * ROWS + i];
}

// One task by one wave (every thread of the workgroup calls it the same
// number of times with the same nb: it contains block barriers, the same
// ones for every wave).  Lane i < nb holds block row i.  RHS (uniform over
// the wave): solve for X = E_1 R0 (node), else only test positive
// definiteness (probe).  Returns (uniformly over the wave) whether every
// factorisation succeeded; a node's lane i leaves Z_i = (H + shift I)^{-1}
// X restricted to block row i in Z.
template <int S>
__device__ bool reduce(const double *__restrict__ sA, const double *__restrict__ sB, const double *__restrict__ sR,
                       int nb, double shift, bool RHS, double *buf, double (&Z)[S][S])
{
    const int i = threadIdx.x % 64;
    const bool live = i < nb;
    double D[S][S], L[S][S], r[S][S];
    zero<S>(L);
    zero<S>(r);
    zero<S>(D);
    if (live) {
#pragma unroll
        for (int a = 0; a < S; a++)
#pragma unroll
            for (int b = 0; b < S; b++) {
                // A_i: lower triangle stored column-major
                D[a][b] = (a >= b) ? sA[i * S * S + a + b * S] : sA[i * S * S + b + a * S];
                if (i > 0) L[a][b] = sB[(i - 1) * S * S + a + b * S];
                if (RHS && i == 0) r[a][b] = sR[a + b * S];
            }
#pragma unroll
        for (int a = 0; a < S; a++) D[a][a] += shift;
    }
    bool ok = true;
    for (int h = 1; h < nb; h *= 2) {
        double inv[S][S];
        ok = chol_inv<S>(D, inv) && ok;
        __syncthreads();   // the previous level's slots are no longer read
        if (live) {
            put_blk<S>(buf, 0, i, inv);
            put_blk<S>(buf, S * S, i, L);
            if (RHS) put_blk<S>(buf, 2 * S * S, i, r);
        }
        __syncthreads();
        double Ln[S][S];
        zero<S>(Ln);
        if (live && i - h >= 0) {
            double im[S][S], Lm[S][S], al[S][S];
            get_blk<S>(buf, 0, i - h, im);
            // a = -L_i inv_{i-h}
#pragma unroll
            for (int a = 0; a < S; a++)
#pragma unroll
                for (int b = 0; b < S; b++) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < S; k++) t += L[a][k] * im[k][b];
                    al[a][b] = -t;
                }
            // D += a L_i^T (symmetric: lower triangle, mirrored)
#pragma unroll
            for (int a = 0; a < S; a++)
#pragma unroll
                for (int b = 0; b <= a; b++) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < S; k++) t += al[a][k] * L[b][k];
                    D[a][b] += t;
                    if (b < a) D[b][a] = D[a][b];
                }
            get_blk<S>(buf, S * S, i - h, Lm);
#pragma unroll
            for (int a = 0; a < S; a++)
#pragma unroll
                for (int b = 0; b < S; b++) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < S; k++) t += al[a][k] * Lm[k][b];
                    Ln[a][b] = t;
                }
            if (RHS) {
                double rm[S][S];
                get_blk<S>(buf, 2 * S * S, i - h, rm);
#pragma unroll
                for (int a = 0; a < S; a++)
#pragma unroll
                    for (int b = 0; b < S; b++) {
                        double t = r[a][b];
#pragma unroll
                        for (int k = 0; k < S; k++) t += al[a][k] * rm[k][b];
                        r[a][b] = t;
                    }
            }
        }
        if (live && i + h < nb) {
            double ip[S][S], Lp[S][S], ga[S][S];
            get_blk<S>(buf, 0, i + h, ip);
            get_blk<S>(buf, S * S, i + h, Lp);
            // g = -L_{i+h}^T inv_{i+h}
#pragma unroll
            for (int a = 0; a < S; a++)
#pragma unroll
                for (int b = 0; b < S; b++) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < S; k++) t += Lp[k][a] * ip[k][b];
                    ga[a][b] = -t;
                }
            // D += g L_{i+h}
#pragma unroll
            for (int a = 0; a < S; a++)
#pragma unroll
                for (int b = 0; b <= a; b++) {
                    double t = 0.0;
#pragma unroll
                    for (int k = 0; k < S; k++) t += ga[a][k] * Lp[k][b];
                    D[a][b] += t;
                    if (b < a) D[b][a] = D[a][b];
                }
            if (RHS) {
                double rp[S][S];
                get_blk<S>(buf, 2 * S * S, i + h, rp);
#pragma unroll
                for (int a = 0; a < S; a++)
#pragma unroll
                    for (int b = 0; b < S; b++) {
                        double t = r[a][b];
#pragma unroll
                        for (int k = 0; k < S; k++) t += ga[a][k] * rp[k][b];
                        r[a][b] = t;
                    }
            }
        }
#pragma unroll
        for (int a = 0; a < S; a++)
#pragma unroll
            for (int b = 0; b < S; b++) L[a][b] = Ln[a][b];
    }
    double inv[S][S];
    ok = chol_inv<S>(D, inv) && ok;
    if (RHS) {
#pragma unroll
        for (int a = 0; a < S; a++)
#pragma unroll
            for (int b = 0; b < S; b++) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < S; k++) t += inv[a][k] * r[k][b];
                Z[a][b] = t;
            }
    }
    // the wave's verdict (lanes past nb hold ok = true)
    ok = __all(live ? ok : true);
    return ok;
}

}  // namespace pcr
}  // namespace bd
