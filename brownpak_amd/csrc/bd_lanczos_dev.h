// bd_lanczos_dev.h -- device helpers shared by the single-chain and the
// batched block-Lanczos kernels: block reductions, S x S Cholesky, the band
// of H_j, banded Cholesky, and the quadrature node table.
#pragma once

#include "bd_internal.h"

namespace bd {

constexpr int GBLK = 120;     // most blocks of the Gram reductions (one per RB rows below that)
constexpr int RB = 256;       // threads per reduction block

// Quadrature table: for kappa <= 10^(d), (N, h) with 2N+1 nodes (d = 1..13).
static __constant__ int c_qN[14] = {20, 20, 27, 34, 42, 50, 57, 66, 79, 87, 97, 109, 110, 125};
static __constant__ double c_qh[14] = {0.205, 0.205, 0.155, 0.125, 0.1, 0.085, 0.075,
                                0.065, 0.055, 0.05, 0.045, 0.04, 0.04, 0.035};

// Main-stream kernels of a Lanczos solve stop once it has converged, failed,
// or a block of its Krylov recurrence could not be formed; the tests on the
// side stream go on through the iterations before that block.
static __device__ inline bool lz_done(const DevStatus *st)
{
    return st->converged_j != 0 || st->ierr != 0 || st->fail_j != 0;
}
static __device__ inline bool lz_tests_done(const DevStatus *st) { return st->converged_j != 0 || st->ierr != 0; }
// Iteration j of a solve is valid (its blocks V_1..V_j exist).
static __device__ inline bool lz_valid(const DevStatus *st, int j) { return st->fail_j == 0 || j < st->fail_j; }
// A CholeskyQR of the Krylov recurrence failed while forming block blk.  The
// recurrence runs ahead of the convergence tests; the tests go on through
// iteration blk - 1, and the solve fails only if none of them converged
// (a converged iteration stands, as in brown_m.f90:497-506, which stops
// there and never forms the later blocks).
static __device__ inline void lz_factor_failed(DevStatus *st, int blk)
{
    int old = st->fail_j;
    while (old == 0 || blk < old) {
        const int prev = atomicCAS(&st->fail_j, old, blk);
        if (prev == old) break;
        old = prev;
    }
}

// The last-block pattern: every workgroup stores its partials with
// device-scope atomic stores (coherent across the XCDs' L2s without any
// cache maintenance), each wave waits until its stores have completed, the
// workgroup synchronises, and one thread bumps a counter; the workgroup that
// completes the count reads the partials with device-scope atomic loads.
// (A release fence instead writes back the whole L2 of the XCD, several
// microseconds per workgroup; an acquire fence invalidates it.)
static __device__ inline void store_dev(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ inline double load_dev(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
static __device__ inline void stores_complete() { __builtin_amdgcn_s_waitcnt(0); }

// ------------------------------------------------------------ reductions
// Sums of the S*S per-thread values acc over the RB threads of the block
// into out (global), in a fixed order: through shared memory, 16 values at
// a time -- thread t adds threads 16g .. 16g+15 for value t % 16, group
// g = t / 16, then 16 threads add the 16 group sums.  (Two barriers per 16
// values; a shuffle tree costs six dependent cross-lane round trips per
// value instead.)
template <int S>
__device__ void block_reduce_store(double (&acc)[S * S], double *out)
{
    constexpr int E = S * S, CH = E < 16 ? E : 16, NG = RB / 16;
    __shared__ double sh[RB][CH + 1];
    __shared__ double s2[NG][CH];
    const int t = threadIdx.x;
#pragma unroll
    for (int e0 = 0; e0 < E; e0 += CH) {
#pragma unroll
        for (int q = 0; q < CH; q++)
            if (e0 + q < E) sh[t][q] = acc[e0 + q];
        __syncthreads();
        if (t < NG * CH) {
            const int o = t % CH, g = t / CH;
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < 16; i++) v += sh[g * 16 + i][o];
            s2[g][o] = v;
        }
        __syncthreads();
        if (t < CH && e0 + t < E) {
            double v = 0.0;
#pragma unroll
            for (int g = 0; g < NG; g++) v += s2[g][t];
            out[e0 + t] = v;
        }
        __syncthreads();
    }
}

// p[0] + p[stride] + ... (cnt terms), added in index order with sixteen
// loads in flight: the same additions as the plain loop, without a memory
// round trip per term (partials written by other XCDs cost ~0.3 us each).
template <bool DEV = false, int U = 16>
__device__ __forceinline__ double sum_strided(const double *__restrict__ p, long stride, int cnt)
{
    double t = 0.0;
    for (int b = 0; b < cnt; b += U) {
        double v[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const double *a = p + (long)min(b + q, cnt - 1) * stride;
            v[q] = DEV ? load_dev(a) : *a;
        }
#pragma unroll
        for (int q = 0; q < U; q++)
            if (b + q < cnt) t += v[q];
    }
    return t;
}

// out[e] = sum over b < cnt of part[b * E + e], e < E <= blockDim.x, by the
// whole block: G = blockDim.x / E groups, group g summing b = g, g + G, ...;
// the group sums are added in group order (a fixed order for a given block
// size).  scr: blockDim.x doubles of shared memory; every thread of the
// block must call; ends with a barrier.  DEV: partials published with
// store_dev by other workgroups of the same launch.
template <bool DEV = false, int U = 16>
__device__ __forceinline__ void block_sum_parts(const double *__restrict__ part, int E, int cnt,
                                                double *scr, double *out)
{
    const int t = threadIdx.x, G = blockDim.x / E, g = t / E, e = t % E;
    double v = 0.0;
    if (g < G && cnt > g) v = sum_strided<DEV, U>(part + (long)g * E + e, (long)G * E, (cnt - g + G - 1) / G);
    scr[t] = v;
    __syncthreads();
    if (t < E) {
        double a = 0.0;
        for (int q = 0; q < G; q++) a += scr[q * E + t];
        out[t] = a;
    }
    __syncthreads();
}

// Sum of n partials of S*S values, fixed order, by the whole block
// (blockDim.x == RB).  Result in g (shared), valid after the call.
template <int S>
__device__ void sum_partials_n(const double *part, int n, double *g)
{
    constexpr int E = S * S;
    constexpr int STR = RB / E;          // stripes over the partial index
    __shared__ double sh[RB];
    const int t = threadIdx.x;
    double v = 0.0;
    if (t < E * STR) {
        const int e = t % E, st = t / E;
        for (int b = st; b < n; b += STR) v += part[b * E + e];
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

// The gridDim.x partials left by a kernel of the same grid.
template <int S>
__device__ void sum_partials(const double *part, double *g)
{
    sum_partials_n<S>(part, (int)gridDim.x, g);
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

// ------------------------------------------------------------ small problem
// H_j (order js = nb * S) is block tridiagonal with S x S blocks: diagonal
// A_b (symmetric, lower triangle used) and sub-diagonal B_{b-1} =
// H(bS.., (b-1)S..), upper triangular (the R factor of CholeskyQR).  For a
// shift sigma, H + sigma I = L L^T with L block lower bidiagonal: diagonal
// blocks L_b (lower) and sub-diagonal blocks G_b = B_{b-1} L_{b-1}^{-T}
// (upper).  One thread owns one shift; its blocks live in registers with
// compile-time indices, and the blocks the backward sweep needs go to a
// node-interleaved scratch array (coalesced over the threads of a block):
// only the triangles of L_b and G_b that can be nonzero, and F_b; the
// diagonal slots of L_b hold 1/diag(L_b).
namespace sp {

template <int S>
constexpr int NT = S * (S + 1) / 2;    // entries of a packed S x S triangle
template <int S>
constexpr int NE = 2 * NT<S> + S * S;  // per block: L (lower), G (upper), F (full)
template <int S>
constexpr int OFF_L = 0;
template <int S>
constexpr int OFF_G = NT<S>;
template <int S>
constexpr int OFF_F = 2 * NT<S>;

// Doubles of small_problem's scratch for H_j of up to nbm blocks and nqp
// node slots: the node-interleaved factor blocks, then one partial sum of Z
// per wave of nodes.
__host__ __device__ constexpr long scratch_doubles(int s, int nbm, int nqp)
{
    return (long)nbm * ((long)(s * (s + 1) + s * s) * nqp + (long)((nqp + 63) / 64) * s * s);
}

template <int S>
__device__ inline void load_blk(const double *__restrict__ p, double (&a)[S][S])
{
#pragma unroll
    for (int c = 0; c < S; c++)
#pragma unroll
        for (int r = 0; r < S; r++) a[r][c] = p[r + c * S];
}

// In-place Cholesky of the lower triangle of a: the strictly lower part of
// L below the diagonal, and inv = 1 / diag(L), which is also left on the
// diagonal of a (diag(L) itself is never needed: every use divides by it).
template <int S>
__device__ inline bool chol(double (&a)[S][S], double (&inv)[S])
{
    bool ok = true;
#pragma unroll
    for (int c = 0; c < S; c++) {
        double d = a[c][c];
#pragma unroll
        for (int k = 0; k < c; k++) d -= a[c][k] * a[c][k];
        ok = ok && (d > 0.0);
        const double r = rsqrt_f64(fmax(d, 1e-300));
        a[c][c] = r;
        inv[c] = r;
#pragma unroll
        for (int i = c + 1; i < S; i++) {
            double t = a[i][c];
#pragma unroll
            for (int k = 0; k < c; k++) t -= a[i][k] * a[c][k];
            a[i][c] = t * r;
        }
    }
    return ok;
}

// G = B Lp^{-T}.  B is upper triangular (the R factor of CholeskyQR, whose
// lower entries are exact zeros), so is G: its upper triangle is formed and
// its lower triangle set to zero, the value the full recurrence gives there.
template <int S>
__device__ inline void gsolve(const double (&B)[S][S], const double (&Lp)[S][S], const double (&ip)[S],
                              double (&G)[S][S])
{
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c < r; c++) G[r][c] = 0.0;
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = r; c < S; c++) {
            double t = B[r][c];
#pragma unroll
            for (int k = 0; k < c; k++) t -= G[r][k] * Lp[c][k];
            G[r][c] = t * ip[c];
        }
}

// A_b + sigma I - G G^T into a (lower), then its Cholesky.
template <int S>
__device__ inline bool next_diag(const double *__restrict__ Ab, double sigma, const double (&G)[S][S],
                                 double (&a)[S][S], double (&inv)[S])
{
    load_blk<S>(Ab, a);
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c <= r; c++) {
            double t = a[r][c] + (r == c ? sigma : 0.0);
#pragma unroll
            for (int k = 0; k < S; k++) t -= G[r][k] * G[c][k];
            a[r][c] = t;
        }
    return chol<S>(a, inv);
}

// Is H + sigma I positive definite?  (no storage)
template <int S>
__device__ bool factor_ok(const double *__restrict__ A, const double *__restrict__ B, int nb, double sigma)
{
    double L[S][S], inv[S];
    load_blk<S>(A, L);
#pragma unroll
    for (int r = 0; r < S; r++) L[r][r] += sigma;
    if (!chol<S>(L, inv)) return false;
    for (int b = 1; b < nb; b++) {
        double Bm[S][S], G[S][S], Ln[S][S], in[S];
        load_blk<S>(B + (b - 1) * S * S, Bm);
        gsolve<S>(Bm, L, inv, G);
        if (!next_diag<S>(A + b * S * S, sigma, G, Ln, in)) return false;
#pragma unroll
        for (int r = 0; r < S; r++) {
            inv[r] = in[r];
#pragma unroll
            for (int c = 0; c < S; c++) L[r][c] = Ln[r][c];
        }
    }
    return true;
}

// Stores / loads of one block's slot in the node-interleaved scratch: the
// full S x S block (column-major), the lower triangle with the diagonal, or
// the upper triangle with the diagonal (packed by columns); get_up sets the
// lower triangle to zero.
template <int S>
__device__ inline void put(double *__restrict__ scr, int nqp, int node, int slot, const double (&a)[S][S])
{
#pragma unroll
    for (int c = 0; c < S; c++)
#pragma unroll
        for (int r = 0; r < S; r++) scr[(long)(slot + r + c * S) * nqp + node] = a[r][c];
}
template <int S>
__device__ inline void get(const double *__restrict__ scr, int nqp, int node, int slot, double (&a)[S][S])
{
#pragma unroll
    for (int c = 0; c < S; c++)
#pragma unroll
        for (int r = 0; r < S; r++) a[r][c] = scr[(long)(slot + r + c * S) * nqp + node];
}
template <int S>
__device__ inline void put_lo(double *__restrict__ scr, int nqp, int node, int slot, const double (&a)[S][S])
{
    int e = slot;
#pragma unroll
    for (int c = 0; c < S; c++)
#pragma unroll
        for (int r = c; r < S; r++, e++) scr[(long)e * nqp + node] = a[r][c];
}
template <int S>
__device__ inline void get_lo(const double *__restrict__ scr, int nqp, int node, int slot, double (&a)[S][S])
{
    int e = slot;
#pragma unroll
    for (int c = 0; c < S; c++)
#pragma unroll
        for (int r = c; r < S; r++, e++) a[r][c] = scr[(long)e * nqp + node];
}
template <int S>
__device__ inline void put_up(double *__restrict__ scr, int nqp, int node, int slot, const double (&a)[S][S])
{
    int e = slot;
#pragma unroll
    for (int c = 0; c < S; c++)
#pragma unroll
        for (int r = 0; r <= c; r++, e++) scr[(long)e * nqp + node] = a[r][c];
}
template <int S>
__device__ inline void get_up(const double *__restrict__ scr, int nqp, int node, int slot, double (&a)[S][S])
{
    int e = slot;
#pragma unroll
    for (int c = 0; c < S; c++)
#pragma unroll
        for (int r = 0; r < S; r++)
            if (r <= c) a[r][c] = scr[(long)(e++) * nqp + node];
            else a[r][c] = 0.0;
}

// One quadrature node: solves (H + sigma I) Z = E1 R0 and hands each block
// Z_b, from the last to the first, to sink(b, Z_b).  The factor blocks go
// through the scratch column of slot node; wr = false leaves the scratch
// unwritten (a lane that repeats another lane's node to keep its wave
// whole).  Returns false if the factorisation fails.  PF: load the next
// block of the backward sweep one step ahead (worth it when the scratch is
// global memory).
template <int S, bool PF = true, typename Sink>
__device__ bool node_solve(const double *__restrict__ A, const double *__restrict__ B, const double *__restrict__ R0,
                           int nb, double sigma, double *__restrict__ scr, int nqp, int node, bool wr, Sink sink)
{
    constexpr int E = NE<S>;
    double L[S][S], inv[S], F[S][S];
    load_blk<S>(A, L);
#pragma unroll
    for (int r = 0; r < S; r++) L[r][r] += sigma;
    bool ok = chol<S>(L, inv);
    {
        double R[S][S];
        load_blk<S>(R0, R);
#pragma unroll
        for (int c = 0; c < S; c++)
#pragma unroll
            for (int i = 0; i < S; i++) {
                double t = R[i][c];
#pragma unroll
                for (int k = 0; k < i; k++) t -= L[i][k] * F[k][c];
                F[i][c] = t * inv[i];
            }
    }
    // forward sweep: factor and F_b = -L_b^{-1} G_b F_{b-1}
    for (int b = 0;; b++) {
        if (wr) {
            put_lo<S>(scr, nqp, node, b * E + OFF_L<S>, L);
            put<S>(scr, nqp, node, b * E + OFF_F<S>, F);
        }
        if (b + 1 >= nb) break;
        double Bm[S][S], G[S][S], Ln[S][S], in[S];
        load_blk<S>(B + b * S * S, Bm);
        gsolve<S>(Bm, L, inv, G);
        ok = next_diag<S>(A + (b + 1) * S * S, sigma, G, Ln, in) && ok;
        if (wr) put_up<S>(scr, nqp, node, (b + 1) * E + OFF_G<S>, G);
        double Fn[S][S];
#pragma unroll
        for (int c = 0; c < S; c++)
#pragma unroll
            for (int i = 0; i < S; i++) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < S; k++) t -= G[i][k] * F[k][c];
#pragma unroll
                for (int k = 0; k < i; k++) t -= Ln[i][k] * Fn[k][c];
                Fn[i][c] = t * in[i];
            }
#pragma unroll
        for (int r = 0; r < S; r++) {
            inv[r] = in[r];
#pragma unroll
            for (int c = 0; c < S; c++) { L[r][c] = Ln[r][c]; F[r][c] = Fn[r][c]; }
        }
    }
    // backward sweep: Z_b = L_b^{-T} (F_b - G_{b+1}^T Z_{b+1}); the last
    // block's L, 1/diag and F are still in registers
    double Z[S][S], Gn[S][S];
#pragma unroll
    for (int r = 0; r < S; r++)
#pragma unroll
        for (int c = 0; c < S; c++) { Gn[r][c] = 0.0; Z[r][c] = 0.0; }
    for (int b = nb - 1; b >= 0; b--) {
        double Gp[S][S], Lp[S][S], Fp[S][S];
        if (!PF && b < nb - 1) {   // block b now (G_{b+1}, L_b, F_b)
            get_up<S>(scr, nqp, node, (b + 1) * E + OFF_G<S>, Gn);
            get_lo<S>(scr, nqp, node, b * E + OFF_L<S>, L);
            get<S>(scr, nqp, node, b * E + OFF_F<S>, F);
#pragma unroll
            for (int r = 0; r < S; r++) inv[r] = L[r][r];
        }
        if (PF && b > 0) {  // prefetch block b-1 while block b is solved
            get_up<S>(scr, nqp, node, b * E + OFF_G<S>, Gp);
            get_lo<S>(scr, nqp, node, (b - 1) * E + OFF_L<S>, Lp);
            get<S>(scr, nqp, node, (b - 1) * E + OFF_F<S>, Fp);
        }
        double Zn[S][S];
#pragma unroll
        for (int c = 0; c < S; c++)
#pragma unroll
            for (int ii = S - 1; ii >= 0; ii--) {
                double t = F[ii][c];
#pragma unroll
                for (int k = 0; k < S; k++) t -= Gn[k][ii] * Z[k][c];
#pragma unroll
                for (int k = ii + 1; k < S; k++) t -= L[k][ii] * Zn[k][c];
                Zn[ii][c] = t * inv[ii];
            }
#pragma unroll
        for (int r = 0; r < S; r++)
#pragma unroll
            for (int c = 0; c < S; c++) Z[r][c] = Zn[r][c];
        sink(b, Zn);
        if (b == 0) break;
        if (PF) {
#pragma unroll
            for (int r = 0; r < S; r++) {
#pragma unroll
                for (int c = 0; c < S; c++) { L[r][c] = Lp[r][c]; F[r][c] = Fp[r][c]; Gn[r][c] = Gp[r][c]; }
                inv[r] = Lp[r][r];
            }
        }
    }
    return ok;
}

// Sum over the 64 lanes of a wave of each entry of the S x S block v, by
// recursive halving: at the step with lane mask m a lane keeps half of its
// entries (the upper half where lane & m) plus the partner's copies of
// them; once one entry per lane is left, the remaining lane bits are summed
// by butterfly.  Lane l returns the wave's total of entry
// e = l >> (6 - log2 P) (entry index r + c S, P = S*S rounded up to a power
// of two; e >= S*S are padding), the same value in all 64/P lanes of its
// group.  The order of the additions is fixed.  Every lane must call.
template <int S>
constexpr int wave_sum_pad()
{
    int p = 1;
    while (p < S * S) p *= 2;
    return p;
}
template <int S>
__device__ inline double wave_block_sum(const double (&v)[S][S], int lane, int &entry)
{
    constexpr int P = wave_sum_pad<S>();
    constexpr int LOGP = P == 1 ? 0 : (P == 2 ? 1 : (P == 4 ? 2 : (P == 8 ? 3 : (P == 16 ? 4 : (P == 32 ? 5 : 6)))));
    double x[P];
#pragma unroll
    for (int e = 0; e < P; e++) x[e] = e < S * S ? v[e % S][e / S] : 0.0;
#pragma unroll
    for (int st = 0; st < LOGP; st++) {
        const int n = P >> st, m = 32 >> st;
        const bool hi = (lane & m) != 0;
#pragma unroll
        for (int i = 0; i < n / 2; i++) {
            const double send = hi ? x[i] : x[i + n / 2], keep = hi ? x[i + n / 2] : x[i];
            x[i] = keep + __shfl_xor(send, m, 64);
        }
    }
#pragma unroll
    for (int m = 32 >> LOGP; m > 0; m >>= 1) x[0] += __shfl_xor(x[0], m, 64);
    entry = lane >> (6 - LOGP);
    return x[0];
}

// Gershgorin bound of H_j by the whole block (any blockDim multiple of 64).
template <int S>
__device__ double gersh_block(const double *Ablk, const double *Bblk, int js, double *red)
{
    double mx = 0.0;
    for (int i = threadIdx.x; i < js; i += blockDim.x) {
        double rs = 0.0;
        for (int l = max(0, i - S); l <= min(js - 1, i + S); l++) {
            double h = (i >= l) ? h_entry<S>(Ablk, Bblk, js, i, l) : h_entry<S>(Ablk, Bblk, js, l, i);
            rs += fabs(h);
        }
        mx = fmax(mx, rs);
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = mx;
    __syncthreads();
    double M = red[0];
    for (int w = 1; w < (int)(blockDim.x / 64); w++) M = fmax(M, red[w]);
    __syncthreads();
    return M;
}

struct SmallOut {
    int nq;
    int err;        // 0 ok, 2 factorisation failed, 3 spectral range beyond the node table
    double lam_lo, lam_hi;
};

// The whole small problem of one Lanczos iteration, by one block:
// M = Gershgorin bound; m = the largest M 2^-(b+1), b < nbound, with H - m I
// positive definite (one thread per probe); DE quadrature
//   H^{-1/2} X = (2/pi) int_0^inf (H + t^2 I)^{-1} X dt,  t = e^u,
//   u = u0 + sinh(v), u0 = (ln m + ln M) / 4, nodes v = k h, |k| <= N,
// one thread per node; Z = sum of w Z_k; C = H Z (so y = V C = V sqrt(H) E1 R0).
// C is written to Cout (js x S, ld js).  scr holds scratch_doubles(S, nb, nqp)
// doubles; nq <= nqp and nq <= blockDim.x (a multiple of 64).
//
// hint (optional, one int kept between the iterations of one solve, -1 at
// its start): the probe index the previous iteration found.  With a hint the
// last wave tests the probes while the other waves solve the nodes placed
// from the hint plus one halving of margin (the smallest Ritz value only
// decreases with j); a placement whose range [m, M] contains the spectrum
// is as accurate as the tight one, so the nodes are solved again only when
// the probes find the spectrum below the placed range.
template <int S>
__device__ SmallOut small_problem(const double *__restrict__ Ag, const double *__restrict__ Bg,
                                  const double *__restrict__ R0g, int nb, int nbound, int dec_max,
                                  double *__restrict__ scr, int nqp, double *__restrict__ Cout,
                                  int *hint = nullptr)
{
    constexpr int E = NE<S>;
    const int js = nb * S, nwv = blockDim.x / 64;
    __shared__ double s_red[16];
    __shared__ int s_ok[64];
    __shared__ int s_fail;
    __shared__ double s_Z[512 * S];      // js <= 512
    // the blocks of H_j and R0 in LDS: the sequential block sweeps below
    // then wait on LDS, not on global memory, at every block
    __shared__ double sA[512 * S], sB[512 * S], sR[S * S];   // nb S^2 <= 512 S
    SmallOut out{0, 0, 0.0, 0.0};
    const int hv = hint ? *hint : -1;    // read by every thread before the first barrier
    for (int e = threadIdx.x; e < nb * S * S; e += blockDim.x) {
        sA[e] = Ag[e];
        sB[e] = Bg[e];
    }
    if ((int)threadIdx.x < S * S) sR[threadIdx.x] = R0g[threadIdx.x];
    __syncthreads();
    const double *A = sA, *B = sB, *R0 = sR;
    const double M = gersh_block<S>(A, B, js, s_red);
    out.lam_hi = M;
    // node placement of probe index kb: (m, N, h); returns dec (> dec_max:
    // beyond the table)
    auto place = [&](int kb, double &m, int &N, double &h) {
        m = ldexp(M, -(kb + 1));
        int dec = (int)ceil(log10(M / m));
        if (dec < 1) dec = 1;
        if (dec > dec_max) return dec_max + 1;
        N = c_qN[dec];
        h = c_qh[dec];
        return dec;
    };
    // one node per thread, whole waves: the lanes of the last wave past the
    // last node repeat it (no scratch writes) and add zeros to the wave sums
    const int node = threadIdx.x, wv = node / 64, lane = node % 64;
    double *wpart = scr + (long)nb * E * nqp;    // wave partial sums of Z, (waves) x (js S)
    auto solve_nodes = [&](double m, int N, double h) {
        const int nq = 2 * N + 1, nw = (nq + 63) / 64;
        if (wv >= nw) return;
        const int nd = min(node, nq - 1);
        const double u0 = 0.25 * (log(m) + log(M));
        const double v = (nd - N) * h;
        const double u = u0 + sinh(v);
        const double w = node < nq ? (2.0 / M_PI) * h * cosh(v) * exp(u) : 0.0;
        constexpr int GRP = 64 / wave_sum_pad<S>();
        auto sink = [&](int b, const double (&z)[S][S]) {
            double wz[S][S];
#pragma unroll
            for (int r = 0; r < S; r++)
#pragma unroll
                for (int c = 0; c < S; c++) wz[r][c] = w * z[r][c];
            int e;
            const double t = wave_block_sum<S>(wz, lane, e);
            if (lane % GRP == 0 && e < S * S) wpart[(long)wv * js * S + b * S * S + e] = t;
        };
        if (!node_solve<S>(A, B, R0, nb, exp(2.0 * u), scr, nqp, nd, node < nq, sink)) s_fail = 1;
    };
    auto fits = [&](int N) { return 2 * N + 1 <= nqp && 2 * N + 1 <= (int)blockDim.x; };
    // the guessed placement, solved beside the probes when it leaves the
    // last wave free
    int kb_guess = -1, N_g = 0;
    double m_g = 0.0, h_g = 0.0;
    if (hint && nbound <= 64 && nwv >= 2) {
        kb_guess = hv >= 0 ? min(hv + 1, nbound - 1) : 0;
        if (place(kb_guess, m_g, N_g, h_g) > dec_max || 2 * N_g + 1 > 64 * (nwv - 1) || !fits(N_g)) kb_guess = -1;
    }
    if (threadIdx.x == 0) s_fail = 0;
    __syncthreads();
    if (kb_guess >= 0) {
        if (wv == nwv - 1) {
            if (lane < nbound) s_ok[lane] = factor_ok<S>(A, B, nb, -ldexp(M, -(lane + 1))) ? 1 : 0;
        } else {
            solve_nodes(m_g, N_g, h_g);
        }
    } else if ((int)threadIdx.x < nbound) {
        s_ok[threadIdx.x] = factor_ok<S>(A, B, nb, -ldexp(M, -(int)(threadIdx.x + 1))) ? 1 : 0;
    }
    __syncthreads();
    int kb = -1;
    for (int b = 0; b < nbound; b++)
        if (s_ok[b]) { kb = b; break; }
    if (kb < 0) { out.err = 2; return out; }
    if (hint && threadIdx.x == 0) *hint = kb;
    // the guess stands when its range reaches below the bound
    const int use = (kb_guess >= 0 && kb <= kb_guess) ? kb_guess : kb;
    double m, h;
    int N;
    if (place(use, m, N, h) > dec_max) { out.err = 3; return out; }
    const int nq = 2 * N + 1;
    out.lam_lo = m;
    out.nq = nq;
    if (!fits(N)) { out.err = 3; return out; }
    if (use != kb_guess) {
        if (threadIdx.x == 0) s_fail = 0;
        __syncthreads();
        solve_nodes(m, N, h);
        __syncthreads();
    }
    const int nw = (nq + 63) / 64;
    if (s_fail) { out.err = 2; return out; }
    // Z = sum of the wave partials (wave order), into LDS
    for (int e = threadIdx.x; e < js * S; e += blockDim.x) {
        double t = 0.0;
        for (int q = 0; q < nw; q++) t += wpart[(long)q * js * S + e];
        const int b = e / (S * S), rc = e % (S * S);     // rc = r + c S inside block b
        const int r = rc % S, c = rc / S;
        s_Z[(b * S + r) + c * js] = t;
    }
    __syncthreads();
    // C = H Z (banded)
    for (int e = threadIdx.x; e < js * S; e += blockDim.x) {
        const int i = e % js, c = e / js;
        double v = 0.0;
        for (int l = max(0, i - S); l <= min(js - 1, i + S); l++) {
            double hv = (i >= l) ? h_entry<S>(A, B, js, i, l) : h_entry<S>(A, B, js, l, i);
            v += hv * s_Z[l + c * js];
        }
        Cout[e] = v;
    }
    return out;
}

// ------------------------------------------------------------ split small problem
// The same small problem spread over several workgroups of two waves:
// wave 0 tests the spectral probes while wave 1 solves this workgroup's
// share of the quadrature nodes, with the band of H_j and every node's
// factor blocks in LDS (so the sequential block sweeps see LDS, not global,
// latency).  The nodes are placed while the probes run, from the largest
// probe index the previous launch found (the smallest Ritz value only
// decreases with j) plus one halving of margin: a placement whose range
// [m, M] contains the spectrum is as accurate as the tight one, so the
// nodes are solved again only when the probes find the spectrum below the
// placed range.  Hints are double-buffered by launch parity (read kb[p ^ 1],
// write kb[p]), so the placement does not depend on the order in which the
// members of one launch finish.  Each workgroup leaves the weighted sum of
// its nodes' solutions in part[wg]; the last one to finish sums the parts
// in workgroup order and forms C = H Z.
template <int S>
constexpr int WG_NBM = 208 / S;      // largest H_j block count served (order <= 208)
template <int S>
constexpr int WG_NPB = S == 1 ? 16 : (S == 2 ? 10 : (S == 3 ? 8 : 6));   // nodes per workgroup
constexpr int WG_T = 128;            // threads: probes in wave 0, nodes in wave 1

// Row i of H (block tridiagonal: A_b symmetric, lower triangle stored;
// H(bS + r, (b-1)S + k) = B_{b-1}(r, k); H(bS + r, (b+1)S + k) = B_b(k, r))
// times the column z, in the order: previous block, diagonal, next block.
template <int S>
__device__ inline double block_row_dot(const double *__restrict__ A, const double *__restrict__ B, int nb, int i,
                                       const double *__restrict__ z)
{
    const int b = i / S, r = i % S;
    const double *Ab = A + b * S * S;
    double v = 0.0;
    if (b > 0) {
        const double *Bp = B + (b - 1) * S * S;
#pragma unroll
        for (int k = 0; k < S; k++) v += Bp[r + k * S] * z[(b - 1) * S + k];
    }
#pragma unroll
    for (int k = 0; k < S; k++) v += (k <= r ? Ab[r + k * S] : Ab[k + r * S]) * z[b * S + k];
    if (b + 1 < nb) {
        const double *Bb = B + b * S * S;
#pragma unroll
        for (int k = 0; k < S; k++) v += Bb[k + r * S] * z[(b + 1) * S + k];
    }
    return v;
}

struct SplitState {
    int cnt;        // workgroups finished (the last resets it)
    int kb[2];      // probe index of the lower bound, by launch parity (-1: none)
    int nq, err;
    double lam_lo, lam_hi;
};

// Returns true (uniformly) in the workgroup that finished last and wrote
// C and the SplitState fields.
template <int S>
__device__ bool small_problem_split(const double *__restrict__ Ag, const double *__restrict__ Bg,
                                    const double *__restrict__ R0, int nb, int nbound, int dec_max,
                                    double *__restrict__ part, double *__restrict__ Cout, SplitState *ss,
                                    const SplitState *hints, int nhints, int parity)
{
    constexpr int E = NE<S>, NBM = WG_NBM<S>, NPB = WG_NPB<S>;
    __shared__ double sA[NBM * S * S], sB[NBM * S * S], sR[S * S];
    __shared__ double sF[NBM * E * NPB];
    __shared__ double s_red[WG_T / 64];
    __shared__ int s_ok[64];
    __shared__ int s_fail, s_last;
    const int js = nb * S, t = threadIdx.x, wv = t / 64, lane = t % 64;
    for (int e = t; e < nb * S * S; e += WG_T) {
        sA[e] = Ag[e];
        sB[e] = Bg[e];
    }
    if (t < S * S) sR[t] = R0[t];
    if (t == 0) s_fail = 0;
    __syncthreads();
    const double M = gersh_block<S>(sA, sB, js, s_red);
    // node placement for a probe index kb: (m, N, h, nq)
    auto place = [&](int kb, double &m, int &N, double &h) {
        m = ldexp(M, -(kb + 1));
        int dec = (int)ceil(log10(M / m));
        if (dec < 1) dec = 1;
        if (dec > dec_max) dec = dec_max + 1;
        N = c_qN[min(dec, 13)];
        h = c_qh[min(dec, 13)];
        return dec;
    };
    int hint = -1;
    for (int g = 0; g < nhints; g++) hint = max(hint, hints[g].kb[parity ^ 1]);
    const int kb_guess = hint >= 0 ? min(hint + 1, nbound - 1) : 0;
    auto solve_nodes = [&](int kb) {
        double m, h;
        int N;
        const int dec = place(kb, m, N, h);
        const int nq = 2 * N + 1;
        if (dec > dec_max) return;
        const int node = blockIdx.x * NPB + lane;
        if (lane < NPB && node < nq) {
            const double u0 = 0.25 * (log(m) + log(M));
            const double v = (node - N) * h;
            const double u = u0 + sinh(v);
            const double w = (2.0 / M_PI) * h * cosh(v) * exp(u);
            auto sink = [&](int b, const double (&z)[S][S]) {
                double wz[S][S];
#pragma unroll
                for (int r = 0; r < S; r++)
#pragma unroll
                    for (int c = 0; c < S; c++) wz[r][c] = w * z[r][c];
                put<S>(sF, NPB, lane, b * E + OFF_F<S>, wz);   // over F_b, read no more
            };
            if (!node_solve<S, false>(sA, sB, sR, nb, exp(2.0 * u), sF, NPB, lane, true, sink)) s_fail = 1;
        }
    };
    if (wv == 0) {
        if (lane < nbound) s_ok[lane] = factor_ok<S>(sA, sB, nb, -ldexp(M, -(lane + 1))) ? 1 : 0;
    } else {
        solve_nodes(kb_guess);
    }
    __syncthreads();
    int kb = -1;
    for (int b = 0; b < nbound; b++)
        if (s_ok[b]) { kb = b; break; }
    int err = 0;
    if (kb < 0) err = 2;
    double m = 0.0, h = 0.0;
    int N = 0;
    // the guessed placement stands when its range reaches below the bound
    // and its node count is in the table
    int use = kb;
    if (kb >= 0 && kb <= kb_guess) {
        double mg, hg;
        int Ng;
        if (place(kb_guess, mg, Ng, hg) <= dec_max) use = kb_guess;
    }
    const int dec = kb < 0 ? 0 : place(use, m, N, h);
    if (!err && dec > dec_max) err = 3;
    const int nq = 2 * N + 1;
    const int nact = err ? 1 : (nq + NPB - 1) / NPB;
    if ((int)blockIdx.x >= nact) return false;   // no nodes here: not counted
    if (!err && use != kb_guess) {
        if (t == 0) s_fail = 0;
        __syncthreads();
        if (wv == 1) solve_nodes(use);
        __syncthreads();
    }
    if (s_fail && !err) err = 2;
    // this workgroup's part: sum over its node slots
    const int nmine = err ? 0 : max(0, min(NPB, nq - (int)blockIdx.x * NPB));
    double *pw = part + (long)blockIdx.x * js * S;
    for (int e = t; e < js * S; e += WG_T) {
        const int b = e / (S * S), rc = e % (S * S);
        double acc = 0.0;
        for (int k = 0; k < nmine; k++) acc += sF[(b * E + OFF_F<S> + rc) * NPB + k];
        store_dev(pw + e, acc);
    }
    stores_complete();
    __syncthreads();
    if (t == 0) s_last = (atomicAdd(&ss->cnt, 1) == nact - 1) ? 1 : 0;
    __syncthreads();
    if (!s_last) return false;
    if (t == 0) {
        ss->cnt = 0;
        ss->kb[parity] = kb;
        ss->nq = nq;
        ss->err = err;
        ss->lam_lo = m;
        ss->lam_hi = M;
    }
    if (err) return true;
    // Z (block-major, S x S blocks col-major) summed over parts, into sF[0..)
    double *sZ = sF;
    for (int e = t; e < js * S; e += WG_T) {
        const double acc = sum_strided<true>(part + e, (long)js * S, nact);
        const int b = e / (S * S), rc = e % (S * S), r = rc % S, c = rc / S;
        sZ[(b * S + r) + c * js] = acc;
    }
    __syncthreads();
    for (int e = t; e < js * S; e += WG_T) {
        const int i = e % js, c = e / js;
        Cout[e] = block_row_dot<S>(sA, sB, nb, i, sZ + (long)c * js);
    }
    return true;
}

}  // namespace sp

}  // namespace bd
