// bd_matfree_dev.h -- device pieces of the matrix-free RPY products shared
// by the single-chain product (bd_matfree.hip) and the batched one
// (bd_batch.hip).
//
// Per ordered bead pair (i, j), r = r_j - r_i:
//   D_ij X_j = c1 X_j + c2 r (r . X_j),
//   r >= 2:  c1 = 3/(4r) + 1/(2r^3),   c2 = 3/(4r^3) - 3/(2r^5)
//   r <  2:  c1 = 1 - 9r/32,           c2 = 3/(32r)
//   i == j:  c1 = 1, c2 = 0
// (bd_solver_m.f90:484-504 with a = 1; c2 multiplies r r^T, not r^ r^T).
//
// A workgroup of MF_WAVES waves owns 64 row beads (one per lane, the same
// rows in every wave); its column beads are staged MF_TILE at a time in LDS
// (coordinates in FP64 and the 3 x S block of X in T) and split between the
// waves, every lane of a wave on the same column bead, so that each record
// is an LDS broadcast read.  The waves' sums are added in LDS in wave order.
#pragma once

#include "bd_internal.h"

namespace bd {

constexpr int MF_CH = 8;        // column beads per FP32 partial sum (and record padding)
constexpr int MF_WAVES = 8;     // waves per workgroup (column-range split)
constexpr int MF_TILE = 256;    // column beads staged in LDS at a time

template <typename T>
__device__ inline T mf_rsqrt(T v);
template <>
__device__ inline float mf_rsqrt<float>(float v)
{
    return __builtin_amdgcn_rsqf(v);
}
template <>
__device__ inline double mf_rsqrt<double>(double v)
{
    return rsqrt_f64(v);
}

// r^2 is offset by this much before 1/sqrt: a bead's pair with itself then
// gives finite overlap coefficients (c1 = 1, the r r^T term multiplying a
// zero vector) without a special case.  The offset changes r^2 only for
// two beads closer than about 1e-7.
template <typename T>
__device__ constexpr T mf_r2_floor();
template <>
__device__ constexpr double mf_r2_floor<double>() { return 1e-30; }
template <>
__device__ constexpr float mf_r2_floor<float>() { return 1e-30f; }

// Coordinates of padding bead j >= nb (records are padded to a multiple of
// MF_CH beads): points 1e6 apart from everything, whose X block is zero.
__device__ inline void mf_far_point(int j, int nb, double &x, double &y, double &z)
{
    x = 1.0e6 * (j - nb + 2);
    y = 1.0e6;
    z = 1.0e6;
}

// LDS record of one column bead: its coordinates and its 3 x S block of X.
template <typename T, int S>
struct alignas(16) MfRec {
    double x, y, z;
    T v[3][S];
};

// The pair term (D_ij X_j)(a, k) added to acc[a][k] for the row bead at
// (xi, yi, zi): both coefficient branches are formed and the one that
// applies selected (no branch); j == i takes the overlap branch at r = 0,
// where c1 = 1 and the r r^T term vanishes, which is the unit diagonal block.
template <typename T, int S, typename A>
__device__ __forceinline__ void mf_pair(const MfRec<T, S> &cur, double xi, double yi, double zi, A (&acc)[3][S])
{
    const T dx = (T)(cur.x - xi), dy = (T)(cur.y - yi), dz = (T)(cur.z - zi);
    const T r2 = dx * dx + dy * dy + dz * dz;
    const T y = mf_rsqrt<T>(r2 + mf_r2_floor<T>()), y2 = y * y;
    const T c1f = y * ((T)0.75 + (T)0.5 * y2);
    const T c2f = y * y2 * ((T)0.75 - (T)1.5 * y2);
    const T c1n = (T)1 - (T)(9.0 / 32.0) * (r2 * y), c2n = (T)(3.0 / 32.0) * y;
    const bool nr = r2 < (T)4;
    const T c1 = nr ? c1n : c1f, c2 = nr ? c2n : c2f;
#pragma unroll
    for (int k = 0; k < S; k++) {
        const T t = c2 * (dx * cur.v[0][k] + dy * cur.v[1][k] + dz * cur.v[2][k]);
        acc[0][k] = fma(c1, cur.v[0][k], acc[0][k]);
        acc[0][k] = fma(t, dx, acc[0][k]);
        acc[1][k] = fma(c1, cur.v[1][k], acc[1][k]);
        acc[1][k] = fma(t, dy, acc[1][k]);
        acc[2][k] = fma(c1, cur.v[2][k], acc[2][k]);
        acc[2][k] = fma(t, dz, acc[2][k]);
    }
}

// LDS a workgroup of the product needs: the staged records, and after the
// last tile the waves' sums (the same bytes).
template <typename T, int S, int NW = MF_WAVES>
constexpr int mf_smem_bytes()
{
    constexpr int red = NW * 64 * (3 * S + 1) * (int)sizeof(double);
    constexpr int til = MF_TILE * (int)sizeof(MfRec<T, S>);
    return red > til ? red : til;
}

// acc += sum over the column chunks [cb, ce) (MF_CH beads each) of the pair
// terms of this lane's row bead (xi, yi, zi), by the whole workgroup of NW
// waves: tile by tile, stage(e, j, rec) fills record e of the tile for
// column bead j, then wave w takes its share of the tile's chunks.  FP32
// terms are summed over each chunk in FP32, then added in FP64.  Every
// thread must call.
template <typename T, int S, int NW = MF_WAVES, typename Stage>
__device__ __forceinline__ void mf_rows_times_cols(double xi, double yi, double zi, int cb, int ce, MfRec<T, S> *tile,
                                                   int w, Stage stage, double (&acc)[3][S])
{
    for (int t0 = cb; t0 < ce; t0 += MF_TILE / MF_CH) {
        const int nct = min(ce - t0, MF_TILE / MF_CH), jb = t0 * MF_CH;
        for (int e = threadIdx.x; e < nct * MF_CH; e += blockDim.x) stage(e, jb + e, tile[e]);
        __syncthreads();
        const int qb = nct * w / NW, qe = nct * (w + 1) / NW;
        if constexpr (sizeof(T) == 8) {
            for (int q = MF_CH * qb; q < MF_CH * qe; q++) mf_pair<T, S>(tile[q], xi, yi, zi, acc);
        } else {
            for (int ch = qb; ch < qe; ch++) {
                T cacc[3][S];
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int k = 0; k < S; k++) cacc[a][k] = (T)0;
#pragma unroll 2
                for (int q = 0; q < MF_CH; q++) mf_pair<T, S>(tile[MF_CH * ch + q], xi, yi, zi, cacc);
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int k = 0; k < S; k++) acc[a][k] += (double)cacc[a][k];
            }
        }
        __syncthreads();
    }
}

// The waves' sums of acc, added in wave order through red (the staging
// bytes, free after mf_rows_times_cols), handed to out(l, a S + k, value)
// for every lane l of the row block.  Every thread must call.
template <int S, int NW = MF_WAVES, typename Out>
__device__ __forceinline__ void mf_wave_sums(const double (&acc)[3][S], double (*red)[64][3 * S + 1], int w, int lane,
                                             Out out)
{
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int k = 0; k < S; k++) red[w][lane][a * S + k] = acc[a][k];
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 3 * S; e += blockDim.x) {
        const int l = e / (3 * S), col = e % (3 * S);
        double v = red[0][l][col];
#pragma unroll
        for (int q = 1; q < NW; q++) v += red[q][l][col];
        out(l, col, v);
    }
    __syncthreads();
}

}  // namespace bd
