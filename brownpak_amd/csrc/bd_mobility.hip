// bd_mobility.hip -- RPY mobility assembly and the symmetric D.V product.
//
// Replaces calc_rpy_tensor (bd_solver_m.f90:458-522) and the dsymv/dsymm('L')
// products (bd_solver_m.f90:446, brown_m.f90:232, 386).
//
// Storage: only the lower triangle of tiles is kept (half the bytes of the
// reference's full n x n matrix).  One product pass streams every stored tile
// once from HBM and uses it twice: tile (I,J) adds D_IJ V_J to output block I
// ("row product") and D_IJ^T V_I to output block J ("column product").  Each
// work unit covers a 2-row x 4-column block of tiles; row products are summed
// over the unit's 4 tile columns in registers before one partial is written
// per (tile row, super-column), and column products are summed over the unit's
// tile rows.  A second kernel adds the partials of every output block in a
// fixed order, so results are deterministic run to run.
#include <map>
#include <mutex>
#include <string>
#include <type_traits>

#include "bd_internal.h"
#include "bd_lanczos_dev.h"

namespace bd {

// Work unit: CHR tile rows x WSC tile columns (a "super-column" is WSC tile
// columns).  2 x 4 is the default and the only shape for blocks wider than 4
// columns; the others serve blocks of at most 4 columns.  A single chain's
// plan takes its shape and super-column widths from plan_pick, a batch's
// shape from plan_tune.
const int kTileShapes[BD_TILE_SHAPES][2] = {{2, 4}, {1, 4}, {4, 4}, {2, 2}, {2, 5}, {2, 6}, {2, 8}};

bool tile_shape_supported(int chr, int wsc)
{
    for (const auto &s : kTileShapes)
        if (s[0] == chr && s[1] == wsc) return true;
    return false;
}

constexpr int NT = 256;  // threads per block of the product kernel
// Work-unit record (16 ints): chain, super-column, Ib, Ie, ntile, nsc,
// valid-slot mask, first column J0 of the super-column, tile base (elements,
// 2 ints), x row offset (2 ints), row-partial offset (2 ints), canonical
// index (the unit's place in super-column order, which indexes its column
// partials), width of the super-column (<= WSC).
constexpr int UREC = 16;

__host__ __device__ inline long tile_index(int I, int J) { return (long)I * (I + 1) / 2 + J; }

// ---------------------------------------------------------------- assembly
// One block per stored tile.  Bead-pair 3x3 blocks for the <= 22 x 22 bead
// pairs the tile touches are computed once into LDS (6 unique entries each),
// then the 64 x 64 entries are written column-major (coalesced along rows).
template <typename E>
__global__ __launch_bounds__(256) void rpy_tile_kernel(const double *__restrict__ crd_all, const int *__restrict__ tiles,
                                                       const int *__restrict__ chain_i, const long *__restrict__ chain_l,
                                                       E *__restrict__ mob_all)
{
    __shared__ double sblk[22 * 22 * 6];
    const int chain = tiles[4 * blockIdx.x], I = tiles[4 * blockIdx.x + 1], J = tiles[4 * blockIdx.x + 2];
    const int nb = chain_i[4 * chain], n = chain_i[4 * chain + 1];
    const double *crd = crd_all + chain_l[3 * chain];
    E *mob = mob_all + chain_l[3 * chain + 1] * (TILE * TILE);
    const int gi0 = I * TILE, gj0 = J * TILE;
    const int bi0 = gi0 / 3, bj0 = gj0 / 3;
    const int nbi = min((gi0 + TILE - 1) / 3, nb - 1) - bi0 + 1;
    const int nbj = min((gj0 + TILE - 1) / 3, nb - 1) - bj0 + 1;

    for (int p = threadIdx.x; p < 22 * 22; p += blockDim.x) {
        int a = p / 22, b = p % 22;
        double *o = sblk + p * 6;
        if (a >= nbi || b >= nbj) continue;
        int bi = bi0 + a, bj = bj0 + b;
        if (bi == bj) {
            o[0] = 1.0; o[1] = 1.0; o[2] = 1.0; o[3] = 0.0; o[4] = 0.0; o[5] = 0.0;
            continue;
        }
        // rij = r_j - r_i; the block is even in rij, so the order is immaterial
        double rx = crd[3 * bj] - crd[3 * bi];
        double ry = crd[3 * bj + 1] - crd[3 * bi + 1];
        double rz = crd[3 * bj + 2] - crd[3 * bi + 2];
        double rijm = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(rx, rx), __dmul_rn(ry, ry)), __dmul_rn(rz, rz)));
        double irijm = 1.0 / rijm;
        double irijm2 = irijm * irijm;
        double c1, c2, cons;
        if (rijm >= 2.0) {
            c1 = 1.0 + (2.0 / 3.0) * irijm2;
            c2 = 1.0 - 2.0 * irijm2;
            cons = 0.75 * irijm;
        } else {
            c1 = 1.0 - 9.0 * rijm / 32.0;
            c2 = 3.0 * rijm / 32.0;
            cons = 1.0;
        }
        o[0] = cons * (c1 + c2 * rx * rx * irijm2);
        o[1] = cons * (c1 + c2 * ry * ry * irijm2);
        o[2] = cons * (c1 + c2 * rz * rz * irijm2);
        o[3] = cons * (c2 * rx * ry * irijm2);
        o[4] = cons * (c2 * rx * rz * irijm2);
        o[5] = cons * (c2 * ry * rz * irijm2);
    }
    __syncthreads();
    E *out = mob + tile_index(I, J) * (TILE * TILE);
    for (int e = threadIdx.x; e < TILE * TILE; e += blockDim.x) {
        int c = e / TILE, r = e % TILE;
        int gi = gi0 + r, gj = gj0 + c;
        double v = 0.0;
        if (gi < n && gj < n) {
            int a = gi / 3 - bi0, b = gj / 3 - bj0;
            int da = gi % 3, db = gj % 3;
            const double *o = sblk + (a * 22 + b) * 6;
            if (da == db) v = o[da];
            else {
                int lo = min(da, db), hi = max(da, db);
                v = o[lo == 0 ? (hi == 1 ? 3 : 4) : 5];
            }
        }
        out[e] = (E)v;
    }
}

// --------------------------------------------------------------- product
// Blocks of 256 threads take work units (CHR tile rows x WSC tile columns of
// one chain) in turn, each described by one precomputed record.  Thread t
// holds a 4 x 4 piece of every tile it touches -- rows 4 (t % 16) .. + 3 of
// columns t / 16 + 16 i, i = 0..3 -- loaded as four 16-byte (FP32) or eight
// 16-byte (FP64) pieces, and computes both products of the tile from its
// registers:
//   row product    y_I[r] += sum_c D(r, c) x_J[c]   (over its 4 columns),
//   column product z_J[c] += sum_r D(r, c) x_I[r]   (over its 4 rows),
// so tiles never pass through shared memory and no barrier separates them.
// The tiles of a unit are visited column by column (PD tiles in flight: the
// registers of a tile are refilled with the tile PD slots ahead as soon as
// its values are converted to FP64).  Column sums are complete after the
// unit's CHR rows and are summed over the 16 row pieces through shared
// memory (one partial per unit and tile column); row sums accumulate over
// the WSC columns and are summed over the 16 column pieces at the end (one
// partial per tile row and super-column).  A second kernel adds the partials
// of every output block in a fixed order, so results are deterministic run
// to run.  rowpart: (ntile * nsc) x TILE x S; colpart: nunits x WSC x TILE x S.
constexpr int TPT = TILE * TILE / NT;   // tile elements per thread (16)
// tiles in flight per block (register sets) and blocks per CU
template <typename E, int S>
constexpr int PDT = (sizeof(E) == 8 || S > 4) ? 2 : 3;
template <typename E, int S, int CHR>
constexpr int OCC = (S > 4 || CHR > 2) ? 1 : 2;   // register budget 256 (512 when 1)

// Register slot 4 i + q of thread tid holds tile element (row 4 (tid % 16) + q,
// column tid / 16 + 16 i), i.e. element 4 (tid + 256 i) + q of the
// column-major tile.
template <typename E>
__device__ inline void load_tile16(const E *__restrict__ tp, int tid, E (&t)[TPT])
{
    if constexpr (sizeof(E) == 4) {
#pragma unroll
        for (int i = 0; i < TPT / 4; i++) {
            const float4 v = reinterpret_cast<const float4 *>(tp)[tid + i * NT];
            t[4 * i] = v.x;
            t[4 * i + 1] = v.y;
            t[4 * i + 2] = v.z;
            t[4 * i + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < TPT / 4; i++) {
            const double2 a = reinterpret_cast<const double2 *>(tp)[2 * (tid + i * NT)];
            const double2 b = reinterpret_cast<const double2 *>(tp)[2 * (tid + i * NT) + 1];
            t[4 * i] = a.x;
            t[4 * i + 1] = a.y;
            t[4 * i + 2] = b.x;
            t[4 * i + 3] = b.y;
        }
    }
}

template <typename E, int S, int CHR, int WSC>
__global__ __launch_bounds__(NT, (OCC<E, S, CHR>)) void symm_tiles_kernel(
    const E *__restrict__ mob_all, const double *__restrict__ x_all, int ldx, const int *__restrict__ units,
    int nunits, const int *__restrict__ chain_i, const long *__restrict__ chain_l, double *__restrict__ rowpart,
    double *__restrict__ colpart, int *__restrict__ ticket, const DevStatus *stop, const int *__restrict__ active,
    XPre pre)
{
    constexpr int NS = CHR * WSC, PD = PDT<E, S>, LDR = TILE * S + 2;
    // x_J[c][k] at xJs[jj][k][(c % 16) * 4 + c / 16] (a thread's 4 columns
    // c = cg + 16 i are adjacent); x_I[r][k] at xIs[ii][k][r] (its 4 rows are)
    __shared__ double xJs[WSC][S][TILE];
    __shared__ double xIs[CHR][S][TILE];
    __shared__ double sred[16 * LDR];
    const int tid = threadIdx.x, cg = tid / 16, rg = tid % 16;
    (void)chain_i;
    (void)chain_l;
    (void)ticket;
    // inside a Lanczos solve: nothing to do once it has converged (the
    // status words are read together with the first unit's record, and
    // tested once both have arrived)
    const int st_conv = stop ? stop->converged_j : 0, st_err = stop ? (stop->ierr | stop->fail_j) : 0;
    // second CholeskyQR pass of the block (XPre): R2 from the Gram partials,
    // by every block (the same fixed-order sum and Cholesky as cholqr_kernel;
    // done before the unit loop: inside it, with the first tiles' registers
    // live, the FP64 variants spill)
    __shared__ double sR2[S * S];
    __shared__ int s_fail;
    if (pre.gpart && st_conv == 0 && st_err == 0) {
        __shared__ double sG[S * S];
        sum_partials_n<S>(pre.gpart, pre.ng, sG);   // ends with a barrier
        if (tid == 0) {
            double R[S][S];
            const bool ok = small_chol<S>(sG, R);
            s_fail = ok ? 0 : 1;
            for (int i = 0; i < S; i++)
                for (int j = 0; j < S; j++) sR2[i + j * S] = R[i][j];
            if (ok && blockIdx.x == 0) {
                // B = R2 R1 (both upper triangular)
                for (int i = 0; i < S; i++)
                    for (int j = 0; j < S; j++) {
                        double t = 0.0;
                        for (int k = 0; k < S; k++) t += R[i][k] * pre.R1[k + j * S];
                        pre.Bout[i + j * S] = t;
                    }
            }
        }
        __syncthreads();
        if (s_fail) {
            if (blockIdx.x == 0 && tid == 0) lz_factor_failed(pre.st, pre.blk);
            return;
        }
    }
    // units are dealt out statically; one record load describes a unit
    for (int u = blockIdx.x; u < nunits; u += gridDim.x) {
        const int *rec = units + (long)u * UREC;
        const int chain = rec[0], sc = rec[1], Ib = rec[2], Ie = rec[3], ntile = rec[4], nsc = rec[5];
        const unsigned mask = (unsigned)rec[6];   // bit ii * WSC + jj: tile (Ib + ii, J0 + jj) stored
        const long tbase = (long)(unsigned)rec[8] | ((long)rec[9] << 32);
        const long xoff = (long)(unsigned)rec[10] | ((long)rec[11] << 32);
        const long rpo = (long)(unsigned)rec[12] | ((long)rec[13] << 32);
        const int J0 = rec[7], cid = rec[14], wdt = rec[15];
        if (st_conv != 0 || st_err != 0) return;
        if (active) {
            // one thread reads the flag (another stream may clear it), so the
            // whole block takes the same branch
            __shared__ int s_act;
            __syncthreads();
            if (tid == 0) s_act = active[chain];
            __syncthreads();
            if (!s_act) continue;
        }
        const double *x = x_all + xoff;
        const E *mob = mob_all + tbase;
        // slot sl = jj * CHR + ii (column by column)
        auto present = [&](int sl) __attribute__((always_inline)) {
            return (mask >> ((sl % CHR) * WSC + sl / CHR)) & 1u;
        };
        auto tile_ptr = [&](int sl) __attribute__((always_inline)) {
            return mob + tile_index(Ib + sl % CHR, J0 + sl / CHR) * (TILE * TILE);
        };
        // x blocks (zeros past the chain's last tile; unconditional loads at
        // clamped addresses), issued ahead of the tiles so that they return
        // first, except where the registers they hold would spill
        constexpr int NX = (WSC + CHR) * TILE * S, XPT = (NX + NT - 1) / NT;
        constexpr bool XFIRST = !(sizeof(E) == 8 && S > 2 && CHR == 2 && WSC >= 4);
        double xv[XPT];
        auto load_x = [&]() __attribute__((always_inline)) {
#pragma unroll
            for (int i = 0; i < XPT; i++) {
                const int e = min(tid + i * NT, NX - 1);
                const int blk = e / (TILE * S), rem = e % (TILE * S), k = rem / TILE, r = rem % TILE;
                const int B = blk < WSC ? min(J0 + blk, ntile - 1) : min(Ib + blk - WSC, Ie - 1);
                xv[i] = x[(long)k * ldx + (long)B * TILE + r];
            }
        };
        if constexpr (XFIRST) load_x();
        E R[PD][TPT];
#pragma unroll
        for (int sl = 0; sl < PD; sl++)
            if (sl < NS && present(sl)) load_tile16<E>(tile_ptr(sl), tid, R[sl]);
        if constexpr (!XFIRST) load_x();
        __syncthreads();   // the previous unit no longer reads xJs / xIs / sred
#pragma unroll
        for (int i = 0; i < XPT; i++) {
            const int e = tid + i * NT;
            if (e < NX) {
                const int blk = e / (TILE * S), rem = e % (TILE * S), k = rem / TILE, r = rem % TILE;
                const bool in = blk < WSC ? (J0 + blk < ntile) : (Ib + blk - WSC < Ie);
                if (blk < WSC) xJs[blk][k][(r % 16) * 4 + r / 16] = in ? xv[i] : 0.0;
                else xIs[blk - WSC][k][r] = in ? xv[i] : 0.0;
            }
        }
        __syncthreads();
        if (pre.gpart) {
            // x rows q -> q R2^{-1} in place (the forward solve of cholqr_kernel)
            for (int rr = tid; rr < (WSC + CHR) * TILE; rr += NT) {
                const int blk = rr / TILE, r = rr % TILE;
                double *xs = blk < WSC ? &xJs[blk][0][(r % 16) * 4 + r / 16] : &xIs[blk - WSC][0][r];
                double q[S];
#pragma unroll
                for (int c = 0; c < S; c++) {
                    double t = xs[c * TILE];
#pragma unroll
                    for (int k = 0; k < c; k++) t -= q[k] * sR2[k + c * S];
                    q[c] = t / sR2[c + c * S];
                }
#pragma unroll
                for (int c = 0; c < S; c++) xs[c * TILE] = q[c];
            }
            __syncthreads();
            // the first super-column's units hold every row block once as
            // their row blocks: they write V
            if (sc == 0) {
                double *V = pre.V + xoff;
                for (int e = tid; e < CHR * TILE * S; e += NT) {
                    const int ii = e / (TILE * S), rem = e % (TILE * S), k = rem / TILE, r = rem % TILE;
                    if (Ib + ii < Ie) V[(long)k * ldx + (long)(Ib + ii) * TILE + r] = xIs[ii][k][r];
                }
            }
        }
        double racc[CHR][4][S];
#pragma unroll
        for (int a = 0; a < CHR; a++)
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int k = 0; k < S; k++) racc[a][q][k] = 0.0;
#pragma unroll
        for (int jj = 0; jj < WSC; jj++) {
            double cacc[4][S];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int k = 0; k < S; k++) cacc[i][k] = 0.0;
            // this tile column's x at the thread's 4 columns (adjacent slots),
            // held across the column's tiles unless registers are short
            constexpr bool HOLD = !(sizeof(E) == 8 && S > 2);
            double xj[S][4];
            if constexpr (HOLD) {
#pragma unroll
                for (int k = 0; k < S; k++)
#pragma unroll
                    for (int i = 0; i < 4; i++) xj[k][i] = xJs[jj][k][cg * 4 + i];
            }
#pragma unroll
            for (int ii = 0; ii < CHR; ii++) {
                const int sl = jj * CHR + ii;
                const bool here = present(sl);
                double d[4][4];   // d[i][q] = D(4 rg + q, cg + 16 i)
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int q = 0; q < 4; q++) d[i][q] = (double)R[sl % PD][4 * i + q];
                if (sl + PD < NS && present(sl + PD)) load_tile16<E>(tile_ptr(sl + PD), tid, R[sl % PD]);
                if (!here) continue;
                const bool offdiag = (Ib + ii) != (J0 + jj);
#pragma unroll
                for (int k = 0; k < S; k++) {
                    double xi[4];
#pragma unroll
                    for (int q = 0; q < 4; q++) xi[q] = xIs[ii][k][4 * rg + q];
                    if constexpr (!HOLD) {
#pragma unroll
                        for (int i = 0; i < 4; i++) xj[k][i] = xJs[jj][k][cg * 4 + i];
                    }
#pragma unroll
                    for (int q = 0; q < 4; q++)
#pragma unroll
                        for (int i = 0; i < 4; i++) racc[ii][q][k] += d[i][q] * xj[k][i];
                    if (offdiag) {
#pragma unroll
                        for (int i = 0; i < 4; i++)
#pragma unroll
                            for (int q = 0; q < 4; q++) cacc[i][k] += d[i][q] * xi[q];
                    }
                }
            }
            // column partial of tile column J0 + jj: the 16 row pieces summed
            // (columns past the super-column's width are never read)
            if (jj >= wdt) continue;
            __syncthreads();
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int k = 0; k < S; k++) sred[rg * LDR + (cg + 16 * i) * S + k] = cacc[i][k];
            __syncthreads();
            double *dst = colpart + ((long)cid * WSC + jj) * (TILE * S);
            for (int e = tid; e < TILE * S; e += NT) {
                double v = 0.0;
#pragma unroll
                for (int g = 0; g < 16; g++) v += sred[g * LDR + e];
                dst[e] = v;
            }
        }
        // row partials of the unit's tile rows: the 16 column pieces summed
#pragma unroll
        for (int ii = 0; ii < CHR; ii++) {
            if (Ib + ii >= Ie) continue;
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 4; q++)
#pragma unroll
                for (int k = 0; k < S; k++) sred[cg * LDR + (4 * rg + q) * S + k] = racc[ii][q][k];
            __syncthreads();
            double *dst = rowpart + (rpo + (long)(Ib + ii) * nsc + sc) * (TILE * S);
            for (int e = tid; e < TILE * S; e += NT) {
                double v = 0.0;
#pragma unroll
                for (int g = 0; g < 16; g++) v += sred[g * LDR + e];
                dst[e] = v;
            }
        }
    }
}

// Output block G (a tile row K of some chain, in super-column sc(K) at
// offset off(K)):
//   y = sum_{sc <= sc(K)} rowpart[K, sc] + sum_{units u of sc(K)} colpart[u, off(K)]
// rowblk (8 ints per block): chain, K, nsc, u0, u1 (canonical indices of
// sc(K)'s units), rp_off, y row offset, sc(K) | off(K) << 16.
// Sums in a fixed order (the row partials, then the column partials); the
// two lists are read as one, PB loads at a time at clamped addresses, so a
// batch is one memory round trip.
struct BlockRec {
    int K, nsc, u0, u1, rpo, yrow, sc, off;   // sc, off: K's super-column and its place in it
};
__device__ inline BlockRec block_rec(const int *__restrict__ rb)
{
    return {rb[1], rb[2], rb[3], rb[4], rb[5], rb[6], rb[7] & 0xffff, rb[7] >> 16};
}

template <int S>
__device__ inline double sum_block_partials(const double *__restrict__ rowpart, const double *__restrict__ colpart,
                                            const BlockRec &br, int e, int wsc, int half = -1)
{
    constexpr int PB = 32;
    const int K = br.K, nsc = br.nsc, u0 = br.u0, u1 = br.u1, rpo = br.rpo;
    const int nrow = br.sc + 1, nall = nrow + (u1 - u0);
    // half 0 / 1: the first / second half of the list only (half -1: all)
    const int ib = half == 1 ? (nall + 1) / 2 : 0, ntot = half == 0 ? (nall + 1) / 2 : nall;
    const double *rp = rowpart + ((long)rpo + (long)K * nsc) * (TILE * S) + e;
    const double *cp = colpart + ((long)u0 * wsc + br.off) * (TILE * S) + e;
    const long cstr = (long)wsc * (TILE * S);
    double v = 0.0;
    for (int i0 = ib; i0 < ntot; i0 += PB) {
        double t[PB];
#pragma unroll
        for (int q = 0; q < PB; q++) {
            const int i = min(i0 + q, ntot - 1);
            t[q] = i < nrow ? rp[(long)i * (TILE * S)] : cp[(long)(i - nrow) * cstr];
        }
#pragma unroll
        for (int q = 0; q < PB; q++)
            if (i0 + q < ntot) v += t[q];
    }
    return v;
}

template <int S>
__global__ __launch_bounds__(256) void symm_reduce_kernel(const double *__restrict__ rowpart, const double *__restrict__ colpart,
                                   const int *__restrict__ rowblk, int nrowblk, double *__restrict__ y, int ldy,
                                   int *__restrict__ ticket, const int *__restrict__ active, int wsc)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) *ticket = 0;  // re-arm the work counter for the next product
    if (g >= (long)nrowblk * TILE * S) return;
    const int G = (int)(g / (TILE * S));
    const int e = (int)(g % (TILE * S));
    const int *rb = rowblk + 8 * G;
    if (active && !active[rb[0]]) return;
    const int r = e / S, k = e % S;
    const BlockRec br = block_rec(rb);
    y[(long)k * ldy + br.yrow + r] = sum_block_partials<S>(rowpart, colpart, br, e, wsc);
}

// Thread sets of symm_reduce_lz_kernel: two (one per half of an output
// block's partial list, added first half + second half) where 2 TILE S
// threads fit the kernel's bound, else one.
template <int S>
constexpr int LZ_HALVES = (2 * TILE * S <= TILE * MAX_S) ? 2 : 1;

// The same sum for one output block per workgroup (LZ_HALVES<S> sets of
// TILE * S threads), fused with the first step of a Lanczos iteration:
//   W1 = W - V_{j-1} B_{j-1}^T  (stored over W),
//   partials of A = V_j^T W1 and P = W1^T W1 (2 S^2 per block, into lzpart).
template <int S>
__global__ __launch_bounds__(TILE * MAX_S) void symm_reduce_lz_kernel(
    const double *__restrict__ rowpart, const double *__restrict__ colpart, const int *__restrict__ rowblk,
    double *__restrict__ W, int ld, const double *__restrict__ Vj, const double *__restrict__ Vp,
    const double *__restrict__ Bp, double *__restrict__ lzpart, const DevStatus *stop, int wsc)
{
    constexpr int NO = 2 * S * S;                 // outputs per block
    constexpr int NGR = TILE * S / NO;            // row groups of the block sum
    constexpr int RPG = (TILE + NGR - 1) / NGR;   // rows per group (the last may be short)
    constexpr int HB = LZ_HALVES<S>;              // thread sets, one per half of the partial list
    __shared__ int s_done;
    __shared__ double sw[TILE][S];
    __shared__ double red[TILE][NO];
    __shared__ double red2[NGR][NO];
    __shared__ double s_hv[HB > 1 ? TILE * S : 1];
    const int G = blockIdx.x, e = threadIdx.x % (TILE * S), hf = threadIdx.x / (TILE * S);
    const bool act = hf == 0;                     // the set that finishes the block
    // the block's record and the status are read together
    const BlockRec br = block_rec(rowblk + 8 * G);
    if (threadIdx.x == 0) s_done = (stop && lz_done(stop)) ? 1 : 0;
    __syncthreads();
    if (s_done) return;
    const int r = e / S, k = e % S;
    const long row = br.yrow + r;
    // this row's vector entries first: their loads overlap the partial sums
    double vj[S], vp[S];
#pragma unroll
    for (int c = 0; c < S; c++) {
        vj[c] = act ? Vj[(long)c * ld + row] : 0.0;
        vp[c] = (act && Vp) ? Vp[(long)c * ld + row] : 0.0;
    }
    double v = sum_block_partials<S>(rowpart, colpart, br, e, wsc, HB > 1 ? hf : -1);
    if constexpr (HB > 1) {
        // first half + second half
        if (!act) s_hv[e] = v;
        __syncthreads();
        if (act) v = v + s_hv[e];
    }
    if (act) {
        if (Vp) {
            double t = 0.0;
#pragma unroll
            for (int kp = 0; kp < S; kp++) t += vp[kp] * Bp[k + kp * S];
            v = v - t;
        }
        W[(long)k * ld + row] = v;
        sw[r][k] = v;
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int c = 0; c < S; c++) {
            red[r][c + k * S] = vj[c] * v;             // A(c, k)
            red[r][S * S + c + k * S] = sw[r][c] * v;  // P(c, k)
        }
    }
    __syncthreads();
    if (act) {
        // column sums over the 64 rows: NGR groups of RPG rows, then the groups
        const int o = e % NO, gq = e / NO;
        if (gq < NGR) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < RPG; q++)
                if (gq * RPG + q < TILE) t += red[gq * RPG + q][o];
            red2[gq][o] = t;
        }
    }
    __syncthreads();
    if (act && e < NO) {
        double t = 0.0;
#pragma unroll
        for (int gq = 0; gq < NGR; gq++) t += red2[gq][e];
        lzpart[(long)G * NO + e] = t;
    }
}

// Expands the tiled lower storage to a dense n x n column-major matrix.
template <typename E>
__global__ void unpack_kernel(const E *__restrict__ mob, int n, double *__restrict__ dense, int ld)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long)n * n) return;
    int j = (int)(g / n), i = (int)(g % n);
    int gi = i, gj = j;
    if (gi < gj) { int t = gi; gi = gj; gj = t; }
    int I = gi / TILE, J = gj / TILE;
    int r = gi % TILE, c = gj % TILE;
    E v;
    if (I == J) v = mob[tile_index(I, J) * (TILE * TILE) + (long)(j % TILE) * TILE + (i % TILE)];
    else v = mob[tile_index(I, J) * (TILE * TILE) + (long)c * TILE + r];
    dense[(long)j * ld + i] = (double)v;
}

// ---------------------------------------------------------------- host side
// Tiles of the unit (rows Ib .. Ie-1, columns J0 .. J0+w-1) that are stored.
static int unit_tiles(int Ib, int Ie, int J0, int w)
{
    int t = 0;
    for (int I = Ib; I < Ie; I++) t += std::max(0, std::min(I, J0 + w - 1) - J0 + 1);
    return t;
}

// Super-column widths of one chain: `first` columns wide for the first `a`
// super-columns, then `rest` wide (the last one takes what is left).
static std::vector<int> make_widths(int nt, int first, int a, int rest)
{
    std::vector<int> w;
    for (int J0 = 0; J0 < nt;) {
        const int x = std::min((int)w.size() < a ? first : rest, nt - J0);
        w.push_back(x);
        J0 += x;
    }
    return w;
}

// Execution order of a single round of units on `slots` block slots of
// which block b and block b + slots/2 share a CU: the largest units run
// alone, the others in pairs of large with small.  Returns positions ->
// unit (canonical) indices; identity when the units do not fit one round.
static std::vector<int> pair_order(const std::vector<int> &tiles, int slots)
{
    const int U = (int)tiles.size(), M = slots / 2;
    std::vector<int> idx(U);
    for (int i = 0; i < U; i++) idx[i] = i;
    if (U <= M || U > 2 * M) return idx;
    std::vector<int> srt = idx;
    std::stable_sort(srt.begin(), srt.end(), [&](int x, int y) { return tiles[x] > tiles[y]; });
    const int k = U - M;   // CUs that get two units
    std::vector<int> pos(U);
    for (int i = 0; i < M - k; i++) pos[k + i] = srt[i];   // singles
    for (int j = 0; j < k; j++) {
        pos[j] = srt[M - k + j];        // larger of the pair
        pos[M + j] = srt[U - 1 - j];    // smaller of the pair
    }
    return pos;
}

// The largest per-CU cost of a single chain's units under pair_order, in
// tile-equivalents: 4 for a CU's start-up, 0.8 per tile on it (measured
// on MI355X: one unit of 10 tiles alone about 13 us, two on one CU about
// 19 us); more than one round costs the rounds times a full CU.
static double plan_cost(const std::vector<int> &tiles, int slots)
{
    const int U = (int)tiles.size(), M = slots / 2;
    int tmax = 0;
    for (int t : tiles) tmax = std::max(tmax, t);
    if (U > 2 * M) return (double)((U + 2 * M - 1) / (2 * M)) * (4.0 + 0.8 * 2 * tmax);
    const std::vector<int> pos = pair_order(tiles, slots);
    double worst = 0.0;
    for (int c = 0; c < std::min(U, M); c++) {
        int t = tiles[pos[c]];
        if (c + M < U) t += tiles[pos[c + M]];
        worst = std::max(worst, 4.0 + 0.8 * t);
    }
    return worst;
}

static std::vector<int> chain_unit_tiles(int nt, int chr, const std::vector<int> &w)
{
    std::vector<int> t;
    for (int J0 = 0, sc = 0; sc < (int)w.size(); J0 += w[sc], sc++)
        for (int Ib = J0; Ib < nt; Ib += chr) t.push_back(unit_tiles(Ib, std::min(Ib + chr, nt), J0, w[sc]));
    return t;
}

// Shape and super-column widths of a single chain's plan: among the unit
// templates (4 x 4 aside, which runs one block per CU) and, for each, the
// uniform widths and every "a super-columns of the full width, then
// narrower ones" mix, the one of least plan_cost on slots = 2 x CUs.  A
// function of the chain length and the CU count only, so a chain gets the
// same plan, and bit-identical products, in every process.
void plan_pick(int nb, int cus, int &chr, int &wsc, std::vector<int> &widths)
{
    const int nt = (3 * nb + TILE - 1) / TILE, slots = 2 * std::max(cus, 1);
    double best = 1e300;
    chr = 2;
    wsc = 4;
    widths = make_widths(nt, 4, 0, 4);
    for (const auto &shp : kTileShapes) {
        if (shp[0] > 2) continue;
        const int c = shp[0], W = shp[1];
        for (int rest = W; rest >= 2; rest--) {
            for (int a = 0; a * W < nt + W; a++) {
                if (rest == W && a > 0) break;   // uniform: one candidate
                const std::vector<int> w = make_widths(nt, W, a, rest);
                const double cost = plan_cost(chain_unit_tiles(nt, c, w), slots);
                if (cost < best - 1e-9) {
                    best = cost;
                    chr = c;
                    wsc = W;
                    widths = w;
                }
            }
        }
    }
}

// Builds the tile list, work units and output-block table of a set of
// chains with n_c rows each, stored back to back (row offsets row_off_c,
// tile offsets in the order given).  widths: per chain, the super-column
// widths (each <= WSC; uniform WSC when null).  slots > 0: order a single
// round of units by pair_order for that many block slots.
// The host half of plan_build: the records only.
struct PlanLayout {
    std::vector<int> tiles, units, rowblk, ci;
    std::vector<long> cl;
    long ntiles = 0, nrowpart = 0;
};

static int plan_layout(PlanLayout &L, const std::vector<int> &nb, const std::vector<long> &row_off, int CHR,
                       int WSC, const std::vector<std::vector<int>> *widths, int slots)
{
    if (!tile_shape_supported(CHR, WSC)) { set_error("unsupported tile unit shape"); return BD_ERR_ARG; }
    const int C = (int)nb.size();
    if (widths && (int)widths->size() != C) { set_error("internal: widths per chain"); return BD_ERR_ARG; }
    std::vector<int> &tiles = L.tiles, &units = L.units, &rowblk = L.rowblk, &ci = L.ci, ntl;
    std::vector<long> &cl = L.cl;
    ci.assign(4 * C, 0);
    cl.assign(3 * C, 0);
    long toff = 0, rpo = 0;
    for (int ch = 0; ch < C; ch++) {
        const int n = 3 * nb[ch], nt = (n + TILE - 1) / TILE;
        const std::vector<int> w = widths ? (*widths)[ch] : make_widths(nt, WSC, 0, WSC);
        int sum = 0;
        for (int x : w) {
            if (x < 1 || x > WSC) { set_error("internal: super-column wider than the unit"); return BD_ERR_ARG; }
            sum += x;
        }
        if (sum != nt) { set_error("internal: super-column widths do not cover the chain"); return BD_ERR_ARG; }
        const int nsc = (int)w.size();
        ci[4 * ch] = nb[ch];
        ci[4 * ch + 1] = n;
        ci[4 * ch + 2] = nt;
        ci[4 * ch + 3] = nsc;
        cl[3 * ch] = row_off[ch];
        cl[3 * ch + 1] = toff;
        cl[3 * ch + 2] = rpo;
        for (int I = 0; I < nt; I++)
            for (int J = 0; J <= I; J++) {
                tiles.push_back(ch); tiles.push_back(I); tiles.push_back(J); tiles.push_back(0);
            }
        std::vector<int> scu(nsc + 1, 0), scJ0(nsc + 1, 0);
        for (int sc = 0; sc < nsc; sc++) {
            scu[sc] = (int)(units.size() / UREC);
            const int J0 = scJ0[sc];
            scJ0[sc + 1] = J0 + w[sc];
            for (int Ib = J0; Ib < nt; Ib += CHR) {
                const int Ie = std::min(Ib + CHR, nt);
                int mask = 0;
                for (int sl = 0; sl < CHR * WSC; sl++) {
                    const int I = Ib + sl / WSC, J = J0 + sl % WSC;
                    if (I < Ie && J <= I && J < J0 + w[sc]) mask |= 1 << sl;
                }
                const long tb = toff * (TILE * TILE), xo = row_off[ch];
                const int cid = (int)(units.size() / UREC);
                const int rec[UREC] = {ch, sc, Ib, Ie, nt, nsc, mask, J0,
                                       (int)(tb & 0xffffffffL), (int)(tb >> 32), (int)(xo & 0xffffffffL),
                                       (int)(xo >> 32), (int)(rpo & 0xffffffffL), (int)(rpo >> 32), cid, w[sc]};
                units.insert(units.end(), rec, rec + UREC);
                ntl.push_back(unit_tiles(Ib, Ie, J0, w[sc]));
            }
        }
        scu[nsc] = (int)(units.size() / UREC);
        for (int K = 0, sc = 0; K < nt; K++) {
            while (K >= scJ0[sc + 1]) sc++;
            rowblk.push_back(ch); rowblk.push_back(K); rowblk.push_back(nsc); rowblk.push_back(scu[sc]);
            rowblk.push_back(scu[sc + 1]); rowblk.push_back((int)rpo); rowblk.push_back((int)(row_off[ch] + (long)K * TILE));
            rowblk.push_back(sc | ((K - scJ0[sc]) << 16));
        }
        toff += (long)nt * (nt + 1) / 2;
        rpo += (long)nt * nsc;
    }
    if (rpo > 0x7fffffffL || row_off.back() + 3L * nb.back() > 0x7fffffffL) {
        set_error("batch too large for 32-bit block offsets");
        return BD_ERR_ARG;
    }
    const int U = (int)(units.size() / UREC);
    if (slots > 0) {
        const std::vector<int> pos = pair_order(ntl, slots);
        std::vector<int> ord(units.size());
        for (int i = 0; i < U; i++)
            std::copy(units.begin() + (long)pos[i] * UREC, units.begin() + (long)(pos[i] + 1) * UREC,
                      ord.begin() + (long)i * UREC);
        units.swap(ord);
    }
    L.ntiles = toff;
    L.nrowpart = rpo;
    return BD_OK;
}

int plan_build(MobPlan &p, const std::vector<int> &nb, const std::vector<long> &row_off, int CHR, int WSC,
               const std::vector<std::vector<int>> *widths, int slots)
{
    PlanLayout L;
    BD_TRY(plan_layout(L, nb, row_off, CHR, WSC, widths, slots));
    const std::vector<int> &tiles = L.tiles, &units = L.units, &rowblk = L.rowblk, &ci = L.ci;
    const std::vector<long> &cl = L.cl;
    const long rpo = L.nrowpart;
    p.nchains = (int)nb.size();
    p.chr = CHR;
    p.wsc = WSC;
    p.ntiles = L.ntiles;
    p.nunits = (int)(units.size() / UREC);
    p.nrowblk = (int)(rowblk.size() / 8);
    auto up = [](auto **dst, const auto &v) -> int {
        using T = typename std::remove_reference<decltype(v)>::type::value_type;
        BD_HIP(hipMalloc(dst, std::max<size_t>(v.size(), 1) * sizeof(T)));
        if (!v.empty()) BD_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return BD_OK;
    };
    BD_TRY(up(&p.tiles, tiles));
    BD_TRY(up(&p.units, units));
    BD_TRY(up(&p.rowblk, rowblk));
    BD_TRY(up(&p.chain_i, ci));
    BD_TRY(up(&p.chain_l, cl));
    BD_HIP(hipMalloc(&p.rowpart, (size_t)std::max(rpo, 1L) * TILE * MAX_S * sizeof(double)));
    BD_HIP(hipMalloc(&p.colpart, (size_t)std::max(p.nunits, 1) * WSC * TILE * MAX_S * sizeof(double)));
    return BD_OK;
}

}  // namespace bd

// Work decomposition of the single-chain plan a context of nbeads beads
// would use on a GPU with cus CUs (host only: no device call).
extern "C" int bd_plan_layout(int nbeads, int cus, int *chr, int *wsc, int *nunits, int *units, int units_cap,
                              int *nrowblk, int *rowblk, int rowblk_cap)
{
    using namespace bd;
    if (nbeads < 1 || cus < 1 || !chr || !wsc || !nunits || !nrowblk) {
        set_error("bd_plan_layout: bad arguments");
        return BD_ERR_ARG;
    }
    std::vector<std::vector<int>> w(1);
    plan_pick(nbeads, cus, *chr, *wsc, w[0]);
    PlanLayout L;
    BD_TRY(plan_layout(L, std::vector<int>{nbeads}, std::vector<long>{0}, *chr, *wsc, &w, 2 * cus));
    *nunits = (int)(L.units.size() / UREC);
    *nrowblk = (int)(L.rowblk.size() / 8);
    if (units && units_cap >= (int)L.units.size()) std::copy(L.units.begin(), L.units.end(), units);
    if (rowblk && rowblk_cap >= (int)L.rowblk.size()) std::copy(L.rowblk.begin(), L.rowblk.end(), rowblk);
    return BD_OK;
}

namespace bd {

void plan_free(MobPlan &p)
{
    int **ip[] = {&p.tiles, &p.units, &p.rowblk, &p.chain_i};
    for (auto q : ip) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
    if (p.chain_l) (void)hipFree(p.chain_l);
    if (p.rowpart) (void)hipFree(p.rowpart);
    if (p.colpart) (void)hipFree(p.colpart);
    p.chain_l = nullptr; p.rowpart = nullptr; p.colpart = nullptr;
}

int mob_alloc(Ctx &c)
{
    c.ntile = (c.n + TILE - 1) / TILE;
    c.npad = c.ntile * TILE;
    c.ntiles_lower = c.ntile * (c.ntile + 1) / 2;
    size_t elems = (size_t)c.ntiles_lower * TILE * TILE;
    if (c.opt.precision == BD_FP32) BD_HIP(hipMalloc(&c.mob32, elems * sizeof(float)));
    else BD_HIP(hipMalloc(&c.mob64, elems * sizeof(double)));
    int cus = 256, chr = 2, wsc = 4;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device);
    std::vector<std::vector<int>> widths(1);
    plan_pick(c.nb, cus, chr, wsc, widths[0]);
    BD_TRY(plan_build(c.plan, std::vector<int>{c.nb}, std::vector<long>{0}, chr, wsc, &widths, 2 * cus));
    c.nunits = c.plan.nunits;
    return BD_OK;
}

void mob_free(Ctx &c)
{
    if (c.mob64) (void)hipFree(c.mob64);
    if (c.mob32) (void)hipFree(c.mob32);
    c.mob64 = nullptr; c.mob32 = nullptr;
    plan_free(c.plan);
    plan_free(c.plan_wide);
}

template <typename E>
int rpy_launch(const MobPlan &p, const double *crd, E *mob, hipStream_t sm)
{
    if (p.ntiles > 0)
        hipLaunchKernelGGL(rpy_tile_kernel<E>, dim3((unsigned)p.ntiles), dim3(256), 0, sm, crd, p.tiles, p.chain_i,
                           p.chain_l, mob);
    BD_HIP(hipGetLastError());
    return BD_OK;
}
template int rpy_launch<float>(const MobPlan &, const double *, float *, hipStream_t);
template int rpy_launch<double>(const MobPlan &, const double *, double *, hipStream_t);

int rpy_assemble_at(Ctx &c, const double *crd)
{
    if (c.mf) {
        BD_TRY(mf_assemble(c, crd));
        c.mob_valid = true;
        return BD_OK;
    }
    if (c.mob32) BD_TRY(rpy_launch<float>(c.plan, crd, c.mob32, c.stream));
    else BD_TRY(rpy_launch<double>(c.plan, crd, c.mob64, c.stream));
    c.mob_valid = true;
    return BD_OK;
}

int rpy_assemble(Ctx &c) { return rpy_assemble_at(c, c.crd); }

// Launches the tile product of a plan: y = D x for every (active) chain.
template <typename E, int S>
int plan_product(const MobPlan &p, const E *mob, const double *x, double *y, int ldx, int ldy, int *ticket,
                 const DevStatus *stop, const int *active, hipStream_t sm, int device, hipEvent_t e0, hipEvent_t e1,
                 const LzEpi *epi, const XPre *pre)
{
    const XPre xp = pre ? *pre : XPre{};
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    // one block per unit up to the occupancy the kernel is built for; beyond
    // that blocks take units in turn
    const int per_cu = (S > 4 || p.chr > 2) ? 1 : 2;   // the kernel's occupancy
    int grid = std::max(1, std::min(p.nunits, cus * per_cu));
    if (e0) BD_HIP(hipEventRecord(e0, sm));
#define BD_TILES(CR, WS)                                                                                         \
    hipLaunchKernelGGL((symm_tiles_kernel<E, S, CR, WS>), dim3(grid), dim3(NT), 0, sm, mob, x, ldx, p.units,      \
                       p.nunits, p.chain_i, p.chain_l, p.rowpart, p.colpart, ticket, stop, active, xp)
    const int shape = p.chr * 16 + p.wsc;
    if (shape == 2 * 16 + 4) BD_TILES(2, 4);
    else if constexpr (S <= 4) {
        if (shape == 1 * 16 + 4) BD_TILES(1, 4);
        else if (shape == 4 * 16 + 4) BD_TILES(4, 4);
        else if (shape == 2 * 16 + 2) BD_TILES(2, 2);
        else if (shape == 2 * 16 + 5) BD_TILES(2, 5);
        else if (shape == 2 * 16 + 6) BD_TILES(2, 6);
        else if (shape == 2 * 16 + 8) BD_TILES(2, 8);
        else { set_error("unsupported tile unit shape"); return BD_ERR_ARG; }
    } else {
        set_error("blocks wider than 4 columns use the 2 x 4 tile unit");
        return BD_ERR_ARG;
    }
#undef BD_TILES
    BD_HIP(hipGetLastError());
    if (e1) BD_HIP(hipEventRecord(e1, sm));
    if (epi) {
        hipLaunchKernelGGL((symm_reduce_lz_kernel<S>), dim3(p.nrowblk), dim3(TILE * S * LZ_HALVES<S>), 0, sm, p.rowpart,
                           p.colpart,
                           p.rowblk, y, ldy, epi->Vj, epi->Vp, epi->Bp, epi->part, stop, p.wsc);
        BD_HIP(hipGetLastError());
        return BD_OK;
    }
    long tot = (long)p.nrowblk * TILE * S;
    hipLaunchKernelGGL((symm_reduce_kernel<S>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, sm, p.rowpart,
                       p.colpart, p.rowblk, p.nrowblk, y, ldy, ticket, active, p.wsc);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

#define BD_INST(E, S)                                                                                          \
    template int plan_product<E, S>(const MobPlan &, const E *, const double *, double *, int, int, int *,     \
                                    const DevStatus *, const int *, hipStream_t, int, hipEvent_t, hipEvent_t,  \
                                    const LzEpi *, const XPre *);
BD_INST(float, 1) BD_INST(float, 2) BD_INST(float, 3) BD_INST(float, 4)
BD_INST(float, 5) BD_INST(float, 6) BD_INST(float, 7) BD_INST(float, 8)
BD_INST(double, 1) BD_INST(double, 2) BD_INST(double, 3) BD_INST(double, 4)
BD_INST(double, 5) BD_INST(double, 6) BD_INST(double, 7) BD_INST(double, 8)
#undef BD_INST

// Autotuner of the work-unit shape: times the product of block width S on
// the plan's own tile layout for every supported shape (five launches after
// two warm-up ones, the fastest of the five counts) and rebuilds the plan
// with the fastest shape.  The time does not depend on the tile values.
// Choices are remembered per (device, element size, S, chain lengths) for
// the life of the process, so a plan built twice gets the same shape (and
// bit-identical results).  times_us[k]: the k-th shape of kTileShapes.

template <typename E, int S>
int plan_tune(MobPlan &p, const std::vector<int> &nb, const std::vector<long> &row_off, const E *mob,
              const double *x, double *y, int ld, int *ticket, hipStream_t sm, int device, double *times_us)
{
    // process-wide, shared by every batch (contexts on other threads too):
    // the lock is held through the timing so one thread times a key once
    static std::map<std::string, std::pair<int, std::vector<double>>> cache;
    static std::mutex cache_mu;
    std::lock_guard<std::mutex> lock(cache_mu);
    std::string key = std::to_string(device) + ":" + std::to_string(sizeof(E)) + ":" + std::to_string(S) + ":";
    for (int v : nb) key += std::to_string(v) + ",";
    auto hit = cache.find(key);
    int best = 0;
    std::vector<double> t(BD_TILE_SHAPES, 0.0);
    if (hit != cache.end()) {
        best = hit->second.first;
        t = hit->second.second;
    } else {
        hipEvent_t e0, e1;
        BD_HIP(hipEventCreateWithFlags(&e0, kTimingEvent));
        BD_HIP(hipEventCreateWithFlags(&e1, kTimingEvent));
        for (int k = 0; k < BD_TILE_SHAPES; k++) {
            plan_free(p);
            BD_TRY(plan_build(p, nb, row_off, kTileShapes[k][0], kTileShapes[k][1]));
            double tmin = 1e30;
            for (int rep = 0; rep < 7; rep++) {
                BD_HIP(hipEventRecord(e0, sm));
                BD_TRY((plan_product<E, S>(p, mob, x, y, ld, ld, ticket, nullptr, nullptr, sm, device, nullptr,
                                           nullptr)));
                BD_HIP(hipEventRecord(e1, sm));
                BD_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                BD_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (rep >= 2) tmin = std::min(tmin, 1e3 * (double)ms);
            }
            t[k] = tmin;
            if (k == 0 || tmin < t[best]) best = k;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        cache[key] = {best, t};
    }
    plan_free(p);
    BD_TRY(plan_build(p, nb, row_off, kTileShapes[best][0], kTileShapes[best][1]));
    if (times_us)
        for (int k = 0; k < BD_TILE_SHAPES; k++) times_us[k] = t[k];
    return BD_OK;
}
#define BD_INST(E, S)                                                                                          \
    template int plan_tune<E, S>(MobPlan &, const std::vector<int> &, const std::vector<long> &, const E *,     \
                                 const double *, double *, int, int *, hipStream_t, int, double *);
BD_INST(float, 1) BD_INST(float, 2) BD_INST(float, 3) BD_INST(float, 4)
BD_INST(double, 1) BD_INST(double, 2) BD_INST(double, 3) BD_INST(double, 4)
#undef BD_INST

int prof_pair(Ctx &c, int w, hipEvent_t *e0, hipEvent_t *e1)
{
    *e0 = *e1 = nullptr;
    if (!c.prof_on || (c.prof_width != 0 && c.prof_width != w)) return BD_OK;
    if (c.prof_used + 2 > c.prof_ev.size()) prof_flush(c);
    if (c.prof_used + 2 > c.prof_ev.size()) {
        hipEvent_t a, b;
        BD_HIP(hipEventCreateWithFlags(&a, kTimingEvent));
        BD_HIP(hipEventCreateWithFlags(&b, kTimingEvent));
        c.prof_ev.push_back(a);
        c.prof_ev.push_back(b);
    }
    *e0 = c.prof_ev[c.prof_used];
    *e1 = c.prof_ev[c.prof_used + 1];
    c.prof_used += 2;
    return BD_OK;
}

template <typename E, int S>
static int product_t(Ctx &c, const E *mob, const double *x, double *y, int ldx, int ldy, const DevStatus *stop,
                     const LzEpi *epi, const XPre *pre)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BD_TRY(prof_pair(c, S, &e0, &e1));
    const MobPlan *pl = &c.plan;
    if constexpr (S > 4) {
        // only the 2 x 4 unit is built for these widths
        if (c.plan.chr != 2 || c.plan.wsc != 4) {
            if (!c.plan_wide.units) BD_TRY(plan_build(c.plan_wide, std::vector<int>{c.nb}, std::vector<long>{0}, 2, 4));
            pl = &c.plan_wide;
        }
    }
    return plan_product<E, S>(*pl, mob, x, y, ldx, ldy, &c.st->ticket, stop, nullptr, c.stream, c.device, e0, e1,
                              epi, pre);
}

template <typename E>
static int product_dispatch(Ctx &c, const E *mob, const double *x, double *y, int s, int ldx, int ldy,
                            const DevStatus *stop, const LzEpi *epi = nullptr, const XPre *pre = nullptr)
{
    switch (s) {
    case 1: return product_t<E, 1>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    case 2: return product_t<E, 2>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    case 3: return product_t<E, 3>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    case 4: return product_t<E, 4>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    case 5: return product_t<E, 5>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    case 6: return product_t<E, 6>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    case 7: return product_t<E, 7>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    case 8: return product_t<E, 8>(c, mob, x, y, ldx, ldy, stop, epi, pre);
    default: set_error("block size s must be in 1..8"); return BD_ERR_ARG;
    }
}

int mob_product(Ctx &c, const double *x, double *y, int s, int ldx, int ldy)
{
    if (!c.mob_valid) { set_error("mobility not assembled"); return BD_ERR_STATE; }
    if (c.mf) return mf_product(c, x, y, s, ldx, ldy, nullptr);
    if (c.mob32) return product_dispatch<float>(c, c.mob32, x, y, s, ldx, ldy, nullptr);
    return product_dispatch<double>(c, c.mob64, x, y, s, ldx, ldy, nullptr);
}

int mob_product_lz(Ctx &c, const double *x, double *y, int s, int ldx, int ldy, const LzEpi *epi, const XPre *pre)
{
    if (!c.mob_valid) { set_error("mobility not assembled"); return BD_ERR_STATE; }
    if (c.mf) {
        if (epi) return mf_product_lz(c, x, y, s, ldx, pre != nullptr, *epi, pre);
        return mf_product(c, x, y, s, ldx, ldy, c.st);
    }
    if (c.mob32) return product_dispatch<float>(c, c.mob32, x, y, s, ldx, ldy, c.st, epi, pre);
    return product_dispatch<double>(c, c.mob64, x, y, s, ldx, ldy, c.st, epi, pre);
}

// Sums the recorded (start, stop) pairs into prof_ms and recycles them.
void prof_flush(Ctx &c)
{
    if (c.prof_used == 0) return;
    (void)hipStreamSynchronize(c.stream);
    for (size_t k = 0; k + 1 < c.prof_used; k += 2) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c.prof_ev[k], c.prof_ev[k + 1]) == hipSuccess) {
            c.prof_ms += ms;
            c.prof_count++;
        }
    }
    c.prof_used = 0;
    if (c.prof_ev.size() < 512) {
        // grow the pool so long runs flush rarely
        for (int k = 0; k < 64; k++) {
            hipEvent_t a;
            if (hipEventCreate(&a) != hipSuccess) break;
            c.prof_ev.push_back(a);
        }
    }
}

int mob_unpack(Ctx &c, double *dense, int ld)
{
    if (c.mf) return mf_dense(c, dense, ld);
    long tot = (long)c.n * c.n;
    if (tot == 0) return BD_OK;
    unsigned grid = (unsigned)((tot + 255) / 256);
    if (c.mob32)
        hipLaunchKernelGGL(unpack_kernel<float>, dim3(grid), dim3(256), 0, c.stream, c.mob32, c.n, dense, ld);
    else
        hipLaunchKernelGGL(unpack_kernel<double>, dim3(grid), dim3(256), 0, c.stream, c.mob64, c.n, dense, ld);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

}  // namespace bd
