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
#include "bd_internal.h"

namespace bd {

constexpr int WSC = 4;   // tile columns per super-column
constexpr int CHR = 2;   // tile rows per work unit
constexpr int NT = 256;  // threads per block of the product kernel

__host__ __device__ inline long tile_index(int I, int J) { return (long)I * (I + 1) / 2 + J; }

// ---------------------------------------------------------------- assembly
// One block per stored tile.  Bead-pair 3x3 blocks for the <= 22 x 22 bead
// pairs the tile touches are computed once into LDS (6 unique entries each),
// then the 64 x 64 entries are written column-major (coalesced along rows).
template <typename E>
__global__ __launch_bounds__(256) void rpy_tile_kernel(const double *__restrict__ crd, int nb,
                                                       int n, int ntile, E *__restrict__ mob)
{
    __shared__ double sblk[22 * 22 * 6];
    // map linear block id to (I, J) with I >= J
    long t = blockIdx.x;
    int I = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((long)I * (I + 1) / 2 > t) I--;
    while ((long)(I + 1) * (I + 2) / 2 <= t) I++;
    int J = (int)(t - (long)I * (I + 1) / 2);

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
template <typename E>
struct TileLds {
    E d[TILE][TILE + 1];            // column-major tile, padded
};

// Persistent kernel; blocks pull work units from a ticket counter.
// rowpart: (ntile * nsc) x TILE x S; colpart: nunits x WSC x TILE x S.
template <typename E, int S>
__global__ __launch_bounds__(NT) void symm_tiles_kernel(
    const E *__restrict__ mob, const double *__restrict__ x, int ldx, const int *__restrict__ units,
    int nunits, int ntile, int nsc, double *__restrict__ rowpart, double *__restrict__ colpart,
    int *__restrict__ ticket, const DevStatus *stop)
{
    // inside a Lanczos solve: nothing to do once it has converged
    if (stop && (stop->converged_j != 0 || stop->ierr != 0)) return;
    // the reduction buffer aliases the staged tile (never live together)
    __shared__ union {
        TileLds<E> t;
        double red[NT / TILE][TILE][S];
    } sU;
    __shared__ double sVJ[WSC][TILE][S];
    __shared__ double sVI[TILE][S];
    __shared__ int s_unit;
    auto &sT = sU.t;
    auto &sred = sU.red;

    const int tid = threadIdx.x;
    const int lane_r = tid % TILE;      // row (row product) / column (column product)
    const int q = tid / TILE;           // quarter: 16 columns (row prod) / 16 rows (col prod)

    for (;;) {
        if (tid == 0) s_unit = atomicAdd(ticket, 1);
        __syncthreads();
        const int u = s_unit;
        if (u >= nunits) break;
        const int sc = units[3 * u], Ib = units[3 * u + 1], Ie = units[3 * u + 2];
        const int J0 = sc * WSC;
        // V_J blocks for the unit's tile columns
        for (int e = tid; e < WSC * TILE * S; e += NT) {
            int jj = e / (TILE * S), rem = e % (TILE * S), r = rem / S, k = rem % S;
            int J = J0 + jj;
            sVJ[jj][r][k] = (J < ntile) ? x[(long)k * ldx + (long)J * TILE + r] : 0.0;
        }
        double colacc[WSC][S];
#pragma unroll
        for (int jj = 0; jj < WSC; jj++)
#pragma unroll
            for (int k = 0; k < S; k++) colacc[jj][k] = 0.0;

        for (int I = Ib; I < Ie; I++) {
            double rowacc[S];
#pragma unroll
            for (int k = 0; k < S; k++) rowacc[k] = 0.0;
            __syncthreads();
            for (int e = tid; e < TILE * S; e += NT) {
                int r = e / S, k = e % S;
                sVI[r][k] = x[(long)k * ldx + (long)I * TILE + r];
            }
            const int jmax = min(WSC, I - J0 + 1);
#pragma unroll
            for (int jj = 0; jj < WSC; jj++) {
                if (jj >= jmax) break;
                const int J = J0 + jj;
                const E *tp = mob + tile_index(I, J) * (TILE * TILE);
                __syncthreads();
                // stage the tile: element e = c*TILE + r
#pragma unroll 4
                for (int e = tid; e < TILE * TILE; e += NT) sT.d[e / TILE][e % TILE] = tp[e];
                __syncthreads();
                // row product: rows lane_r, columns 16q..16q+15
#pragma unroll 4
                for (int cc = 0; cc < TILE / 4; cc++) {
                    int c = q * (TILE / 4) + cc;
                    double dv = (double)sT.d[c][lane_r];
#pragma unroll
                    for (int k = 0; k < S; k++) rowacc[k] += dv * sVJ[jj][c][k];
                }
                if (I != J) {
                    // column product: column lane_r, rows 16q..16q+15
#pragma unroll 4
                    for (int rr = 0; rr < TILE / 4; rr++) {
                        int r = q * (TILE / 4) + rr;
                        double dv = (double)sT.d[lane_r][r];
#pragma unroll
                        for (int k = 0; k < S; k++) colacc[jj][k] += dv * sVI[r][k];
                    }
                }
            }
            // reduce row partial over quarters, write one partial per (I, sc)
            __syncthreads();
#pragma unroll
            for (int k = 0; k < S; k++) sred[q][lane_r][k] = rowacc[k];
            __syncthreads();
            for (int e = tid; e < TILE * S; e += NT) {
                int r = e / S, k = e % S;
                double v = 0.0;
#pragma unroll
                for (int qq = 0; qq < NT / TILE; qq++) v += sred[qq][r][k];
                rowpart[((long)I * nsc + sc) * (TILE * S) + e] = v;
            }
        }
        // column partials of this unit
        for (int jj = 0; jj < WSC; jj++) {
            __syncthreads();
#pragma unroll
            for (int k = 0; k < S; k++) sred[q][lane_r][k] = colacc[jj][k];
            __syncthreads();
            for (int e = tid; e < TILE * S; e += NT) {
                int r = e / S, k = e % S;
                double v = 0.0;
#pragma unroll
                for (int qq = 0; qq < NT / TILE; qq++) v += sred[qq][r][k];
                colpart[((long)u * WSC + jj) * (TILE * S) + e] = v;
            }
        }
        __syncthreads();
    }
}

// y(K block) = sum_{sc <= K/WSC} rowpart[K, sc] + sum_{units of sc(K)} colpart[u, K % WSC]
template <int S>
__global__ void symm_reduce_kernel(const double *__restrict__ rowpart, const double *__restrict__ colpart,
                                   const int *__restrict__ sc_unit_ptr, int ntile, int nsc,
                                   double *__restrict__ y, int ldy, int *__restrict__ ticket)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) *ticket = 0;  // re-arm the work counter for the next product
    if (g >= (long)ntile * TILE * S) return;
    int K = (int)(g / (TILE * S));
    int e = (int)(g % (TILE * S));
    int r = e / S, k = e % S;
    double v = 0.0;
    for (int sc = 0; sc <= K / WSC; sc++) v += rowpart[((long)K * nsc + sc) * (TILE * S) + e];
    int scK = K / WSC, jj = K % WSC;
    for (int u = sc_unit_ptr[scK]; u < sc_unit_ptr[scK + 1]; u++)
        v += colpart[((long)u * WSC + jj) * (TILE * S) + e];
    y[(long)k * ldy + (long)K * TILE + r] = v;
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
int mob_alloc(Ctx &c)
{
    c.ntile = (c.n + TILE - 1) / TILE;
    c.npad = c.ntile * TILE;
    c.ntiles_lower = c.ntile * (c.ntile + 1) / 2;
    size_t elems = (size_t)c.ntiles_lower * TILE * TILE;
    if (c.opt.precision == BD_FP32) BD_HIP(hipMalloc(&c.mob32, elems * sizeof(float)));
    else BD_HIP(hipMalloc(&c.mob64, elems * sizeof(double)));

    const int nsc = (c.ntile + WSC - 1) / WSC;
    std::vector<int> units, scptr(nsc + 1, 0);
    for (int sc = 0; sc < nsc; sc++) {
        scptr[sc] = (int)(units.size() / 3);
        int J0 = sc * WSC;
        for (int Ib = J0; Ib < c.ntile; Ib += CHR) {
            units.push_back(sc);
            units.push_back(Ib);
            units.push_back(std::min(Ib + CHR, c.ntile));
        }
    }
    scptr[nsc] = (int)(units.size() / 3);
    c.nunits = (int)(units.size() / 3);
    BD_HIP(hipMalloc(&c.units, units.size() * sizeof(int)));
    BD_HIP(hipMemcpy(c.units, units.data(), units.size() * sizeof(int), hipMemcpyHostToDevice));
    BD_HIP(hipMalloc(&c.colunit_ptr, scptr.size() * sizeof(int)));
    BD_HIP(hipMemcpy(c.colunit_ptr, scptr.data(), scptr.size() * sizeof(int), hipMemcpyHostToDevice));
    BD_HIP(hipMalloc(&c.rowpart, (size_t)c.ntile * nsc * TILE * MAX_S * sizeof(double)));
    BD_HIP(hipMalloc(&c.colpart, (size_t)c.nunits * WSC * TILE * MAX_S * sizeof(double)));
    return BD_OK;
}

void mob_free(Ctx &c)
{
    if (c.mob64) (void)hipFree(c.mob64);
    if (c.mob32) (void)hipFree(c.mob32);
    if (c.units) (void)hipFree(c.units);
    if (c.colunit_ptr) (void)hipFree(c.colunit_ptr);
    if (c.rowpart) (void)hipFree(c.rowpart);
    if (c.colpart) (void)hipFree(c.colpart);
    c.mob64 = nullptr; c.mob32 = nullptr; c.units = nullptr; c.colunit_ptr = nullptr;
    c.rowpart = nullptr; c.colpart = nullptr;
}

int rpy_assemble_at(Ctx &c, const double *crd)
{
    if (c.nb > 0) {
        if (c.mob32)
            hipLaunchKernelGGL(rpy_tile_kernel<float>, dim3(c.ntiles_lower), dim3(256), 0, c.stream,
                               crd, c.nb, c.n, c.ntile, c.mob32);
        else
            hipLaunchKernelGGL(rpy_tile_kernel<double>, dim3(c.ntiles_lower), dim3(256), 0, c.stream,
                               crd, c.nb, c.n, c.ntile, c.mob64);
        BD_HIP(hipGetLastError());
    }
    c.mob_valid = true;
    return BD_OK;
}

int rpy_assemble(Ctx &c) { return rpy_assemble_at(c, c.crd); }

template <typename E, int S>
static int product_t(Ctx &c, const E *mob, const double *x, double *y, int ldx, int ldy, const DevStatus *stop)
{
    const int nsc = (c.ntile + WSC - 1) / WSC;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c.device);
    int grid = std::min(c.nunits, cus * 3);
    int *ticket = &c.st->ticket;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c.prof_on) {
        if (c.prof_used + 2 > c.prof_ev.size()) prof_flush(c);
        if (c.prof_used + 2 > c.prof_ev.size()) {
            hipEvent_t a, b;
            BD_HIP(hipEventCreate(&a));
            BD_HIP(hipEventCreate(&b));
            c.prof_ev.push_back(a);
            c.prof_ev.push_back(b);
        }
        e0 = c.prof_ev[c.prof_used];
        e1 = c.prof_ev[c.prof_used + 1];
        c.prof_used += 2;
        BD_HIP(hipEventRecord(e0, c.stream));
    }
    hipLaunchKernelGGL((symm_tiles_kernel<E, S>), dim3(grid), dim3(NT), 0, c.stream, mob, x, ldx,
                       c.units, c.nunits, c.ntile, nsc, c.rowpart, c.colpart, ticket, stop);
    BD_HIP(hipGetLastError());
    if (c.prof_on) BD_HIP(hipEventRecord(e1, c.stream));
    long tot = (long)c.ntile * TILE * S;
    hipLaunchKernelGGL((symm_reduce_kernel<S>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       c.stream, c.rowpart, c.colpart, c.colunit_ptr, c.ntile, nsc, y, ldy, ticket);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

template <typename E>
static int product_dispatch(Ctx &c, const E *mob, const double *x, double *y, int s, int ldx, int ldy,
                            const DevStatus *stop)
{
    switch (s) {
    case 1: return product_t<E, 1>(c, mob, x, y, ldx, ldy, stop);
    case 2: return product_t<E, 2>(c, mob, x, y, ldx, ldy, stop);
    case 3: return product_t<E, 3>(c, mob, x, y, ldx, ldy, stop);
    case 4: return product_t<E, 4>(c, mob, x, y, ldx, ldy, stop);
    case 5: return product_t<E, 5>(c, mob, x, y, ldx, ldy, stop);
    case 6: return product_t<E, 6>(c, mob, x, y, ldx, ldy, stop);
    case 7: return product_t<E, 7>(c, mob, x, y, ldx, ldy, stop);
    case 8: return product_t<E, 8>(c, mob, x, y, ldx, ldy, stop);
    default: set_error("block size s must be in 1..8"); return BD_ERR_ARG;
    }
}

int mob_product(Ctx &c, const double *x, double *y, int s, int ldx, int ldy)
{
    if (!c.mob_valid) { set_error("mobility not assembled"); return BD_ERR_STATE; }
    if (c.mob32) return product_dispatch<float>(c, c.mob32, x, y, s, ldx, ldy, nullptr);
    return product_dispatch<double>(c, c.mob64, x, y, s, ldx, ldy, nullptr);
}

int mob_product_lz(Ctx &c, const double *x, double *y, int s, int ldx, int ldy)
{
    if (!c.mob_valid) { set_error("mobility not assembled"); return BD_ERR_STATE; }
    if (c.mob32) return product_dispatch<float>(c, c.mob32, x, y, s, ldx, ldy, c.st);
    return product_dispatch<double>(c, c.mob64, x, y, s, ldx, ldy, c.st);
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
