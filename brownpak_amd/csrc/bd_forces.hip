// bd_forces.hip -- bond and pair forces over a device neighbour list.
//
// Replaces ia_calc_forces (interaction_m.f90:84-147) with
// ia_add_bond_forces (219-272) / ia_get_bond_force (ia_bond_m.f90:85-362),
// ia_add_vdw_forces (interaction_m.f90:151-215) / vdw_lj (ia_vdw_m.f90:155-
// 180), and the pair-table builders of pairtab_m.fpp (DIR 152-190, VER
// 194-252) with exclusions from connectivity_m.f90:43-317.
//
// The reference accumulates +f on i and -f on j pair by pair.  Here every
// bead gathers its own terms (each pair is evaluated from both ends), so no
// atomics are needed and the sum order is fixed.  The pair predicate
// r^2 < rc^2 (or (rc+skin)^2 for VER) uses the reference's expression
// ((x*x + y*y) + z*z) with separately rounded operations and the nint
// minimum image, so the device pair set is bit-identical to a scalar FP64
// evaluation of the reference's loop.
#include <algorithm>

#include "bd_internal.h"
#include "bd_forces_dev.h"

namespace bd {

constexpr int NBR_CAP0 = 64;        // initial neighbours per bead (full list)
constexpr int FB = 256;             // threads per force block
constexpr int FW = FB / 64;         // fewest beads per force_gather_kernel block (64 lanes each)
constexpr int N2_WAVES = 8;         // beads per nbr_n2_kernel block (a wavefront each)
constexpr int N2_CHUNK = 2048;      // coordinates staged in LDS per pass of nbr_n2_kernel

__device__ inline bool excluded(const int *ex_ptr, const int *ex_col, int i, int j);

// Slot of neighbour k of bead i in the list: bead-major, each bead's row of
// cap entries contiguous (a group of lanes on one bead reads and writes
// neighbouring words).
__device__ __forceinline__ long nbr_slot(int i, int k, int cap) { return (long)i * cap + k; }

// ------------------------------------------------------------ energy sums
// acc layout: 0 e_bond, 1 e_vdw, 2..10 stress(3x3 col-major), 11 bndlen sum,
// 12 bndlen min, 13 bndlen max, 14 e_angle, 15 e_tether, 16 e_external
constexpr int NACC = 17;

__device__ void block_acc_store(double (&a)[NACC], double *out)
{
    __shared__ double sh[FB / 64][NACC];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    for (int e = 0; e < NACC; e++) {
        double v = a[e];
        for (int off = 32; off > 0; off >>= 1) {
            double o = __shfl_xor(v, off, 64);
            v = (e == 12) ? fmin(v, o) : (e == 13) ? fmax(v, o) : v + o;
        }
        if (lane == 0) sh[wv][e] = v;
    }
    __syncthreads();
    if (threadIdx.x < NACC) {
        int e = threadIdx.x;
        double v = sh[0][e];
        for (int w = 1; w < FB / 64; w++) v = (e == 12) ? fmin(v, sh[w][e]) : (e == 13) ? fmax(v, sh[w][e]) : v + sh[w][e];
        out[blockIdx.x * NACC + e] = v;
    }
}

// ------------------------------------------------------------ forces
// Per-bead gather of bond and pair terms, LPB lanes per bead (a group of
// 8, 16 or 64 lanes of one wavefront, chosen from the neighbour list's
// capacity): lane l of the group takes the bead's bonds l, l + LPB, ... and
// then its neighbours l, l + LPB, ... (list order), and the group adds the
// lanes' sums in a fixed butterfly order (a handful of terms per lane
// instead of a chain of dependent loads per bead; a short list leaves few
// lanes idle, and the beads of a wavefront read neighbouring words of the
// list).  Energies/stress/bond lengths are counted once per term: bonds
// from their first bead, pairs from the lower index.
// Angles, tethers, external fields and charges (bd_set_angles, bd_set_tethers,
// bd_set_externals, bd_set_charges): per-bead CSR lists so that every bead's
// wavefront gathers its own terms.
struct ExtraArg {
    int nangles, ntethers;            // 0: the term's lists are not set (null)
    const int *ang_ptr, *ang_idx;     // bead -> 4 * angle + role (0: i-1, 1: i, 2: i+1)
    const int *ang_atoms, *ang_type;  // (3, nangles) 0-based beads; type 0-based
    const int *ang_styles;
    const double *ang_params;
    const int *teth_ptr, *teth_idx;   // bead -> tether
    const int *teth_type, *teth_styles;
    const double *teth_pts, *teth_params;
    int nexternals;
    const int *ext_styles;
    const double *ext_params;
    const double *charge;             // null: all charges 0
};

template <bool EXTRA, bool GEN, int LPB>
__global__ __launch_bounds__(FB) void force_gather_kernel(
    const double *__restrict__ crd, int nb, BoxArg bx, const int *__restrict__ bond_ptr,
    const int *__restrict__ bond_idx, const int *__restrict__ bond_ij, const int *__restrict__ bond_type,
    const int *__restrict__ bstyles, const double *__restrict__ bparams, int use_pairs,
    const int *__restrict__ nbr_cnt, const int *__restrict__ nbr_idx, int nbr_cap,
    PairScan ps, const int *__restrict__ ex_ptr,
    const int *__restrict__ ex_col, const int *__restrict__ atype, int natype, const int *__restrict__ ptype, const int *__restrict__ vstyles,
    const double *__restrict__ vparams, ExtraArg xa, double *__restrict__ frc, int want_energy,
    double *__restrict__ epart, DevStatus *st)
{
    const int lane = threadIdx.x % LPB;
    const int i = blockIdx.x * (FB / LPB) + threadIdx.x / LPB;
    double acc[NACC];
    for (int e = 0; e < NACC; e++) acc[e] = 0.0;
    acc[12] = 1.0e300;
    acc[13] = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (i < nb) {
        // the bead's bond range and neighbour count in one round trip
        const int kb0 = bond_ptr[i], kb1 = bond_ptr[i + 1];
        const int cnt = use_pairs ? nbr_cnt[i] : 0;
        const int ati = use_pairs ? atype[i] : 0;
        for (int k = kb0 + lane; k < kb1; k += LPB) {
            int b = bond_idx[k];
            int bi = bond_ij[2 * b], bj = bond_ij[2 * b + 1];
            double x, y, z;
            pair_vec(crd, bi, bj, bx, x, y, z);
            double r = sqrt(x * x + y * y + z * z);
            int t = bond_type[b];
            double en, f;
            if (bond_force(r, bstyles[t], bparams + t * MXPARAM, en, f)) atomicExch(&st->ierr, 1);
            double gx = f * x / r, gy = f * y / r, gz = f * z / r;
            double sg = (bi == i) ? 1.0 : -1.0;
            if (bi == i && bj == i) sg = 0.0;  // degenerate self bond
            fx += sg * gx; fy += sg * gy; fz += sg * gz;
            if (want_energy && bi == i) {
                acc[0] += en;
                double rv[3] = {x, y, z}, fv[3] = {gx, gy, gz};
                for (int bb = 0; bb < 3; bb++)
                    for (int a = 0; a < 3; a++) acc[2 + a + 3 * bb] -= rv[a] * fv[bb];
                acc[11] += r;
                acc[12] = fmin(acc[12], r);
                acc[13] = fmax(acc[13], r);
            }
        }
        if (use_pairs) {
            const bool full = cnt > nbr_cap;  // list overflowed: scan all beads
            const int kend = full ? nb : cnt;
            for (int k = lane; k < kend; k += LPB) {
                int j, mult = 1;
                if (full) {
                    j = k;
                    if (j == i) continue;
                    mult = scan_mult(ps, bx, i, j);
                    if (mult == 0 || excluded(ex_ptr, ex_col, i, j)) continue;
                } else {
                    j = nbr_idx[nbr_slot(i, k, nbr_cap)];
                }
                int t = ptype[ati * natype + atype[j]];
                const int sty = vstyles[t];
                if (GEN ? sty == 0 : sty != 1) continue;
                double x, y, z;
                pair_vec(crd, i, j, bx, x, y, z);
                double r = sqrt(x * x + y * y + z * z);
                double en, f;
                if (GEN) {
                    const double qq = xa.charge ? xa.charge[i] * xa.charge[j] : 0.0;
                    vdw_force(r, qq, sty, vparams + t * MXPARAM, en, f);
                } else {
                    lj_force(r, vparams + t * MXPARAM, en, f);
                }
                double gx = f * x / r, gy = f * y / r, gz = f * z / r;
                if (mult != 1) {
                    gx *= mult; gy *= mult; gz *= mult; en *= mult;
                }
                fx += gx; fy += gy; fz += gz;
                if (want_energy && i < j) {
                    acc[1] += en;
                    double rv[3] = {x, y, z}, fv[3] = {gx, gy, gz};
                    for (int bb = 0; bb < 3; bb++)
                        for (int a = 0; a < 3; a++) acc[2 + a + 3 * bb] -= rv[a] * fv[bb];
                }
            }
        }
        if (EXTRA) {
            // ia_add_angle_forces (interaction_m.f90:276-325): the bead's share
            // of each of its angles; energy and stress once, at the centre bead
            const int ka0 = xa.nangles > 0 ? xa.ang_ptr[i] : 0, ka1 = xa.nangles > 0 ? xa.ang_ptr[i + 1] : 0;
            for (int k = ka0 + lane; k < ka1; k += LPB) {
                const int code = xa.ang_idx[k], a = code >> 2, role = code & 3;
                const int b0 = xa.ang_atoms[3 * a], b1 = xa.ang_atoms[3 * a + 1], b2 = xa.ang_atoms[3 * a + 2];
                double q1[3], q2[3];
                pair_vec(crd, b0, b1, bx, q1[0], q1[1], q1[2]);
                pair_vec(crd, b1, b2, bx, q2[0], q2[1], q2[2]);
                const int t = xa.ang_type[a];
                double en, f0[3], f2[3];
                angle_force(q1, q2, xa.ang_styles[t], xa.ang_params + t * MXPARAM, en, f0, f2);
                if (role == 0) { fx += f0[0]; fy += f0[1]; fz += f0[2]; }
                if (role == 2) { fx += f2[0]; fy += f2[1]; fz += f2[2]; }
                if (role == 1) {
                    fx -= f0[0] + f2[0]; fy -= f0[1] + f2[1]; fz -= f0[2] + f2[2];
                    if (want_energy) {
                        acc[14] += en;
                        for (int bb = 0; bb < 3; bb++)
                            for (int aa = 0; aa < 3; aa++)
                                acc[2 + aa + 3 * bb] += -q1[aa] * f0[bb] + q2[aa] * f2[bb];
                    }
                }
            }
            // ia_add_tether_forces (interaction_m.f90:375-418), no imaging
            const int kt0 = xa.ntethers > 0 ? xa.teth_ptr[i] : 0, kt1 = xa.ntethers > 0 ? xa.teth_ptr[i + 1] : 0;
            for (int k = kt0 + lane; k < kt1; k += LPB) {
                const int tt = xa.teth_idx[k], t = xa.teth_type[tt];
                double q[3];
                for (int d = 0; d < 3; d++) q[d] = crd[3 * i + d] - xa.teth_pts[3 * tt + d];
                const double qm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                double en, f;
                if (tether_force(qm, xa.teth_styles[t], xa.teth_params + t * MXPARAM, en, f))
                    atomicCAS(&st->ierr, 0, 5);
                const double fj[3] = {-f * q[0] / qm, -f * q[1] / qm, -f * q[2] / qm};
                fx += fj[0]; fy += fj[1]; fz += fj[2];
                if (want_energy) {
                    acc[15] += en;
                    for (int bb = 0; bb < 3; bb++)
                        for (int aa = 0; aa < 3; aa++) acc[2 + aa + 3 * bb] += q[aa] * fj[bb];
                }
            }
            // ia_add_external_forces (ia_external_m.f90:61-107)
            if (lane == 0) {
                for (int e = 0; e < xa.nexternals; e++) {
                    const double *p = xa.ext_params + e * MXPARAM;
                    if (xa.ext_styles[e] == 1 && (int)p[0] - 1 == i) {
                        fx += p[1];
                        if (want_energy) acc[16] += -p[1] * (crd[3 * i] - crd[0]);
                    } else if (xa.ext_styles[e] == 2) {
                        const double x = crd[3 * i + (int)p[0] - 1];
                        if (p[2] > 0.0 ? x < p[1] : x > p[1]) atomicCAS(&st->ierr, 0, 6);
                    }
                }
            }
        }
    }
    // the group's sum (every lane ends with the same value: each step adds
    // the same two operands on both lanes of a pair; the partners stay
    // inside the group)
    for (int off = LPB / 2; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off, 64);
        fy += __shfl_xor(fy, off, 64);
        fz += __shfl_xor(fz, off, 64);
    }
    if (i < nb && lane == 0) {
        frc[3 * i] = fx;
        frc[3 * i + 1] = fy;
        frc[3 * i + 2] = fz;
    }
    if (want_energy) block_acc_store(acc, epart);
}

__global__ void energy_final_kernel(const double *__restrict__ epart, int nblk, int nbonds, int imcon, double vol,
                                    double *__restrict__ out)
{
    // thread e combines quantity e over the blocks in block order
    __shared__ double sa[NACC];
    const int e = threadIdx.x;
    if (e < NACC) {
        double a = epart[e];
        for (int b = 1; b < nblk; b++) {
            const double v = epart[b * NACC + e];
            a = (e == 12) ? fmin(a, v) : (e == 13) ? fmax(a, v) : a + v;
        }
        sa[e] = a;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double a[NACC];
    for (int q = 0; q < NACC; q++) a[q] = sa[q];
    out[0] = a[0];
    out[1] = a[1];
    // energy_tot in the reference's order: bond, angle, dihedral (0), vdw,
    // tether, external (interaction_m.f90:141-142)
    out[2] = a[0] + a[14] + 0.0 + a[1] + a[15] + a[16];
    for (int k = 0; k < 9; k++) out[3 + k] = imcon ? a[2 + k] / vol : a[2 + k];
    out[12] = nbonds > 0 ? a[11] / nbonds : 0.0;
    out[13] = nbonds > 0 ? a[12] : 0.0;
    out[14] = a[13];
    out[24] = a[14];
    out[25] = 0.0;
    out[26] = a[15];
    out[27] = a[16];
}

// ------------------------------------------------------------ pair lists
__device__ inline bool excluded(const int *ex_ptr, const int *ex_col, int i, int j)
{
    // rows are sorted ascending
    for (int k = ex_ptr[i]; k < ex_ptr[i + 1]; k++)
        if (ex_col[k] == j) return true;
    return false;
}

// Verlet trigger (pairtab_m.fpp:211-222): rebuild iff first call or
// 4 max|dr|^2 >= skin^2.  One thread per bead; each block takes the
// maximum of its beads (fmax, so a NaN displacement is passed over as in a
// serial maximum) and flags vmove when 4 max >= skin^2, which holds for some
// block exactly when it holds for the maximum over all beads.  The last block
// to finish (vdone ticket) sets rebuild, counts it (nrebuild) and clears
// both counters.
__global__ __launch_bounds__(256) void verlet_check_kernel(const double *__restrict__ crd,
                                                           const double *__restrict__ save, int nb, BoxArg bx,
                                                           double tskin_sq, int first, DevStatus *st)
{
    __shared__ double sh[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double mx = 0.0;
    if (i < nb) {
        double dx = __dsub_rn(crd[3 * i], save[3 * i]);
        double dy = __dsub_rn(crd[3 * i + 1], save[3 * i + 1]);
        double dz = __dsub_rn(crd[3 * i + 2], save[3 * i + 2]);
        if (bx.imcon) { dx = img(dx, bx.L[0]); dy = img(dy, bx.L[1]); dz = img(dz, bx.L[2]); }
        mx = fmax(mx, rsq3(dx, dy, dz));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); w++) m = fmax(m, sh[w]);
        if (!(4.0 * m < tskin_sq)) __hip_atomic_fetch_or(&st->vmove, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int done = __hip_atomic_fetch_add(&st->vdone, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (done == (int)gridDim.x - 1) {
            const int mv = __hip_atomic_load(&st->vmove, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            st->rebuild = (first || mv) ? 1 : 0;
            if (first || mv) st->nrebuild += 1;
            __hip_atomic_store(&st->vmove, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&st->vdone, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// All-pairs neighbour build: one wavefront per bead scans every j != i.
// Keeps j with r^2 < rc2 that is not excluded.  Entries are in ascending j.
// The block's N2_WAVES beads scan coordinates staged in LDS (as x, y, z
// arrays, N2_CHUNK beads per pass), the same differences and predicate as
// pair_vec / rsq3.  check_flag 0 (a build on every call): the kernel also
// reports the rebuild (st->rebuild = 1).
__global__ __launch_bounds__(64 * N2_WAVES) void nbr_n2_kernel(const double *__restrict__ crd, int nb, BoxArg bx,
                                                               double rc2, const int *__restrict__ ex_ptr,
                                                               const int *__restrict__ ex_col,
                                                               int *__restrict__ nbr_cnt, int *__restrict__ nbr_idx,
                                                               int cap, int check_flag, double *__restrict__ save,
                                                               DevStatus *st)
{
    if (check_flag && !st->rebuild) return;
    if (!check_flag && blockIdx.x == 0 && threadIdx.x == 0) st->rebuild = 1;
    constexpr int PER = 3 * N2_CHUNK / (64 * N2_WAVES);   // staged values per thread and pass
    static_assert(PER * 64 * N2_WAVES == 3 * N2_CHUNK, "staging split");
    __shared__ double sc[3][N2_CHUNK];
    const int lane = threadIdx.x % 64;
    const int i = blockIdx.x * N2_WAVES + threadIdx.x / 64;
    const bool live = i < nb;   // uniform over the wavefront
    const int ii = live ? i : 0;
    const double xi = crd[3 * ii], yi = crd[3 * ii + 1], zi = crd[3 * ii + 2];
    const int ex0 = ex_ptr[ii], ex1 = ex_ptr[ii + 1];   // the bead's exclusions (ascending)
    int cnt = 0;
    for (int c0 = 0; c0 < nb; c0 += N2_CHUNK) {
        const int cn = min(N2_CHUNK, nb - c0);
        __syncthreads();   // the previous pass no longer reads sc
        double v[PER];   // all loads in flight before the first store
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = threadIdx.x + q * 64 * N2_WAVES;
            v[q] = t < 3 * cn ? crd[3L * c0 + t] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = threadIdx.x + q * 64 * N2_WAVES;
            if (t < 3 * cn) sc[t % 3][t / 3] = v[q];
        }
        __syncthreads();
        if (!live) continue;
        for (int j0 = 0; j0 < cn; j0 += 64) {
            const int jl = j0 + lane, j = c0 + jl;
            bool keep = false;
            if (jl < cn && j != i) {
                double x = __dsub_rn(sc[0][jl], xi), y = __dsub_rn(sc[1][jl], yi), z = __dsub_rn(sc[2][jl], zi);
                if (bx.imcon) {
                    x = img(x, bx.L[0]);
                    y = img(y, bx.L[1]);
                    z = img(z, bx.L[2]);
                }
                if (rsq3(x, y, z) < rc2) {
                    keep = true;
                    for (int k = ex0; k < ex1; k++)
                        if (ex_col[k] == j) keep = false;
                }
            }
            unsigned long long m = __ballot(keep);
            if (keep) {
                int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
                if (pos < cap) nbr_idx[nbr_slot(i, pos, cap)] = j;
            }
            cnt += __popcll(m);
        }
    }
    if (!live) return;
    if (lane == 0) {
        nbr_cnt[i] = cnt;
        if (cnt > cap) atomicExch(&st->overflow, 1);
    }
    if (save && lane < 3) save[3 * i + lane] = crd[3 * i + lane];
}

// ---- cell-list accelerated Verlet build (periodic boxes with >= 3 cells)
__global__ void cell_assign_kernel(const double *__restrict__ crd, int nb, BoxArg bx, int ncx, int ncy, int ncz,
                                   int *__restrict__ bead_cell, int *__restrict__ cell_cnt, int check_flag,
                                   const DevStatus *st)
{
    if (check_flag && !st->rebuild) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    int nc[3] = {ncx, ncy, ncz}, ic[3];
    for (int d = 0; d < 3; d++) {
        double L = bx.L[d];
        double x = crd[3 * i + d];
        double f = x / L;
        f -= floor(f);
        int c = (int)(f * nc[d]);
        if (c >= nc[d]) c = nc[d] - 1;
        if (c < 0) c = 0;
        ic[d] = c;
    }
    int cell = (ic[2] * ncy + ic[1]) * ncx + ic[0];
    bead_cell[i] = cell;
    atomicAdd(&cell_cnt[cell], 1);
}

// exclusive scan of cell counts (single block, any size) + reset fill counters
__global__ __launch_bounds__(1024) void cell_scan_kernel(int *__restrict__ cell_cnt, int *__restrict__ cell_ptr, int ncell,
                                                         int check_flag, const DevStatus *st)
{
    if (check_flag && !st->rebuild) return;
    __shared__ int sums[1024];
    const int t = threadIdx.x;
    const int per = (ncell + 1023) / 1024;
    int b = t * per, e = min(ncell, b + per);
    int s = 0;
    for (int k = b; k < e; k++) s += cell_cnt[k];
    sums[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = (t >= off) ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    int run = sums[t] - s;
    for (int k = b; k < e; k++) {
        cell_ptr[k] = run;
        run += cell_cnt[k];
        cell_cnt[k] = 0;  // reused as fill counter
    }
    if (t == 1023) cell_ptr[ncell] = sums[1023];
}

__global__ void cell_fill_kernel(const int *__restrict__ bead_cell, int nb, const int *__restrict__ cell_ptr,
                                 int *__restrict__ taken, int *__restrict__ out, int check_flag,
                                 const DevStatus *st)
{
    if (check_flag && !st->rebuild) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    int c = bead_cell[i];
    int pos = cell_ptr[c] + atomicAdd(&taken[c], 1);
    out[pos] = i;
}

// One workgroup of NC_WAVES waves per cell.  The candidates of the 27
// surrounding cells are taken in the order of a serial scan (cells dz, dy,
// dx from -1 to 1, beads of a cell in ascending order), so every list is
// the one that scan gives: wave 0's lanes 0-26 read the cells' bead ranges
// and a prefix sum lays them end to end; the candidates and their
// coordinates are staged in LDS, NC_CAP at a time, once for all the beads
// of the cell; each wave then takes beads of the cell in turn and tests a
// staged chunk 64 candidates at a time, packing those that pass in order by
// a ballot (a bead's running count stays in LDS between chunks).  Each
// workgroup also clears its cell's count, which the next build's
// assignment then finds at zero.
constexpr int NC_WAVES = 4;
constexpr int NC_CAP = 512;      // candidates staged at a time
constexpr int NC_BEADS = 256;    // beads of one cell counted at a time
constexpr int NC_EX = 8;         // excluded partners of a bead held in registers
__global__ __launch_bounds__(64 * NC_WAVES) void nbr_cell_kernel(
    const double *__restrict__ crd, int nb, BoxArg bx, double rc2, int ncx, int ncy, int ncz,
    const int *__restrict__ cell_ptr, const int *__restrict__ cell_beads, const int *__restrict__ ex_ptr,
    const int *__restrict__ ex_col, int *__restrict__ nbr_cnt, int *__restrict__ nbr_idx, int cap,
    double *__restrict__ save, int *__restrict__ cell_cnt, DevStatus *st)
{
    __shared__ int s_end[32], s_beg[32];
    __shared__ int s_j[NC_CAP];
    __shared__ double s_x[NC_CAP], s_y[NC_CAP], s_z[NC_CAP];
    __shared__ int s_cnt[NC_BEADS];
    if (!st->rebuild) return;
    const int t = threadIdx.x, lane = t % 64, wv = t / 64;
    const int c = blockIdx.x;
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    if (wv == 0) {
        int b0 = 0, len = 0;
        if (lane < 27) {
            const int dz = lane / 9 - 1, dy = (lane / 3) % 3 - 1, dx = lane % 3 - 1;
            const int jx = (cx + dx + ncx) % ncx, jy = (cy + dy + ncy) % ncy, jz = (cz + dz + ncz) % ncz;
            const int jc = (jz * ncy + jy) * ncx + jx;
            b0 = cell_ptr[jc];
            len = cell_ptr[jc + 1] - b0;
        }
        int incl = len;   // inclusive prefix over lanes 0..31
#pragma unroll
        for (int off = 1; off < 32; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane < 32) {
            s_end[lane] = incl;
            s_beg[lane] = b0 - (incl - len);   // cell_beads index = s_beg + candidate number
        }
    }
    __syncthreads();
    const int total = s_end[26];
    const int m0 = cell_ptr[c], nm = cell_ptr[c + 1] - m0;   // the cell's own beads
    if (t == 0) cell_cnt[c] = 0;   // the fill count of this build: zero again for the next one
    for (int q0 = 0; q0 < nm; q0 += NC_BEADS) {
        const int nq = min(NC_BEADS, nm - q0);
        for (int q = t; q < nq; q += 64 * NC_WAVES) s_cnt[q] = 0;
        for (int k0 = 0; k0 < total; k0 += NC_CAP) {
            const int nk = min(NC_CAP, total - k0);
            __syncthreads();   // the previous chunk's readers are done; s_cnt is set
            for (int e = t; e < nk; e += 64 * NC_WAVES) {
                const int k = k0 + e;
                int lo = 0, hi = 26;   // first cell whose end exceeds k
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (s_end[mid] > k) hi = mid;
                    else lo = mid + 1;
                }
                const int j = cell_beads[s_beg[lo] + k];
                s_j[e] = j;
                s_x[e] = crd[3 * j];
                s_y[e] = crd[3 * j + 1];
                s_z[e] = crd[3 * j + 2];
            }
            __syncthreads();
            for (int q = wv; q < nq; q += NC_WAVES) {
                const int i = cell_beads[m0 + q0 + q];
                const double xi = crd[3 * i], yi = crd[3 * i + 1], zi = crd[3 * i + 2];
                // bead i's excluded partners: the first NC_EX of its row in
                // registers, read once for the whole scan
                const int eb = ex_ptr[i], ee = ex_ptr[i + 1];
                int exl[NC_EX];
#pragma unroll
                for (int u = 0; u < NC_EX; u++) exl[u] = eb + u < ee ? ex_col[eb + u] : -1;
                auto is_excluded = [&](int j) {
                    bool r = false;
#pragma unroll
                    for (int u = 0; u < NC_EX; u++) r = r || exl[u] == j;
                    for (int k = eb + NC_EX; k < ee && !r; k++) r = ex_col[k] == j;
                    return r;
                };
                int cnt = s_cnt[q];
                for (int e0 = 0; e0 < nk; e0 += 64) {
                    const int e = e0 + lane;
                    int j = -1;
                    bool take = false;
                    if (e < nk) {
                        j = s_j[e];
                        if (j != i) {
                            // pair_vec on the staged coordinates (the same operations)
                            double x = __dsub_rn(s_x[e], xi), y = __dsub_rn(s_y[e], yi), z = __dsub_rn(s_z[e], zi);
                            if (bx.imcon) {
                                x = img(x, bx.L[0]);
                                y = img(y, bx.L[1]);
                                z = img(z, bx.L[2]);
                            }
                            take = rsq3(x, y, z) < rc2 && !is_excluded(j);
                        }
                    }
                    const unsigned long long m = __ballot(take);
                    if (take) {
                        const int pos = cnt + __popcll(m & ((1ull << lane) - 1ull));
                        if (pos < cap) nbr_idx[nbr_slot(i, pos, cap)] = j;
                    }
                    cnt += __popcll(m);
                }
                if (lane == 0) s_cnt[q] = cnt;
            }
        }
        __syncthreads();
        for (int q = t; q < nq; q += 64 * NC_WAVES) {
            const int i = cell_beads[m0 + q0 + q];
            const int cnt = s_cnt[q];
            nbr_cnt[i] = cnt;
            if (cnt > cap) atomicExch(&st->overflow, 1);
            save[3 * i] = crd[3 * i];
            save[3 * i + 1] = crd[3 * i + 1];
            save[3 * i + 2] = crd[3 * i + 2];
        }
        __syncthreads();   // s_cnt is reused by the next bead batch
    }
}

// Beads of every cell in ascending index order (the fill above places them
// in atomic order): bead i goes to its cell's start plus the number of the
// cell's beads with a smaller index.  Pair lists, and so the force sums, are
// then the same run to run.
__global__ void cell_rank_kernel(const int *__restrict__ bead_cell, int nb, const int *__restrict__ cell_ptr,
                                 const int *__restrict__ unsorted, int *__restrict__ cell_beads, int check_flag,
                                 const DevStatus *st)
{
    if (check_flag && !st->rebuild) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int c = bead_cell[i], b = cell_ptr[c], e = cell_ptr[c + 1];
    int rank = 0;
    for (int p = b; p < e; p++) rank += unsorted[p] < i ? 1 : 0;
    cell_beads[b + rank] = i;
}

// cl_build cell index (cell_list_m.f90:186-204): cells of side
// cs = L / floor(L / rc), index int(x / cs) truncated and clamped to nc - 1;
// a coordinate with x / cs <= -1 (or NaN) is an error there (the
// reference's cell lists need x >= 0), flagged here as ierr 4.
__global__ void cl_assign_kernel(const double *__restrict__ crd, int nb, double csx, double csy, double csz,
                                 int ncx, int ncy, int ncz, int *__restrict__ bead_cell, int *__restrict__ cell_cnt,
                                 DevStatus *st)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const double cs[3] = {csx, csy, csz};
    const int nc[3] = {ncx, ncy, ncz};
    int ic[3];
    for (int d = 0; d < 3; d++) {
        const double q = crd[3 * i + d] / cs[d];
        if (!(q > -1.0)) { atomicExch(&st->ierr, 4); ic[d] = 0; continue; }
        int k = (int)q;   // truncation toward zero, as Fortran int()
        if (k > nc[d] - 1) k = nc[d] - 1;
        ic[d] = k;
    }
    const int cell = (ic[2] * ncy + ic[1]) * ncx + ic[0];
    bead_cell[i] = cell;
    atomicAdd(&cell_cnt[cell], 1);
}

// pt_build CL (pairtab_m.fpp:305-385): every pair of beads in one cell, and
// every pair between a cell and each of its 13 "forward" neighbours
// (cl_build_cell_nbrs, cell_list_m.f90:206-228), with no distance test and
// with the repeats a grid of fewer than 3 cells per side produces.  As a
// full list: for bead i in cell c, the other beads of c, the beads of
// c + d_k and the beads of c - d_k for the 13 offsets d_k, each time it
// occurs, exclusions skipped.
__constant__ int c_cl_dst[13][3] = {{1, 0, 0},  {1, 1, 0},  {-1, 1, 0}, {0, 1, 0},  {0, 0, 1},
                                    {-1, 0, 1}, {1, 0, 1},  {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                                    {-1, 1, 1}, {0, 1, 1},  {1, 1, 1}};

__global__ __launch_bounds__(256) void cl_nbr_kernel(int nb, int ncx, int ncy, int ncz,
                                                     const int *__restrict__ bead_cell, const int *__restrict__ cell_ptr,
                                                     const int *__restrict__ cell_beads, const int *__restrict__ ex_ptr,
                                                     const int *__restrict__ ex_col, int *__restrict__ nbr_cnt,
                                                     int *__restrict__ nbr_idx, int cap, DevStatus *st)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const int c = bead_cell[i];
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    int cnt = 0;
    auto take = [&](int jc, bool same) {
        for (int p = cell_ptr[jc]; p < cell_ptr[jc + 1]; p++) {
            const int j = cell_beads[p];
            if (same && j == i) continue;
            if (excluded(ex_ptr, ex_col, i, j)) continue;
            if (cnt < cap) nbr_idx[nbr_slot(i, cnt, cap)] = j;
            cnt++;
        }
    };
    take(c, true);
    for (int sgn = 1; sgn >= -1; sgn -= 2)
        for (int k = 0; k < 13; k++) {
            const int jx = ((cx + sgn * c_cl_dst[k][0]) % ncx + ncx) % ncx;
            const int jy = ((cy + sgn * c_cl_dst[k][1]) % ncy + ncy) % ncy;
            const int jz = ((cz + sgn * c_cl_dst[k][2]) % ncz + ncz) % ncz;
            take((jz * ncy + jy) * ncx + jx, false);
        }
    nbr_cnt[i] = cnt;
    if (cnt > cap) atomicExch(&st->overflow, 1);
}

__global__ void zero_int_kernel(int *p, int n, int check_flag, const DevStatus *st)
{
    if (check_flag && !st->rebuild) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ------------------------------------------------------------ adjoint
// dfrc/dr of the bond styles (the derivative of the force magnitude that
// ia_get_bond_force returns) and of the shifted LJ force.
__device__ inline double bond_dforce(double r, int sty, const double *p)
{
    switch (sty) {
    case 1: return p[0];
    case 2: {
        double e = r - p[2], q = e * e / p[3];
        return p[0] * (1.0 + q) / ((1.0 - q) * (1.0 - q));
    }
    case 3: {
        double q = r * r / p[4];
        double d = p[0] * (1.0 + q) / ((1.0 - q) * (1.0 - q));
        if (r < p[5]) {
            double sir = p[3] / r, s6 = sir * sir * sir * sir * sir * sir, s12 = s6 * s6;
            d += 24 * p[2] * (26 * s12 - 7 * s6) / (r * r);
        }
        return d;
    }
    case 4: {
        double x = r / p[1], rr = 1.0 / (1.0 - x);
        return (-1.0 - 0.5 * rr * rr * rr) / (p[1] * p[0]);
    }
    default: return 0.0;
    }
}

__device__ inline double lj_dforce(double r, const double *p)
{
    if (!(r < p[2])) return 0.0;
    double sir = p[1] / r, s6 = sir * sir * sir * sir * sir * sir, s12 = s6 * s6;
    return 24 * p[0] * (26 * s12 - 7 * s6) / (r * r);
}

// J_g(x) u for the pair force g(x) = frc(|x|) x/|x|.
__device__ inline void jg_apply(double x, double y, double z, double r, double f, double df, const double *u,
                                double &ox, double &oy, double &oz)
{
    double a = f / r, b = (df - a) / (r * r);
    double xu = x * u[0] + y * u[1] + z * u[2];
    ox += a * u[0] + b * x * xu;
    oy += a * u[1] + b * y * xu;
    oz += a * u[2] + b * z * xu;
}

// out_i = sum_j J_g(r_j - r_i)(u_j - u_i)  (the force Jacobian, symmetric),
// tpart[blk*3 + m] = sum_i u_i . dF_i/dtheta_m for theta = (eps, sigma) of
// vdw type 1 and k of bond type 1 (FENE/harmonic).  tie_rc: rc = 2^(1/6) sigma.
__global__ __launch_bounds__(FB) void adj_force_kernel(
    const double *__restrict__ crd, int nb, BoxArg bx, const int *__restrict__ bond_ptr,
    const int *__restrict__ bond_idx, const int *__restrict__ bond_ij, const int *__restrict__ bond_type,
    const int *__restrict__ bstyles, const double *__restrict__ bparams, int use_pairs,
    const int *__restrict__ nbr_cnt, const int *__restrict__ nbr_idx, int nbr_cap,
    PairScan ps, const int *__restrict__ ex_ptr,
    const int *__restrict__ ex_col, const int *__restrict__ atype, int natype, const int *__restrict__ ptype,
    const int *__restrict__ vstyles, const double *__restrict__ vparams, int tie_rc, const double *__restrict__ u,
    double *__restrict__ out, double *__restrict__ tpart)
{
    const int i = blockIdx.x * FB + threadIdx.x;
    double t[3] = {0.0, 0.0, 0.0};
    if (i < nb) {
        double ox = 0.0, oy = 0.0, oz = 0.0;
        const double ui[3] = {u[3 * i], u[3 * i + 1], u[3 * i + 2]};
        for (int k = bond_ptr[i]; k < bond_ptr[i + 1]; k++) {
            int b = bond_idx[k];
            int bi = bond_ij[2 * b], bj = bond_ij[2 * b + 1];
            int j = (bi == i) ? bj : bi;
            if (j == i) continue;
            double x, y, z;
            pair_vec(crd, i, j, bx, x, y, z);
            double r = sqrt(x * x + y * y + z * z);
            int ty = bond_type[b];
            const double *p = bparams + ty * MXPARAM;
            double en, f;
            bond_force(r, bstyles[ty], p, en, f);
            double df = bond_dforce(r, bstyles[ty], p);
            double du[3] = {u[3 * j] - ui[0], u[3 * j + 1] - ui[1], u[3 * j + 2] - ui[2]};
            jg_apply(x, y, z, r, f, df, du, ox, oy, oz);
            if (ty == 0 && (bstyles[ty] == 1 || bstyles[ty] == 2) && p[0] != 0.0) {
                double dk = f / p[0];  // both styles are linear in k
                t[2] += dk * (x * ui[0] + y * ui[1] + z * ui[2]) / r;
            }
        }
        if (use_pairs) {
            const int ati = atype[i];
            const int cnt = nbr_cnt[i];
            const bool full = cnt > nbr_cap;
            const int kend = full ? nb : cnt;
            for (int k = 0; k < kend; k++) {
                int j, mult = 1;
                if (full) {
                    j = k;
                    if (j == i) continue;
                    mult = scan_mult(ps, bx, i, j);
                    if (mult == 0 || excluded(ex_ptr, ex_col, i, j)) continue;
                } else {
                    j = nbr_idx[nbr_slot(i, k, nbr_cap)];
                }
                int ty = ptype[ati * natype + atype[j]];
                if (vstyles[ty] != 1) continue;
                const double *p = vparams + ty * MXPARAM;
                double x, y, z;
                pair_vec(crd, i, j, bx, x, y, z);
                double r = sqrt(x * x + y * y + z * z);
                double en, f;
                lj_force(r, p, en, f);
                double df = lj_dforce(r, p);
                f *= mult;   // a pair the list would hold mult times
                df *= mult;
                double du[3] = {u[3 * j] - ui[0], u[3 * j + 1] - ui[1], u[3 * j + 2] - ui[2]};
                jg_apply(x, y, z, r, f, df, du, ox, oy, oz);
                if (ty == 0 && r < p[2]) {
                    double eps = p[0], sig = p[1];
                    double proj = mult * (x * ui[0] + y * ui[1] + z * ui[2]) / r;
                    t[0] += (f / mult / eps) * proj;
                    double sr = sig / r, s6 = sr * sr * sr * sr * sr * sr, s12 = s6 * s6;
                    double dsig = -144.0 * eps * (4 * s12 - s6) / (sig * r);
                    if (!tie_rc) {
                        double sc = sig / p[2], c6 = sc * sc * sc * sc * sc * sc, c12 = c6 * c6;
                        dsig -= -144.0 * eps * (4 * c12 - c6) / (sig * p[2]);
                    }
                    t[1] += dsig * proj;
                }
            }
        }
        out[3 * i] = ox;
        out[3 * i + 1] = oy;
        out[3 * i + 2] = oz;
    }
    __shared__ double sh[FB / 64][3];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    for (int m = 0; m < 3; m++) {
        double v = t[m];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sh[wv][m] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double v = 0.0;
        for (int w = 0; w < FB / 64; w++) v += sh[w][threadIdx.x];
        tpart[blockIdx.x * 3 + threadIdx.x] = v;
    }
}

// ------------------------------------------------------------ host
static std::vector<int> host_exclusions(int nb, const std::vector<int> &bi, const std::vector<int> &bj, int excl,
                                        std::vector<int> &ptr)
{
    // 1-ring adjacency (connectivity_m.f90:202-246), unique and sorted
    std::vector<std::vector<int>> adj(nb);
    for (size_t b = 0; b < bi.size(); b++) {
        if (bi[b] != bj[b]) {
            adj[bi[b]].push_back(bj[b]);
            adj[bj[b]].push_back(bi[b]);
        }
    }
    for (auto &v : adj) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    std::vector<int> col;
    ptr.assign(nb + 1, 0);
    std::vector<int> buf;
    for (int i = 0; i < nb; i++) {
        buf.clear();
        if (excl > 0) {
            buf.push_back(i);
            for (int j : adj[i]) {
                buf.push_back(j);
                if (excl > 1)
                    for (int j2 : adj[j]) {
                        buf.push_back(j2);
                        if (excl > 2)
                            for (int j3 : adj[j2]) buf.push_back(j3);
                    }
            }
        }
        std::sort(buf.begin(), buf.end());
        buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
        col.insert(col.end(), buf.begin(), buf.end());
        ptr[i + 1] = (int)col.size();
    }
    return col;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &v)
{
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    BD_HIP(hipMalloc(dst, bytes));
    if (!v.empty()) BD_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return BD_OK;
}

// The derived slots of ia_vdw_setup's setters (ia_vdw_m.f90:85-151, 185-206,
// 270-306, 356-389): cut-off energies (and LJ's cut-off force) per style.
static void vdw_derived(int sty, double *p)
{
    auto lj_at = [](double eps, double sigma, double rc) {
        const double sir = sigma / rc, sir2 = sir * sir, sir6 = sir2 * sir2 * sir2, sir12 = sir6 * sir6;
        return std::make_pair(4 * eps * (sir12 - sir6), -24 * eps * (2 * sir12 - sir6) / rc);
    };
    switch (sty) {
    case 1: {  // vdw_lj_set
        auto v = lj_at(p[0], p[1], p[2]);
        p[3] = v.first;
        p[4] = v.second;
        break;
    }
    case 2: {  // vdw_gaussian_set
        const double ex = std::exp(-p[1] * (p[2] * p[2]));
        p[3] = p[0] * ex;
        p[4] = -2.0 * p[0] * p[1] * p[2] * ex;
        break;
    }
    case 4:  // vdw_lj_coul_debye_set
        p[6] = lj_at(p[0], p[1], p[2]).first;
        p[7] = p[3] > 0.0 ? p[4] * std::exp(-p[5] * p[3]) / p[3] : 0.0;
        break;
    case 5:  // vdw_lj_coul_set
        p[5] = lj_at(p[0], p[1], p[2]).first;
        p[6] = p[3] > 0.0 ? p[4] / p[3] : 0.0;
        break;
    default:
        break;
    }
}

int forces_setup(Ctx &c, int nbonds, const int *bonds, int nbond_types, const int *bond_styles,
                 const double *bond_params, int natom_types, const int *atom_types, int nvdw_types,
                 const int *vdw_styles, const double *vdw_params)
{
    const int nb = c.nb;
    std::vector<int> bij(2 * std::max(nbonds, 0)), btype(std::max(nbonds, 0));
    std::vector<int> bi(nbonds), bj(nbonds);
    for (int b = 0; b < nbonds; b++) {
        int t = bonds[3 * b] - 1, i = bonds[3 * b + 1] - 1, j = bonds[3 * b + 2] - 1;
        if (t < 0 || t >= nbond_types || i < 0 || i >= nb || j < 0 || j >= nb) {
            set_error("bond " + std::to_string(b + 1) + " has an out-of-range type or bead index");
            return BD_ERR_ARG;
        }
        bij[2 * b] = i; bij[2 * b + 1] = j; btype[b] = t;
        bi[b] = i; bj[b] = j;
    }
    std::vector<int> bstyle(std::max(nbond_types, 1), 0);
    std::vector<double> bpar((size_t)std::max(nbond_types, 1) * MXPARAM, 0.0);
    for (int t = 0; t < nbond_types; t++) {
        bstyle[t] = bond_styles[t];
        if (bstyle[t] < 0 || bstyle[t] > 4) { set_error("bond style must be 0..4"); return BD_ERR_ARG; }
        for (int k = 0; k < MXPARAM; k++) bpar[t * MXPARAM + k] = bond_params[t * MXPARAM + k];
        double *p = &bpar[t * MXPARAM];
        if (bstyle[t] == 2) p[3] = p[1] * p[1];                         // bond_fene_set
        if (bstyle[t] == 3) { p[4] = p[1] * p[1]; p[5] = 1.122462048309373 * p[3]; }  // bond_kg_set
    }
    // bead -> bonds CSR in bond order (atbo_build, connectivity_m.f90:43-90)
    std::vector<int> bptr(nb + 1, 0), bidx;
    {
        std::vector<std::vector<int>> per(nb);
        for (int b = 0; b < nbonds; b++) {
            per[bi[b]].push_back(b);
            if (bj[b] != bi[b]) per[bj[b]].push_back(b);
        }
        for (int i = 0; i < nb; i++) {
            bidx.insert(bidx.end(), per[i].begin(), per[i].end());
            bptr[i + 1] = (int)bidx.size();
        }
    }
    std::vector<int> atype(nb, 0);
    for (int i = 0; i < nb; i++) {
        atype[i] = atom_types ? atom_types[i] - 1 : 0;
        if (atype[i] < 0 || atype[i] >= std::max(natom_types, 1)) {
            set_error("atom type out of range");
            return BD_ERR_ARG;
        }
    }
    const int nat = std::max(natom_types, 1);
    std::vector<int> ptype(nat * nat, 0);
    for (int a = 1; a <= nat; a++)
        for (int b = 1; b <= nat; b++) {
            // interaction_m.f90:182-186
            int typ = (a < b) ? b + (2 * nat - a) * (a - 1) / 2 : a + (2 * nat - b) * (b - 1) / 2;
            ptype[(a - 1) * nat + (b - 1)] = typ - 1;
        }
    std::vector<int> vstyle(std::max(nvdw_types, 1), 0);
    bool general = false;
    std::vector<double> vpar((size_t)std::max(nvdw_types, 1) * MXPARAM, 0.0);
    if (c.lvdw && nvdw_types > 0 && nvdw_types < nat * (nat + 1) / 2) {
        set_error("VDW section must list all nat*(nat+1)/2 type pairs");
        return BD_ERR_ARG;
    }
    for (int t = 0; t < nvdw_types; t++) {
        vstyle[t] = vdw_styles[t];
        if (vstyle[t] < 0 || vstyle[t] > 6) {
            set_error("vdw style " + std::to_string(vstyle[t]) + " is not one of 0..6 (ia_vdw_m.f90:21-47)");
            return BD_ERR_ARG;
        }
        for (int k = 0; k < MXPARAM; k++) vpar[t * MXPARAM + k] = vdw_params[t * MXPARAM + k];
        vdw_derived(vstyle[t], &vpar[t * MXPARAM]);
        if (vstyle[t] >= 2) general = true;
    }
    std::vector<int> eptr;
    std::vector<int> ecol = host_exclusions(nb, bi, bj, c.excluded_atoms, eptr);

    c.h_bpar = bpar;
    c.h_vpar = vpar;
    c.h_bsty = bstyle;
    c.h_vsty = vstyle;
    c.nbonds = nbonds;
    c.nbond_types = nbond_types;
    c.natom_types = nat;
    c.nvdw_types = nvdw_types;
    c.vdw_general = general;
    BD_TRY(upload(&c.bond_ij, bij));
    BD_TRY(upload(&c.bond_type, btype));
    BD_TRY(upload(&c.bead_bond_ptr, bptr));
    BD_TRY(upload(&c.bead_bond_idx, bidx));
    BD_TRY(upload(&c.bond_styles, bstyle));
    BD_TRY(upload(&c.bond_params, bpar));
    BD_TRY(upload(&c.atom_types, atype));
    BD_TRY(upload(&c.vdw_styles, vstyle));
    BD_TRY(upload(&c.vdw_params, vpar));
    BD_TRY(upload(&c.ex_ptr, eptr));
    BD_TRY(upload(&c.ex_col, ecol));
    BD_TRY(upload(&c.scan_tmp, ptype));  // pair-type table (nat x nat)
    c.ex_total = (int)ecol.size();

    if (!c.nbr_cnt) {
        BD_HIP(hipMalloc(&c.nbr_cnt, (size_t)std::max(nb, 1) * sizeof(int)));
        BD_HIP(hipMemset(c.nbr_cnt, 0, (size_t)std::max(nb, 1) * sizeof(int)));
    }
    if (!c.nbr_idx) {
        c.pair_cap = NBR_CAP0;
        BD_HIP(hipMalloc(&c.nbr_idx, (size_t)c.pair_cap * std::max(nb, 1) * sizeof(int)));
    }
    if (!c.crd_save) BD_HIP(hipMalloc(&c.crd_save, (size_t)3 * std::max(nb, 1) * sizeof(double)));
    int nblk = (nb + FW - 1) / FW;   // force_gather_kernel blocks
    if (c.red_cap < nblk * NACC) {
        if (c.red) (void)hipFree(c.red);
        c.red_cap = std::max(nblk, 1) * NACC;
        BD_HIP(hipMalloc(&c.red, (size_t)c.red_cap * sizeof(double)));
    }
    c.verlet_first = true;
    c.pairs_valid = false;
    // periodic cell grid for the Verlet build (cells of at least rc + skin)
    c.ncell_tot = 0;
    if (c.mth_ptgen == BD_PT_CL && c.lvdw && nvdw_types > 0) {
        // cl_set_cell_size (cell_list_m.f90:91-132): nc = floor(L / rc)
        if (!c.imcon) { set_error("mth_ptgen CL needs a periodic box (imcon 1)"); return BD_ERR_ARG; }
        int nc[3];
        for (int d = 0; d < 3; d++) nc[d] = (int)std::floor(c.box[d] / c.rcutoff);
        if (nc[0] < 2 || nc[1] < 2 || nc[2] < 2) {
            // one cell along an axis would pair every bead with itself
            set_error("mth_ptgen CL needs at least 2 cells of rcutoff per box edge");
            return BD_ERR_ARG;
        }
        for (int d = 0; d < 3; d++) {
            c.ncell[d] = nc[d];
            c.cl_cs[d] = c.box[d] / nc[d];
        }
        c.ncell_tot = nc[0] * nc[1] * nc[2];
        if (c.cell_cnt) (void)hipFree(c.cell_cnt);
        if (c.cell_ptr) (void)hipFree(c.cell_ptr);
        if (!c.cell_beads) BD_HIP(hipMalloc(&c.cell_beads, (size_t)nb * sizeof(int)));
        if (!c.cell_fill) BD_HIP(hipMalloc(&c.cell_fill, (size_t)nb * sizeof(int)));
        if (!c.bead_cell) BD_HIP(hipMalloc(&c.bead_cell, (size_t)nb * sizeof(int)));
        BD_HIP(hipMalloc(&c.cell_cnt, (size_t)c.ncell_tot * sizeof(int)));
        BD_HIP(hipMalloc(&c.cell_ptr, (size_t)(c.ncell_tot + 1) * sizeof(int)));
    }
    if (c.imcon && c.mth_ptgen == BD_PT_VER) {
        double rl = c.rcutoff + c.tskin;
        int nc[3];
        for (int d = 0; d < 3; d++) nc[d] = (int)std::floor(c.box[d] / rl);
        if (nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3) {
            for (int d = 0; d < 3; d++) c.ncell[d] = nc[d];
            c.ncell_tot = nc[0] * nc[1] * nc[2];
            if (c.cell_cnt) (void)hipFree(c.cell_cnt);
            if (c.cell_ptr) (void)hipFree(c.cell_ptr);
            if (!c.cell_beads) BD_HIP(hipMalloc(&c.cell_beads, (size_t)nb * sizeof(int)));
            if (!c.cell_fill) BD_HIP(hipMalloc(&c.cell_fill, (size_t)nb * sizeof(int)));
            if (!c.bead_cell) BD_HIP(hipMalloc(&c.bead_cell, (size_t)nb * sizeof(int)));
            BD_HIP(hipMalloc(&c.cell_cnt, (size_t)c.ncell_tot * sizeof(int)));
            BD_HIP(hipMemset(c.cell_cnt, 0, (size_t)c.ncell_tot * sizeof(int)));
            BD_HIP(hipMalloc(&c.cell_ptr, (size_t)(c.ncell_tot + 1) * sizeof(int)));
        }
    }
    c.have_topology = true;
    return BD_OK;
}

void forces_free(Ctx &c)
{
    int **ip[] = {&c.bond_ij, &c.bond_type, &c.bead_bond_ptr, &c.bead_bond_idx, &c.bond_styles, &c.atom_types,
                  &c.vdw_styles, &c.ex_ptr, &c.ex_col, &c.nbr_cnt, &c.nbr_ptr, &c.nbr_idx, &c.cell_cnt,
                  &c.cell_ptr, &c.cell_beads, &c.cell_fill, &c.bead_cell, &c.scan_tmp};
    for (auto p : ip) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    int **ip2[] = {&c.ang_ptr, &c.ang_idx, &c.ang_atoms, &c.ang_type, &c.ang_styles, &c.teth_ptr, &c.teth_idx,
                   &c.teth_type, &c.teth_styles, &c.ext_styles};
    for (auto p : ip2) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    double **dp[] = {&c.bond_params, &c.vdw_params, &c.crd_save, &c.red, &c.ang_params, &c.teth_pts,
                     &c.teth_params, &c.ext_params, &c.charge};
    for (auto p : dp) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
}

static BoxArg box_arg(const Ctx &c)
{
    BoxArg b;
    b.imcon = c.imcon;
    for (int d = 0; d < 3; d++) b.L[d] = c.box[d];
    return b;
}

__global__ void set_rebuild_kernel(DevStatus *st) { st->rebuild = 1; }

// Builds (or keeps) the neighbour list for the current coordinates.  With
// sync the overflow flag is checked on the host and the capacity grown;
// without it (the step loop) nothing waits: beads whose list overflowed are
// handled by the all-pairs fallback of force_gather_kernel.
static int build_pairs(Ctx &c, bool sync)
{
    const int nb = c.nb;
    hipStream_t sm = c.stream;
    BoxArg bx = box_arg(c);
    const bool ver = (c.mth_ptgen == BD_PT_VER);
    const double rl = ver ? c.rcutoff + c.tskin : c.rcutoff;
    const double rc2 = rl * rl;
    for (int attempt = 0; attempt < 8; attempt++) {
        // DIR builds on every call and nbr_n2_kernel reports that itself
        const bool n2_dir = !ver && c.mth_ptgen != BD_PT_CL;
        if (ver)
            hipLaunchKernelGGL(verlet_check_kernel, dim3(std::max((nb + 255) / 256, 1)), dim3(256), 0, sm, c.crd,
                               c.crd_save, nb, bx,
                               c.tskin * c.tskin, c.verlet_first ? 1 : 0, c.st);
        else if (!n2_dir)
            hipLaunchKernelGGL(set_rebuild_kernel, dim3(1), dim3(1), 0, sm, c.st);  // CL: every call
        if (c.mth_ptgen == BD_PT_CL) {
            const int nbk = (nb + 255) / 256, nck = (c.ncell_tot + 255) / 256;
            hipLaunchKernelGGL(zero_int_kernel, dim3(nck), dim3(256), 0, sm, c.cell_cnt, c.ncell_tot, 0, c.st);
            hipLaunchKernelGGL(cl_assign_kernel, dim3(nbk), dim3(256), 0, sm, c.crd, nb, c.cl_cs[0], c.cl_cs[1],
                               c.cl_cs[2], c.ncell[0], c.ncell[1], c.ncell[2], c.bead_cell, c.cell_cnt, c.st);
            hipLaunchKernelGGL(cell_scan_kernel, dim3(1), dim3(1024), 0, sm, c.cell_cnt, c.cell_ptr, c.ncell_tot, 0,
                               c.st);
            hipLaunchKernelGGL(cell_fill_kernel, dim3(nbk), dim3(256), 0, sm, c.bead_cell, nb, c.cell_ptr, c.cell_cnt,
                               c.cell_fill, 0, c.st);
            hipLaunchKernelGGL(cell_rank_kernel, dim3(nbk), dim3(256), 0, sm, c.bead_cell, nb, c.cell_ptr, c.cell_fill,
                               c.cell_beads, 0, c.st);
            hipLaunchKernelGGL(cl_nbr_kernel, dim3(nbk), dim3(256), 0, sm, nb, c.ncell[0], c.ncell[1], c.ncell[2],
                               c.bead_cell, c.cell_ptr, c.cell_beads, c.ex_ptr, c.ex_col, c.nbr_cnt, c.nbr_idx,
                               c.pair_cap, c.st);
        } else if (ver && c.ncell_tot > 0) {
            int nbk = (nb + 255) / 256;
            // cell_cnt is zero here: set so at allocation, and each cell's
            // workgroup of nbr_cell_kernel clears its count after a build
            hipLaunchKernelGGL(cell_assign_kernel, dim3(nbk), dim3(256), 0, sm, c.crd, nb, bx, c.ncell[0], c.ncell[1],
                               c.ncell[2], c.bead_cell, c.cell_cnt, 1, c.st);
            hipLaunchKernelGGL(cell_scan_kernel, dim3(1), dim3(1024), 0, sm, c.cell_cnt, c.cell_ptr, c.ncell_tot, 1,
                               c.st);
            hipLaunchKernelGGL(cell_fill_kernel, dim3(nbk), dim3(256), 0, sm, c.bead_cell, nb, c.cell_ptr, c.cell_cnt,
                               c.cell_fill, 1, c.st);
            hipLaunchKernelGGL(cell_rank_kernel, dim3(nbk), dim3(256), 0, sm, c.bead_cell, nb, c.cell_ptr, c.cell_fill,
                               c.cell_beads, 1, c.st);
            hipLaunchKernelGGL(nbr_cell_kernel, dim3(c.ncell_tot), dim3(64 * NC_WAVES), 0, sm, c.crd, nb, bx, rc2,
                               c.ncell[0], c.ncell[1], c.ncell[2], c.cell_ptr, c.cell_beads, c.ex_ptr, c.ex_col,
                               c.nbr_cnt, c.nbr_idx, c.pair_cap, c.crd_save, c.cell_cnt, c.st);
        } else {
            hipLaunchKernelGGL(nbr_n2_kernel, dim3((nb + N2_WAVES - 1) / N2_WAVES), dim3(64 * N2_WAVES), 0, sm, c.crd,
                               nb, bx, rc2, c.ex_ptr,
                               c.ex_col, c.nbr_cnt, c.nbr_idx, c.pair_cap, n2_dir ? 0 : 1, ver ? c.crd_save : nullptr,
                               c.st);
        }
        BD_HIP(hipGetLastError());
        c.verlet_first = false;
        c.pairs_valid = true;
        if (!sync) return BD_OK;
        BD_HIP(hipMemcpyAsync(c.st_host, c.st, sizeof(DevStatus), hipMemcpyDeviceToHost, sm));
        BD_HIP(hipStreamSynchronize(sm));
        c.last_rebuilt = c.st_host->rebuild;
        if (!c.st_host->overflow) return BD_OK;
        // grow and rebuild from scratch
        BD_TRY(grow_pairs(c));
        c.verlet_first = true;
    }
    set_error("neighbour list capacity could not be satisfied");
    return BD_ERR_CAPACITY;
}

// Doubles the list capacity.  The new list holds nothing, so the next build
// is forced (a Verlet check would otherwise keep the old counts over it).
int grow_pairs(Ctx &c)
{
    int zero = 0;
    BD_HIP(hipStreamSynchronize(c.stream));
    BD_HIP(hipMemcpy(&c.st->overflow, &zero, sizeof(int), hipMemcpyHostToDevice));
    (void)hipFree(c.nbr_idx);
    c.nbr_idx = nullptr;
    c.pair_cap *= 2;
    BD_HIP(hipMalloc(&c.nbr_idx, (size_t)c.pair_cap * c.nb * sizeof(int)));
    c.verlet_first = true;
    c.pairs_valid = false;
    return BD_OK;
}

static PairScan pair_scan(const Ctx &c)
{
    const bool ver = (c.mth_ptgen == BD_PT_VER), cl = (c.mth_ptgen == BD_PT_CL);
    const double rl = ver ? c.rcutoff + c.tskin : c.rcutoff;
    return PairScan{ver ? c.crd_save : c.crd, rl * rl, cl ? c.bead_cell : nullptr, {c.ncell[0], c.ncell[1], c.ncell[2]}};
}

int compute_forces(Ctx &c, bool want_energy, bool sync)
{
    if (!c.have_topology) { set_error("topology not set"); return BD_ERR_STATE; }
    const int nb = c.nb;
    const bool pairs = c.lvdw && c.nvdw_types > 0;
    if (pairs) BD_TRY(build_pairs(c, sync));
    const bool ver = (c.mth_ptgen == BD_PT_VER);
    const double rl = ver ? c.rcutoff + c.tskin : c.rcutoff;
    BoxArg bx = box_arg(c);
    // lanes per bead: a group about as wide as the list is long (the
    // capacity bounds every bead's listed count; a bead whose list
    // overflowed is scanned against all beads by the same group)
    const int lpb = !pairs ? 8 : (c.pair_cap <= 32 ? 8 : (c.pair_cap <= 128 ? 16 : 64));
    int nblk = (nb + FB / lpb - 1) / (FB / lpb);
    if ((long)std::max(nblk, 1) * NACC > c.red_cap) {
        set_error("internal: force partial buffer smaller than the launch");
        return BD_ERR_STATE;
    }
    ExtraArg xa{c.nangles, c.ntethers, c.ang_ptr, c.ang_idx, c.ang_atoms, c.ang_type, c.ang_styles, c.ang_params, c.teth_ptr, c.teth_idx,
                c.teth_type, c.teth_styles, c.teth_pts, c.teth_params, c.nexternals, c.ext_styles, c.ext_params,
                c.charge};
    const bool extra = c.nangles > 0 || c.ntethers > 0 || c.nexternals > 0;
    auto pick = [&](auto lanes) {
        constexpr int L = decltype(lanes)::value;
        return extra ? (c.vdw_general ? force_gather_kernel<true, true, L> : force_gather_kernel<true, false, L>)
                     : (c.vdw_general ? force_gather_kernel<false, true, L> : force_gather_kernel<false, false, L>);
    };
    auto kern = lpb == 8 ? pick(std::integral_constant<int, 8>{})
                         : (lpb == 16 ? pick(std::integral_constant<int, 16>{}) : pick(std::integral_constant<int, 64>{}));
    hipLaunchKernelGGL(kern, dim3(std::max(nblk, 1)), dim3(FB), 0, c.stream, c.crd, nb, bx, c.bead_bond_ptr,
                       c.bead_bond_idx, c.bond_ij, c.bond_type, c.bond_styles, c.bond_params, pairs ? 1 : 0,
                       c.nbr_cnt, c.nbr_idx, c.pair_cap, pair_scan(c), c.ex_ptr, c.ex_col,
                       c.atom_types, c.natom_types, c.scan_tmp, c.vdw_styles, c.vdw_params, xa, c.frc,
                       want_energy ? 1 : 0, c.red, c.st);
    BD_HIP(hipGetLastError());
    if (want_energy) {
        double vol = c.box[0] * c.box[1] * c.box[2];
        hipLaunchKernelGGL(energy_final_kernel, dim3(1), dim3(64), 0, c.stream, c.red, std::max(nblk, 1), c.nbonds,
                           c.imcon, vol, c.scal);
        BD_HIP(hipGetLastError());
    }
    return BD_OK;
}

// Force Jacobian-vector product and theta partials at the device
// coordinates (the neighbour list must be the one of the last compute_forces).
int adj_force(Ctx &c, const double *u, double *out, double *tpart, int tie_rc)
{
    const int nb = c.nb;
    const bool pairs = c.lvdw && c.nvdw_types > 0;
    const bool ver = (c.mth_ptgen == BD_PT_VER);
    const double rl = ver ? c.rcutoff + c.tskin : c.rcutoff;
    BoxArg bx = box_arg(c);
    int nblk = std::max((nb + FB - 1) / FB, 1);
    hipLaunchKernelGGL(adj_force_kernel, dim3(nblk), dim3(FB), 0, c.stream, c.crd, nb, bx, c.bead_bond_ptr,
                       c.bead_bond_idx, c.bond_ij, c.bond_type, c.bond_styles, c.bond_params, pairs ? 1 : 0,
                       c.nbr_cnt, c.nbr_idx, c.pair_cap, pair_scan(c), c.ex_ptr, c.ex_col,
                       c.atom_types, c.natom_types, c.scan_tmp, c.vdw_styles, c.vdw_params, tie_rc, u, out, tpart);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

// Updates eps, sigma (vdw type 1; rc = 2^(1/6) sigma when tie_rc) and k
// (bond type 1) on the device, re-deriving the shifted-force constants.
int set_theta(Ctx &c, double eps, double sigma, double kb, int tie_rc)
{
    if (c.h_vpar.size() >= (size_t)MXPARAM && c.nvdw_types > 0 && c.h_vsty[0] == 1) {
        double *p = &c.h_vpar[0];
        p[0] = eps;
        p[1] = sigma;
        if (tie_rc) p[2] = 1.122462048309373 * sigma;
        double sir = p[1] / p[2], sir2 = sir * sir, sir6 = sir2 * sir2 * sir2, sir12 = sir6 * sir6;
        p[3] = 4 * p[0] * (sir12 - sir6);
        p[4] = -24 * p[0] * (2 * sir12 - sir6) / p[2];
        BD_HIP(hipMemcpy(c.vdw_params, c.h_vpar.data(), c.h_vpar.size() * sizeof(double), hipMemcpyHostToDevice));
        if (tie_rc) c.rcutoff = p[2];
    }
    if (c.h_bpar.size() >= (size_t)MXPARAM && c.nbond_types > 0) {
        c.h_bpar[0] = kb;
        BD_HIP(hipMemcpy(c.bond_params, c.h_bpar.data(), c.h_bpar.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    c.verlet_first = true;
    return BD_OK;
}

// ---------------------------------------------- angles, tethers, externals

// Parameters of a bonded-term type table (12 x ntypes), styles in lo..hi.
static int check_types(const char *what, int ntypes, const int *styles, const double *params, int lo, int hi)
{
    if (ntypes > 0 && (!styles || !params)) {
        set_error(std::string(what) + ": null style or parameter table");
        return BD_ERR_ARG;
    }
    for (int t = 0; t < ntypes; t++)
        if (styles[t] < lo || styles[t] > hi) {
            set_error(std::string(what) + " style " + std::to_string(styles[t]) + " is not one of " +
                      std::to_string(lo) + ".." + std::to_string(hi));
            return BD_ERR_ARG;
        }
    return BD_OK;
}

int set_angles(Ctx &c, int nangles, const int *angles, int ntypes, const int *styles, const double *params)
{
    if (nangles < 0 || ntypes < 0 || (nangles > 0 && !angles)) { set_error("bad angle arguments"); return BD_ERR_ARG; }
    BD_TRY(check_types("angle", ntypes, styles, params, 0, 1));
    const int nb = c.nb;
    std::vector<int> atoms(3 * (size_t)nangles), type(nangles);
    std::vector<std::vector<int>> per(nb);
    for (int a = 0; a < nangles; a++) {
        const int t = angles[4 * a] - 1;
        if (t < 0 || t >= ntypes) { set_error("angle " + std::to_string(a + 1) + ": type out of range"); return BD_ERR_ARG; }
        type[a] = t;
        for (int r = 0; r < 3; r++) {
            const int b = angles[4 * a + 1 + r] - 1;
            if (b < 0 || b >= nb) { set_error("angle " + std::to_string(a + 1) + ": bead out of range"); return BD_ERR_ARG; }
            atoms[3 * a + r] = b;
            per[b].push_back(4 * a + r);
        }
    }
    std::vector<int> ptr(nb + 1, 0), idx;
    for (int i = 0; i < nb; i++) {
        idx.insert(idx.end(), per[i].begin(), per[i].end());
        ptr[i + 1] = (int)idx.size();
    }
    std::vector<int> sty(styles, styles + ntypes);
    std::vector<double> par(params, params + (size_t)ntypes * MXPARAM);   // ang_cos_set: k = params(1)
    BD_TRY(upload(&c.ang_ptr, ptr));
    BD_TRY(upload(&c.ang_idx, idx));
    BD_TRY(upload(&c.ang_atoms, atoms));
    BD_TRY(upload(&c.ang_type, type));
    BD_TRY(upload(&c.ang_styles, sty));
    BD_TRY(upload(&c.ang_params, par));
    c.nangles = nangles;
    return BD_OK;
}

int set_tethers(Ctx &c, int ntethers, const int *tethers, const double *points, int ntypes, const int *styles,
                const double *params)
{
    if (ntethers < 0 || ntypes < 0 || (ntethers > 0 && (!tethers || !points))) {
        set_error("bad tether arguments");
        return BD_ERR_ARG;
    }
    BD_TRY(check_types("tether", ntypes, styles, params, 0, 2));
    const int nb = c.nb;
    std::vector<int> type(ntethers);
    std::vector<double> pts(points, points + 3 * (size_t)ntethers);
    std::vector<std::vector<int>> per(nb);
    for (int k = 0; k < ntethers; k++) {
        const int t = tethers[2 * k] - 1, b = tethers[2 * k + 1] - 1;
        if (t < 0 || t >= ntypes || b < 0 || b >= nb) {
            set_error("tether " + std::to_string(k + 1) + ": type or bead out of range");
            return BD_ERR_ARG;
        }
        type[k] = t;
        per[b].push_back(k);
    }
    std::vector<int> ptr(nb + 1, 0), idx;
    for (int i = 0; i < nb; i++) {
        idx.insert(idx.end(), per[i].begin(), per[i].end());
        ptr[i + 1] = (int)idx.size();
    }
    std::vector<int> sty(styles, styles + ntypes);
    std::vector<double> par(params, params + (size_t)ntypes * MXPARAM);
    BD_TRY(upload(&c.teth_ptr, ptr));
    BD_TRY(upload(&c.teth_idx, idx));
    BD_TRY(upload(&c.teth_type, type));
    BD_TRY(upload(&c.teth_styles, sty));
    BD_TRY(upload(&c.teth_params, par));
    BD_TRY(upload(&c.teth_pts, pts));
    c.ntethers = ntethers;
    return BD_OK;
}

int set_externals(Ctx &c, int nexternals, const int *styles, const double *params)
{
    if (nexternals < 0) { set_error("bad external-field arguments"); return BD_ERR_ARG; }
    BD_TRY(check_types("external field", nexternals, styles, params, 0, 2));
    for (int e = 0; e < nexternals; e++) {
        const double *p = params + (size_t)e * MXPARAM;
        if (styles[e] == 1 && ((int)p[0] < 1 || (int)p[0] > c.nb)) {
            set_error("external field " + std::to_string(e + 1) + ": atom index out of range");
            return BD_ERR_ARG;
        }
        if (styles[e] == 2 && ((int)p[0] < 1 || (int)p[0] > 3)) {
            set_error("external field " + std::to_string(e + 1) + ": component must be 1..3");
            return BD_ERR_ARG;
        }
    }
    std::vector<int> sty(styles, styles + nexternals);
    std::vector<double> par(params, params + (size_t)nexternals * MXPARAM);
    BD_TRY(upload(&c.ext_styles, sty));
    BD_TRY(upload(&c.ext_params, par));
    c.nexternals = nexternals;
    return BD_OK;
}

int set_charges(Ctx &c, const double *charge)
{
    if (!charge) {
        if (c.charge) (void)hipFree(c.charge);
        c.charge = nullptr;
        return BD_OK;
    }
    std::vector<double> q(charge, charge + c.nb);
    return upload(&c.charge, q);
}

int pair_table_host(Ctx &c, std::vector<int> &pairs)
{
    pairs.clear();
    if (!c.pairs_valid) { set_error("no pair table has been built"); return BD_ERR_STATE; }
    const int nb = c.nb;
    std::vector<int> cnt(nb), idx((size_t)c.pair_cap * nb);
    BD_HIP(hipStreamSynchronize(c.stream));
    BD_HIP(hipMemcpy(cnt.data(), c.nbr_cnt, nb * sizeof(int), hipMemcpyDeviceToHost));
    BD_HIP(hipMemcpy(idx.data(), c.nbr_idx, idx.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < nb; i++) {
        if (cnt[i] > c.pair_cap) { set_error("pair list overflowed; call bd_forces first"); return BD_ERR_CAPACITY; }
        std::vector<int> row;
        for (int k = 0; k < cnt[i]; k++) {
            int j = idx[(size_t)i * c.pair_cap + k];
            if (j > i) row.push_back(j);
        }
        std::sort(row.begin(), row.end());
        for (int j : row) {
            pairs.push_back(i + 1);
            pairs.push_back(j + 1);
        }
    }
    return BD_OK;
}

}  // namespace bd
