// bd_stats.hip -- run statistics accumulated on the device every step.
//
// Restates stats_accumulate, compute_ic0 and compute_ic1 (stats_m.f90:
// 299-488) for linear molecules (no branches): after every step the device
// computes the instantaneous structure of the end-of-step coordinates
//   imcon 0 (compute_ic0): one molecule of all beads -- span (max - min per
//     axis), end-to-end vector reev = r_N - r_1, reedsq = |reev|^2,
//     rgsq = mean |r_i - com|^2;
//   imcon 1 (compute_ic1): per molecule of more than one bead, the same
//     quantities of its beads after the reference's imaging loop, averaged
//     over those molecules.  That loop (stats_m.f90:383-386) subtracts the
//     first bead in place, first bead included, so the first bead becomes 0
//     before the others are shifted: beads 2..natm are imaged about the
//     origin, not about bead 1.  The statistics restate that as it is;
// and updates running means over bins of stats_binsize(1) steps (energies,
// rgsq, reedsq, span) and stats_binsize(2) steps (stress), in the order and
// with the update formula of stats_accumulate (accu += (x - accu) / bnpop).
// The energies, stress and bond lengths are those of the step's force
// evaluation (calc_drift, at the start-of-step coordinates), as in bds_run.
// Sums run in a fixed order, so the statistics repeat bit for bit.
#include "bd_forces_dev.h"
#include "bd_internal.h"

namespace bd {

// device state (double words): [0] bnpop, [1] bnpop_stress, [2..8] energy
// accumulators (bond, angle, dihedral, vdw, tether, external, total),
// [9..17] stress accumulator (3 x 3, column-major), [18] rgsq, [19] reedsq,
// [20..22] span accumulators, [23..25] reev, [26..28] bndlen mean/min/max;
// [32..39] the instantaneous structure of the step (rgsq, reedsq, span(3),
// reev(3)).
constexpr int ST_WORDS = 40;
constexpr int ST_INST = 32;
constexpr int SB = 1024;   // threads of the single-molecule kernel

// fixed-order block reduction of n values per thread (sum, min or max)
template <int NV>
__device__ inline void block_reduce(double (&v)[NV], int op, double *sh)
{
    // op 0: sum, 1: min, 2: max; sh: NV * blockDim.x doubles
    const int t = threadIdx.x, nt = blockDim.x;
#pragma unroll
    for (int q = 0; q < NV; q++) sh[q * nt + t] = v[q];
    __syncthreads();
    for (int h = nt / 2; h > 0; h >>= 1) {
        if (t < h) {
#pragma unroll
            for (int q = 0; q < NV; q++) {
                const double a = sh[q * nt + t], b = sh[q * nt + t + h];
                sh[q * nt + t] = op == 0 ? a + b : (op == 1 ? fmin(a, b) : fmax(a, b));
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = sh[q * nt];
    __syncthreads();
}

// compute_ic0 (stats_m.f90:409-488 without branches): one block.
__global__ __launch_bounds__(SB) void stats_ic0_kernel(const double *__restrict__ crd, int nb, double *st)
{
    __shared__ double sh[3 * SB];
    double s[3] = {0, 0, 0}, mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) {
            const double x = crd[3 * i + d];
            s[d] += x;
            mn[d] = fmin(mn[d], x);
            mx[d] = fmax(mx[d], x);
        }
    block_reduce<3>(s, 0, sh);
    block_reduce<3>(mn, 1, sh);
    block_reduce<3>(mx, 2, sh);
    const double com[3] = {s[0] / nb, s[1] / nb, s[2] / nb};
    double g[1] = {0.0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const double dx = crd[3 * i] - com[0], dy = crd[3 * i + 1] - com[1], dz = crd[3 * i + 2] - com[2];
        g[0] += dx * dx + dy * dy + dz * dz;
    }
    block_reduce<1>(g, 0, sh);
    if (threadIdx.x != 0) return;
    double *o = st + ST_INST;
    o[0] = g[0] / nb;
    double re[3];
    for (int d = 0; d < 3; d++) re[d] = crd[3 * (nb - 1) + d] - crd[d];
    o[1] = re[0] * re[0] + re[1] * re[1] + re[2] * re[2];
    for (int d = 0; d < 3; d++) {
        o[2 + d] = mx[d] - mn[d];
        o[5 + d] = re[d];
    }
}

// compute_ic1 (stats_m.f90:367-405), one block per molecule: part[m] =
// (span(3), reedsq, rgsq, reev(3), counted).  mol: (natm, first bead,
// 0-based) per molecule.  Only the aligned-box form of get_image
// (simbox_m.f90:114-117) is needed: the device box is rectangular.
__global__ __launch_bounds__(256) void stats_ic1_kernel(const double *__restrict__ crd, const int *__restrict__ mol,
                                                        BoxArg bx, double *__restrict__ part)
{
    __shared__ double sh[3 * 256];
    const int m = blockIdx.x, natm = mol[2 * m], b0 = mol[2 * m + 1];
    double *o = part + 9L * m;
    if (natm <= 1) {
        if (threadIdx.x < 9) o[threadIdx.x] = 0.0;
        return;
    }
    // molbuf after the imaging loop: bead 1 at 0, bead i > 1 at img(r_i)
    auto rel = [&](int i, int d) { return i == 0 ? 0.0 : img(crd[3 * (b0 + i) + d], bx.L[d]); };
    double s[3] = {0, 0, 0}, mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = threadIdx.x; i < natm; i += blockDim.x)
        for (int d = 0; d < 3; d++) {
            const double x = rel(i, d);
            s[d] += x;
            mn[d] = fmin(mn[d], x);
            mx[d] = fmax(mx[d], x);
        }
    block_reduce<3>(s, 0, sh);
    block_reduce<3>(mn, 1, sh);
    block_reduce<3>(mx, 2, sh);
    const double com[3] = {s[0] / natm, s[1] / natm, s[2] / natm};
    double g[1] = {0.0};
    for (int i = threadIdx.x; i < natm; i += blockDim.x) {
        const double dx = rel(i, 0) - com[0], dy = rel(i, 1) - com[1], dz = rel(i, 2) - com[2];
        g[0] += dx * dx + dy * dy + dz * dz;
    }
    block_reduce<1>(g, 0, sh);
    if (threadIdx.x != 0) return;
    double re[3];
    for (int d = 0; d < 3; d++) {
        re[d] = rel(natm - 1, d);
        o[d] = mx[d] - mn[d];
        o[5 + d] = re[d];
    }
    o[3] = re[0] * re[0] + re[1] * re[1] + re[2] * re[2];
    o[4] = g[0] / natm;
    o[8] = 1.0;
}

// The running means of stats_accumulate (one thread).  e: the step's
// energies/stress/bond lengths (c.scal layout: [0] bond, [1] vdw, [2] total,
// [3..11] stress, [12..14] bond length mean/min/max, [24..27] angle,
// dihedral, tether, external energies); part/nmol: molecule partials of
// compute_ic1 (nmol < 0: imcon 0, the structure is in st[ST_INST..]).
__global__ void stats_accumulate_kernel(const double *__restrict__ e, const double *__restrict__ part, int nmol,
                                        int bnsiz, int bnsiz_stress, double *st, const DevStatus *status)
{
    if (status->ierr) return;   // bds_run returns before stats_accumulate
    double *in = st + ST_INST;
    if (nmol >= 0) {
        // compute_ic1: sums over the counted molecules, in molecule order
        // (none counted: 0/0, as the reference divides by count(natm > 1))
        double sp[3] = {0, 0, 0}, rq = 0, rg = 0, re[3] = {0, 0, 0};
        int cnt = 0;
        for (int m = 0; m < nmol; m++) {
            const double *o = part + 9L * m;
            if (o[8] == 0.0) continue;
            for (int d = 0; d < 3; d++) {
                sp[d] += o[d];
                re[d] = o[5 + d];
            }
            rq += o[3];
            rg += o[4];
            cnt++;
        }
        in[0] = rg / cnt;
        in[1] = rq / cnt;
        for (int d = 0; d < 3; d++) {
            in[2 + d] = sp[d] / cnt;
            in[5 + d] = re[d];
        }
    }
    int bnpop = (int)st[0], bnpop_s = (int)st[1];
    if (bnpop == bnsiz) {
        bnpop = 0;
        for (int k = 2; k <= 8; k++) st[k] = 0.0;
        for (int k = 18; k <= 22; k++) st[k] = 0.0;
    }
    bnpop++;
    const double en[7] = {e[0], e[24], e[25], e[1], e[26], e[27], e[2]};
    for (int k = 0; k < 7; k++) st[2 + k] = st[2 + k] + ((en[k] - st[2 + k]) / bnpop);
    st[18] = st[18] + ((in[0] - st[18]) / bnpop);
    st[19] = st[19] + ((in[1] - st[19]) / bnpop);
    for (int d = 0; d < 3; d++) st[20 + d] = st[20 + d] + ((in[2 + d] - st[20 + d]) / bnpop);
    if (bnpop_s == bnsiz_stress) {
        bnpop_s = 0;
        for (int k = 9; k <= 17; k++) st[k] = 0.0;
    }
    bnpop_s++;
    for (int k = 0; k < 9; k++) st[9 + k] = st[9 + k] + ((e[3 + k] - st[9 + k]) / bnpop_s);
    for (int d = 0; d < 3; d++) st[23 + d] = in[5 + d];
    for (int k = 0; k < 3; k++) st[26 + k] = e[12 + k];
    st[0] = bnpop;
    st[1] = bnpop_s;
}

int stats_step(Ctx &c)
{
    if (c.imcon == 0) {
        hipLaunchKernelGGL(stats_ic0_kernel, dim3(1), dim3(SB), 0, c.stream, c.crd, c.nb, c.stats);
    } else if (c.stats_nmol > 0) {
        BoxArg bx;
        bx.imcon = 1;
        for (int d = 0; d < 3; d++) bx.L[d] = c.box[d];
        hipLaunchKernelGGL(stats_ic1_kernel, dim3(c.stats_nmol), dim3(256), 0, c.stream, c.crd, c.stats_mol, bx,
                           c.stats_part);
    }
    hipLaunchKernelGGL(stats_accumulate_kernel, dim3(1), dim3(1), 0, c.stream, c.scal, c.stats_part,
                       c.imcon == 0 ? -1 : c.stats_nmol, c.stats_bnsiz, c.stats_bnsiz_stress, c.stats, c.st);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

void stats_free(Ctx &c)
{
    if (c.stats) (void)hipFree(c.stats);
    if (c.stats_mol) (void)hipFree(c.stats_mol);
    if (c.stats_part) (void)hipFree(c.stats_part);
    c.stats = nullptr;
    c.stats_mol = nullptr;
    c.stats_part = nullptr;
    c.stats_on = false;
}

}  // namespace bd

using namespace bd;

extern "C" int bd_stats_enable(bd_ctx *h, int binsize, int binsize_stress, int nmolecules, const int *molecules)
{
    if (!h || binsize < 0 || binsize_stress < 0 || nmolecules < 0 || (nmolecules > 0 && !molecules)) {
        set_error("bd_stats_enable: bad arguments");
        return BD_ERR_ARG;
    }
    Ctx &c = h->c;
    std::vector<int> mol(2 * (size_t)nmolecules);
    for (int m = 0; m < nmolecules; m++) {
        const int natm = molecules[2 * m], first = molecules[2 * m + 1] - 1;
        if (natm < 0 || first < 0 || (natm > 0 && first + natm > c.nb)) {
            set_error("bd_stats_enable: molecule outside the bead range");
            return BD_ERR_ARG;
        }
        mol[2 * m] = natm;
        mol[2 * m + 1] = first;
    }
    BD_HIP(hipStreamSynchronize(c.stream));
    stats_free(c);
    BD_HIP(hipMalloc(&c.stats, ST_WORDS * sizeof(double)));
    BD_HIP(hipMemset(c.stats, 0, ST_WORDS * sizeof(double)));
    if (nmolecules > 0) {
        BD_HIP(hipMalloc(&c.stats_mol, mol.size() * sizeof(int)));
        BD_HIP(hipMemcpy(c.stats_mol, mol.data(), mol.size() * sizeof(int), hipMemcpyHostToDevice));
        BD_HIP(hipMalloc(&c.stats_part, 9 * (size_t)nmolecules * sizeof(double)));
    }
    c.stats_nmol = nmolecules;
    c.stats_bnsiz = binsize;
    c.stats_bnsiz_stress = binsize_stress;
    c.stats_on = true;
    return BD_OK;
}

extern "C" int bd_stats_get(bd_ctx *h, double *out)
{
    if (!h || !out) { set_error("null argument"); return BD_ERR_ARG; }
    Ctx &c = h->c;
    if (!c.stats_on) { set_error("statistics not enabled (bd_stats_enable)"); return BD_ERR_STATE; }
    double st[ST_WORDS], com[3];
    BD_HIP(hipStreamSynchronize(c.stream));
    BD_HIP(hipMemcpy(st, c.stats, sizeof(st), hipMemcpyDeviceToHost));
    BD_HIP(hipMemcpy(com, c.scal + 16, sizeof(com), hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; k++) out[k] = st[26 + k];        // bndlen, bndlen_min, bndlen_max
    for (int k = 0; k < 7; k++) out[3 + k] = st[2 + k];     // energies
    for (int k = 0; k < 9; k++) out[10 + k] = st[9 + k];    // stress
    out[19] = st[18];                                       // rgsq
    out[20] = st[19];                                       // reedsq
    for (int k = 0; k < 3; k++) out[21 + k] = st[20 + k];   // span
    for (int k = 0; k < 3; k++) out[24 + k] = st[23 + k];   // reev
    for (int k = 0; k < 3; k++) out[27 + k] = com[k];       // molecule_com
    return BD_OK;
}
