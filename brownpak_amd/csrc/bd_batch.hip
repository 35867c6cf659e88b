// bd_batch.hip -- many independent chains advanced together.
//
// Each chain is one BROWNPAK run (linear bead-spring chain, one bond type,
// one LJ/WCA type, RPY, Cholesky-free block Lanczos with s = nts_mobsam,
// unbounded domain): the same EM step as bds_run (bd_solver_m.f90:128-238),
// chain by chain, but every kernel covers all chains of the batch in one
// launch.  The mobility tiles of all chains share one MobPlan (work units of
// all chains feed one persistent product kernel); each Lanczos iteration runs
// the Krylov recurrence of every still-active chain with one block per chain,
// and the small problem of every chain with one block per chain (one thread
// per quadrature node, sp::small_problem).  Chains converge independently: a chain that passed the
// reference's test (brown_m.f90:497-506) is switched off for the rest of the
// solve, exactly as if it had left its own loop.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bd_forces_dev.h"
#include "bd_internal.h"
#include "bd_lanczos_dev.h"
#include "bd_matfree_dev.h"

namespace bd {

constexpr int BT = 256;            // threads per per-chain block
constexpr int BNBOUND = 32;        // spectral-bound probes per chain
constexpr int BDEC = 8;            // node table rows used (kappa <= 1e8)
constexpr int BNQP = 160;          // node slots per chain (2 * 79 + 1 <= 160)
constexpr int BJS = 512;           // largest j s of a chain's y combination (the order limit below)

struct BChainDev {
    int nb, n, npad, pad;
    long row;       // row offset of the chain in the padded vectors
};

struct Batch {
    int C = 0, NB = 0;
    long NP = 0;
    bd_opts opt{};
    int s = 1, dimk_max = 0;
    hipStream_t stream = nullptr, side = nullptr;
    int device = 0;
    std::vector<int> nb;
    std::vector<long> row;
    MobPlan plan;
    double tune_us[BD_TILE_SHAPES] = {};   // unit-shape timings of the autotuner
    float *mob32 = nullptr;
    double *mob64 = nullptr;
    // matrix-free products (no stored tiles): the bead records of every
    // chain (padded to MF_CH beads), each chain's first record, and the work
    // list of (chain, 64-bead row block), largest chains first
    bool mf = false;
    double *mf_rec = nullptr;
    int *mf_bo = nullptr;
    int2 *mf_items = nullptr;
    int mf_nitems = 0;
    // work list of the force kernel: (chain, BF_ROWS-bead block)
    int2 *bf_items = nullptr;
    int bf_nitems = 0;
    BChainDev *chains = nullptr;
    int *bead_chain = nullptr;    // per bead (global): chain
    int *bead_local = nullptr;    // per bead: index inside its chain
    long *bead_row = nullptr;     // per bead: row of its x component
    double *crd = nullptr, *frc = nullptr, *dfr = nullptr, *diff = nullptr, *xi = nullptr;
    double *V = nullptr, *W = nullptr, *Y = nullptr;
    double *A = nullptr, *B = nullptr, *R0 = nullptr;
    double *scr = nullptr, *Cc = nullptr;   // small-problem scratch and coefficients, per chain
    long scr_stride = 0;
    // active: the chain still takes Lanczos iterations (cleared by the
    // convergence test); run: its Krylov recurrence still advances (also
    // cleared on a breakdown); failed: iteration whose CholeskyQR broke down
    int *active = nullptr, *run = nullptr, *failed = nullptr, *conv = nullptr;
    int *kb_hint = nullptr;       // per chain: probe index of the small problem's last iteration
    int *ready = nullptr;         // per chain: last iteration whose small problem left C
    int *ycnt = nullptr;          // per chain: row blocks of the y combination finished (0 between launches)
    double *ypart = nullptr;      // per chain and row block: column sums of the convergence test
    int yrb = 1;                  // row blocks of the largest chain
    int *counters = nullptr;      // [0] active chains, [1] error code, [2] iterations, [3] not converged
    int *host_counters = nullptr; // pinned, 2 slots x 2
    DevStatus *st = nullptr;
    hipEvent_t ev[LZ_NEV] = {};
    hipEvent_t chunk_ev[2] = {};
    // model
    int bond_style = 2;
    double bpar[MXPARAM] = {}, vpar[MXPARAM] = {};
    double *dev_bpar = nullptr, *dev_vpar = nullptr;
    double rcutoff = 0.0;
    int excl = 0, lvdw = 1;
    bool have_model = false;
    int cntr = 0;
    long long nts = 0;
    int n_updates = 0, total_iters = 0, n_not_conv = 0;
    uint64_t seed = 0x5eed;
    uint64_t *dev_rng = nullptr;  // [0] generator stream position, [1] its copy at a block start
    std::vector<double> stage;    // host staging for injected xi
    // one mobility block (s steps) with device noise as a HIP graph
    hipGraphExec_t gexec = nullptr;
    hipEvent_t join_ev = nullptr;
    double *crd_snap = nullptr;   // coordinates at the start of the block being replayed
    // timing of the product
    bool prof_on = false;
    std::vector<hipEvent_t> prof_ev;
    size_t prof_used = 0;
    double prof_ms = 0.0;
    long long prof_count = 0;
};

// ------------------------------------------------------------ kernels
// Gaussian block for all chains: rows beyond a chain's n are zero.
__global__ void b_gauss_kernel(double *out, long NP, int s, const int *__restrict__ rowblk,
                               const BChainDev *__restrict__ ch, uint64_t seed, const uint64_t *__restrict__ offp)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NP * s) return;
    const uint64_t offset = *offp;   // the stream position lives on the device (graph replays)
    long rowg = g % NP;
    int c = rowblk[8 * (rowg / TILE)];
    long loc = rowg - ch[c].row;
    if (loc >= ch[c].n) {
        out[g] = 0.0;
        return;
    }
    uint64_t ctr = offset + (uint64_t)(g >> 1);
    uint32_t q[4] = {(uint32_t)ctr, (uint32_t)(ctr >> 32), 0x13198A2Eu, 0x03707344u};
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * q[0], p1 = (uint64_t)0xCD9E8D57u * q[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ q[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ q[3] ^ k1;
        q[0] = n0; q[1] = (uint32_t)p1; q[2] = n2; q[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    uint64_t a = ((uint64_t)q[0] << 32) | q[1], b = ((uint64_t)q[2] << 32) | q[3];
    double u1 = ((a >> 11) + 1) * 0x1.0p-53, u2 = (b >> 11) * 0x1.0p-53;
    double rr = sqrt(-2.0 * log(u1)), sn, cs;
    sincospi(2.0 * u2, &sn, &cs);
    out[g] = (g & 1) ? rr * sn : rr * cs;
}

// Advances the generator's stream position (one thread).
__global__ void b_rng_advance_kernel(uint64_t *offp, uint64_t inc) { *offp += inc; }

// Per-chain Lanczos iteration counts of a solve into counters[2] (sum) and
// counters[3] (chains that hit the cap), and counters[1] = 5 if a chain's
// Krylov basis broke down before it converged (the eager path's error).
__global__ void b_tally_kernel(const int *__restrict__ conv, const int *__restrict__ failed, int C, int dmax,
                               int *counters)
{
    __shared__ int s_it[256], s_nc[256];
    int it = 0, nc = 0, bad = 0;
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        it += conv[c] ? conv[c] : dmax;
        nc += conv[c] ? 0 : 1;
        bad |= (failed[c] && !conv[c]) ? 1 : 0;
    }
    s_it[threadIdx.x] = it;
    s_nc[threadIdx.x] = nc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0, b = 0;
        for (int t = 0; t < (int)blockDim.x; t++) { a += s_it[t]; b += s_nc[t]; }
        counters[2] += a;
        counters[3] += b;
    }
    if (bad) atomicCAS(&counters[1], 0, 5);
}

// Block-wide sum of an S x S matrix accumulated per thread (BT threads).
template <int S>
__device__ void bsum(double (&acc)[S * S], double *out)
{
    __shared__ double sh[BT / 64][S * S];
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
        for (int w = 0; w < BT / 64; w++) v += sh[w][threadIdx.x];
        out[threadIdx.x] = v;
    }
    __syncthreads();
}

// CholeskyQR2 of X (rows of one chain) into Q; R = R2 R1 -> Rout.
template <int S>
__device__ bool b_cholqr2(const double *X, double *Q, long ld, int rows, double *Rout)
{
    __shared__ double g[S * S];
    double R1[S][S], R2[S][S];
    for (int pass = 0; pass < 2; pass++) {
        const double *src = pass == 0 ? X : Q;
        double acc[S * S];
#pragma unroll
        for (int e = 0; e < S * S; e++) acc[e] = 0.0;
        for (int r = threadIdx.x; r < rows; r += BT) {
            double v[S];
#pragma unroll
            for (int k = 0; k < S; k++) v[k] = src[(long)k * ld + r];
#pragma unroll
            for (int cj = 0; cj < S; cj++)
#pragma unroll
                for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += v[ci] * v[cj];
        }
        bsum<S>(acc, g);
        double (&R)[S][S] = pass == 0 ? R1 : R2;
        if (!small_chol<S>(g, R)) return false;
        for (int r = threadIdx.x; r < rows; r += BT) {
            double x[S], q[S];
#pragma unroll
            for (int k = 0; k < S; k++) x[k] = src[(long)k * ld + r];
#pragma unroll
            for (int c = 0; c < S; c++) {
                double t = x[c];
#pragma unroll
                for (int k = 0; k < c; k++) t -= q[k] * R[k][c];
                q[c] = t / R[c][c];
            }
#pragma unroll
            for (int k = 0; k < S; k++) Q[(long)k * ld + r] = q[k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int i = 0; i < S; i++)
            for (int j = 0; j < S; j++) {
                double t = 0.0;
                for (int k = 0; k < S; k++) t += R2[i][k] * R1[k][j];
                Rout[i + j * S] = t;
            }
    return true;
}

// Initial block: xi = V_1 R0 for every chain.
template <int S>
__global__ __launch_bounds__(BT) void b_qr0_kernel(const double *xi, double *V, long NP, const BChainDev *ch,
                                                   double *R0, int *counters)
{
    const int c = blockIdx.x;
    const long r0 = ch[c].row;
    if (!b_cholqr2<S>(xi + r0, V + r0, NP, ch[c].npad, R0 + (long)c * S * S) && threadIdx.x == 0)
        atomicCAS(&counters[1], 0, 2);
}

// One Krylov step for every active chain (one block per chain):
//   W -= V_{j-1} B_{j-1}^T; A_j = V_j^T W; W -= V_j symL(A_j); W = V_{j+1} B_j.
template <int S>
__global__ __launch_bounds__(BT) void b_krylov_kernel(double *W, double *V, long NP, int dmax, int j,
                                                      const BChainDev *ch, double *A, double *B, int *run,
                                                      int *failed)
{
    const int c = blockIdx.x;
    if (!run[c]) return;
    const long r0 = ch[c].row;
    const int rows = ch[c].npad;
    const long ldv = NP;
    double *w = W + r0;
    const double *vj = V + (long)(j - 1) * S * ldv + r0;
    const double *vp = (j > 1) ? V + (long)(j - 2) * S * ldv + r0 : nullptr;
    const double *bp = (j > 1) ? B + ((long)c * dmax + (j - 2)) * S * S : nullptr;
    double *aj = A + ((long)c * dmax + (j - 1)) * S * S;
    __shared__ double sA[S * S];
    double acc[S * S];
#pragma unroll
    for (int e = 0; e < S * S; e++) acc[e] = 0.0;
    for (int r = threadIdx.x; r < rows; r += BT) {
        double x[S], v[S];
#pragma unroll
        for (int k = 0; k < S; k++) { x[k] = w[(long)k * NP + r]; v[k] = vj[(long)k * ldv + r]; }
        if (vp) {
#pragma unroll
            for (int cc = 0; cc < S; cc++) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < S; k++) t += vp[(long)k * ldv + r] * bp[cc + k * S];
                x[cc] = x[cc] - t;
            }
#pragma unroll
            for (int k = 0; k < S; k++) w[(long)k * NP + r] = x[k];
        }
#pragma unroll
        for (int cj = 0; cj < S; cj++)
#pragma unroll
            for (int ci = 0; ci < S; ci++) acc[ci + cj * S] += v[ci] * x[cj];
    }
    bsum<S>(acc, sA);
    if (threadIdx.x < S * S) aj[threadIdx.x] = sA[threadIdx.x];
    if (j >= dmax) return;
    for (int r = threadIdx.x; r < rows; r += BT) {
        double x[S], v[S];
#pragma unroll
        for (int k = 0; k < S; k++) { x[k] = w[(long)k * NP + r]; v[k] = vj[(long)k * ldv + r]; }
#pragma unroll
        for (int cc = 0; cc < S; cc++) {
            double t = 0.0;
#pragma unroll
            for (int k = 0; k < S; k++) t += v[k] * ((k >= cc) ? sA[k + cc * S] : sA[cc + k * S]);
            x[cc] = x[cc] - t;
        }
#pragma unroll
        for (int k = 0; k < S; k++) w[(long)k * NP + r] = x[k];
    }
    __syncthreads();
    double *vn = V + (long)j * S * ldv + r0;
    if (!b_cholqr2<S>(w, vn, NP, rows, B + ((long)c * dmax + (j - 1)) * S * S) && threadIdx.x == 0) {
        // breakdown: V_{j+1} does not exist; harmless if iteration j converges
        failed[c] = j;
        run[c] = 0;
    }
}

// H_j of chain c exists if no CholeskyQR before iteration j broke down
__device__ inline bool b_live(const int *active, const int *failed, int c, int j)
{
    return active[c] && !(failed[c] && failed[c] < j);
}

// Per chain (one block): the small problem (sp::small_problem: bounds,
// quadrature, C = sqrt(H_j) E1 R0 into Cbuf); ready[c] = j once C is there.
template <int S>
__global__ __launch_bounds__(BT) void b_small_kernel(const double *A, const double *B, const double *R0, int dmax,
                                                     int j, double *scr, long scr_stride, double *Cbuf,
                                                     int *active, const int *failed, int *kb_hint, int *ready,
                                                     int *counters)
{
    const int c = blockIdx.x;
    __shared__ int s_state;
    if (threadIdx.x == 0) s_state = !active[c] ? 0 : (b_live(active, failed, c, j) ? 1 : 2);
    __syncthreads();
    if (s_state == 0) return;
    if (s_state == 2) {
        // the Krylov basis ended before this chain converged: stop it here
        // and report it as failed at the end of the solve
        if (threadIdx.x == 0) {
            active[c] = 0;
            atomicSub(&counters[0], 1);
        }
        return;
    }
    const double *Ab = A + (long)c * dmax * S * S, *Bb = B + (long)c * dmax * S * S;
    double *Cc = Cbuf + (long)c * dmax * S * S;
    const sp::SmallOut o =
        sp::small_problem<S>(Ab, Bb, R0 + (long)c * S * S, j, BNBOUND, BDEC, scr + c * scr_stride, BNQP, Cc, kb_hint + c);
    if (threadIdx.x != 0) return;
    if (o.err) atomicCAS(&counters[1], 0, o.err);
    else ready[c] = j;
}

// y_j = V_{1:j} C of every chain whose small problem of iteration j is
// done (ready[c] == j), two adjacent rows per thread (16-byte loads; chain
// rows start on 64-row boundaries and npad is a multiple of 64; the basis is
// loaded non-temporally, as nothing else reads its older blocks) and YRR rows
// per workgroup (grid: chains x row blocks), the basis columns added in
// order; from j = 3 on, the convergence test: per column the sums over the
// chain's rows of |y_j - y_{j-1}| and |y_{j-1}| (squares for s = 1), by each
// workgroup in a fixed order, then by the chain's last workgroup to finish
// over the workgroups in row-block order; the chain stops when max_k sum|d| /
// max_k sum|y| (the 2-norm ratio for s = 1) is <= tol.
constexpr int YR = 256;          // threads per workgroup
constexpr int YRR = 2 * YR;      // rows per workgroup
template <int S>
__global__ __launch_bounds__(YR) void b_ycomb_kernel(const double *Cbuf, int dmax, int j, const double *V, long NP,
                                                     const BChainDev *ch, double *Y, const int *ready, int *active,
                                                     int *run, int *conv, double *ypart, int yrb, int *ycnt,
                                                     double tol, int *counters)
{
    const int c = blockIdx.x, rb = blockIdx.y, js = j * S;
    if (ready[c] != j) return;
    const int rows = ch[c].npad, nrb = (rows + YRR - 1) / YRR;
    if (rb >= nrb) return;
    __shared__ double sC[BJS * S];      // js S <= BJS S values
    const double *Cc = Cbuf + (long)c * dmax * S * S;
    for (int e = threadIdx.x; e < js * S; e += YR) sC[e] = Cc[e];
    __syncthreads();
    const long r0 = ch[c].row;
    const int r = rb * YRR + 2 * threadIdx.x;   // rows r and r + 1 (rows is even)
    double *y = Y + (long)(j % 2) * S * NP + r0;
    const double *yo = Y + (long)((j + 1) % 2) * S * NP + r0;
    const bool want = j > 2;
    double dn[S], on[S];
#pragma unroll
    for (int k = 0; k < S; k++) { dn[k] = 0.0; on[k] = 0.0; }
    if (r < rows) {
        double2 acc[S];
#pragma unroll
        for (int k = 0; k < S; k++) acc[k] = make_double2(0.0, 0.0);
        // sixteen basis columns in flight, added in column order
        for (int m0 = 0; m0 < js; m0 += 16) {
            double2 vm[16];
#pragma unroll
            for (int q = 0; q < 16; q++)
            {
                typedef double v2d __attribute__((ext_vector_type(2)));
                const v2d t = __builtin_nontemporal_load(
                    reinterpret_cast<const v2d *>(V + (long)min(m0 + q, js - 1) * NP + r0 + r));
                vm[q] = make_double2(t.x, t.y);
            }
#pragma unroll
            for (int q = 0; q < 16; q++)
                if (m0 + q < js) {
#pragma unroll
                    for (int k = 0; k < S; k++) {
                        const double cq = sC[(m0 + q) + k * js];
                        acc[k].x += vm[q].x * cq;
                        acc[k].y += vm[q].y * cq;
                    }
                }
        }
#pragma unroll
        for (int k = 0; k < S; k++) {
            *reinterpret_cast<double2 *>(y + (long)k * NP + r) = acc[k];
            if (want) {
                const double2 o = *reinterpret_cast<const double2 *>(yo + (long)k * NP + r);
                const double d0 = acc[k].x - o.x, d1 = acc[k].y - o.y;
                if (S == 1) { dn[k] = d0 * d0 + d1 * d1; on[k] = o.x * o.x + o.y * o.y; }
                else { dn[k] = fabs(d0) + fabs(d1); on[k] = fabs(o.x) + fabs(o.y); }
            }
        }
    }
    if (!want) return;
    __shared__ double sh[YR / 64][2 * S];
    __shared__ int s_last;
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < 2 * S; k++) {
        double v = k < S ? dn[k] : on[k - S];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sh[wv][k] = v;
    }
    __syncthreads();
    double *pc = ypart + (long)c * yrb * 2 * S;
    if (threadIdx.x == 0)
        for (int k = 0; k < 2 * S; k++) {
            double v = 0.0;
            for (int w = 0; w < YR / 64; w++) v += sh[w][k];
            store_dev(pc + rb * 2 * S + k, v);
        }
    stores_complete();
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(&ycnt[c], 1) == nrb - 1) ? 1 : 0;
    __syncthreads();
    if (!s_last || threadIdx.x != 0) return;
    ycnt[c] = 0;
    double tot[2 * S];
    for (int k = 0; k < 2 * S; k++) tot[k] = 0.0;
    for (int q = 0; q < nrb; q++)
        for (int k = 0; k < 2 * S; k++) tot[k] += load_dev(pc + q * 2 * S + k);
    double num = 0.0, den = 0.0;
    if (S == 1) {
        num = sqrt(tot[0]);
        den = sqrt(tot[1]);
    } else {
        for (int k = 0; k < S; k++) { num = fmax(num, tot[k]); den = fmax(den, tot[S + k]); }
    }
    if (num / den <= tol) {
        conv[c] = j;
        active[c] = 0;
        run[c] = 0;
        atomicSub(&counters[0], 1);
    }
}

// Forces of every bead: chain bonds (i-1, i+1) and the chain's LJ pairs
// r^2 < rc^2 with |i - j| > excl (the 1..3-ring exclusions of a linear
// chain), in that order, partners j in increasing order.  One workgroup per
// (chain, BF_ROWS-bead block) from items; the chain's coordinates are staged
// in LDS BF_TILE beads at a time, so every lane of a wave reads the same
// partner bead as an LDS broadcast; a chunk of BF_CH partners whose
// bounding box lies beyond the cutoff from every bead of the wave is passed
// over (it holds no pair in range, so the sums are the same).
constexpr int BF_ROWS = 256;
constexpr int BF_TILE = 1024;
constexpr int BF_CH = 16;        // candidate beads per bounding box
__global__ __launch_bounds__(BF_ROWS) void b_force_lds_kernel(const double *__restrict__ crd,
                                                              const int2 *__restrict__ items,
                                                              const BChainDev *__restrict__ ch, int bsty,
                                                              const double *bp, int lvdw, const double *vp,
                                                              double rc2, int excl, double *__restrict__ frc,
                                                              int *counters)
{
    __shared__ double sx[BF_TILE], sy[BF_TILE], sz[BF_TILE];
    __shared__ double sbox[BF_TILE / BF_CH][6];   // per chunk of BF_CH staged beads: low x, y, z, high x, y, z
    const int2 it = items[blockIdx.x];
    const int c = it.x, nbc = ch[c].nb;
    const int i = it.y * BF_ROWS + (int)threadIdx.x;
    const bool mine = i < nbc;
    const double *cr = crd + ch[c].row;
    const int ic = mine ? i : nbc - 1;
    const double xi = cr[3 * ic], yi = cr[3 * ic + 1], zi = cr[3 * ic + 2];
    double f[3] = {0, 0, 0};
    if (mine) {
        for (int side = -1; side <= 1; side += 2) {
            int j = i + side;
            if (j < 0 || j >= nbc) continue;
            double x = __dsub_rn(cr[3 * j], xi), y = __dsub_rn(cr[3 * j + 1], yi), z = __dsub_rn(cr[3 * j + 2], zi);
            double r = sqrt(rsq3(x, y, z)), en, fr;
            if (bond_force(r, bsty, bp, en, fr)) atomicCAS(&counters[1], 0, 1);
            f[0] += fr * x / r;
            f[1] += fr * y / r;
            f[2] += fr * z / r;
        }
    }
    if (lvdw) {
        for (int j0 = 0; j0 < nbc; j0 += BF_TILE) {
            const int nt = min(BF_TILE, nbc - j0);
            __syncthreads();
            for (int e = threadIdx.x; e < nt; e += BF_ROWS) {
                sx[e] = cr[3 * (j0 + e)];
                sy[e] = cr[3 * (j0 + e) + 1];
                sz[e] = cr[3 * (j0 + e) + 2];
            }
            __syncthreads();
            // the bounding box of each chunk of BF_CH staged beads
            for (int b = threadIdx.x; b * BF_CH < nt; b += BF_ROWS) {
                double lo[3] = {sx[b * BF_CH], sy[b * BF_CH], sz[b * BF_CH]}, hi[3] = {lo[0], lo[1], lo[2]};
                for (int q = b * BF_CH + 1; q < min(nt, (b + 1) * BF_CH); q++) {
                    lo[0] = fmin(lo[0], sx[q]); hi[0] = fmax(hi[0], sx[q]);
                    lo[1] = fmin(lo[1], sy[q]); hi[1] = fmax(hi[1], sy[q]);
                    lo[2] = fmin(lo[2], sz[q]); hi[2] = fmax(hi[2], sz[q]);
                }
                for (int a = 0; a < 3; a++) {
                    sbox[b][a] = lo[a];
                    sbox[b][3 + a] = hi[a];
                }
            }
            __syncthreads();
            if (!mine) continue;
            for (int b = 0; b * BF_CH < nt; b++) {
                // a chunk farther than the cutoff from every bead of the
                // wave holds no pair in range (the squared distance to its
                // box is a lower bound; the margin covers the rounding of
                // both sums), and adds nothing: skipping it keeps the sums
                const double d0 = fmax(0.0, fmax(sbox[b][0] - xi, xi - sbox[b][3]));
                const double d1 = fmax(0.0, fmax(sbox[b][1] - yi, yi - sbox[b][4]));
                const double d2 = fmax(0.0, fmax(sbox[b][2] - zi, zi - sbox[b][5]));
                if (__all(d0 * d0 + d1 * d1 + d2 * d2 > rc2 * (1.0 + 1e-6))) continue;
                for (int q = b * BF_CH; q < min(nt, (b + 1) * BF_CH); q++) {
                    const int j = j0 + q;
                    if (j == i || abs(j - i) <= excl) continue;
                    double x = __dsub_rn(sx[q], xi), y = __dsub_rn(sy[q], yi), z = __dsub_rn(sz[q], zi);
                    double r2 = rsq3(x, y, z);
                    if (!(r2 < rc2)) continue;
                    double r = sqrt(r2), en, fr;
                    lj_force(r, vp, en, fr);
                    f[0] += fr * x / r;
                    f[1] += fr * y / r;
                    f[2] += fr * z / r;
                }
            }
        }
    }
    if (mine) {
        double *o = frc + ch[c].row + 3L * i;
        o[0] = f[0];
        o[1] = f[1];
        o[2] = f[2];
    }
}

// r += dt * (D F) + diff (diff may be null); one thread per bead.  Nothing
// moves once an error is flagged (counters[1]): bds_run returns before the
// update of the failing step (bd_solver_m.f90:174).
__global__ void b_update_kernel(double *crd, const double *dfr, const double *diff, int NB, const long *bead_row,
                                double dt, const int *counters)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NB || counters[1] != 0) return;
    long o = bead_row[g];
    for (int d = 0; d < 3; d++) {
        double r = crd[o + d] + dt * dfr[o + d];
        if (diff) r = r + diff[o + d];
        crd[o + d] = r;
    }
}

// centre of mass to the origin, one block per chain (simbox_m.f90:182-202)
__global__ __launch_bounds__(256) void b_com_kernel(double *crd, const BChainDev *ch, const int *counters)
{
    if (counters[1] != 0) return;
    const int c = blockIdx.x;
    double *cr = crd + ch[c].row;
    const int nb = ch[c].nb;
    __shared__ double sh[3][4];
    __shared__ double m[3];
    double s[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) s[d] += cr[3 * i + d];
    for (int d = 0; d < 3; d++) {
        double v = s[d];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[d][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) m[threadIdx.x] = (sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3]) / nb;
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        for (int d = 0; d < 3; d++) cr[3 * i + d] = cr[3 * i + d] - m[d];
}

__global__ void b_scale_copy_kernel(double *dst, const double *src, long n, double a)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = a * src[i];
}

__global__ void b_lz_init_kernel(int *active, int *run, int *failed, int *conv, int *kb_hint, int *ready, int C,
                                 int *counters)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C) {
        active[c] = 1;
        run[c] = 1;
        failed[c] = 0;
        conv[c] = 0;
        kb_hint[c] = -1;
        ready[c] = 0;
    }
    if (c == 0) counters[0] = C;
}

// Records of every chain for the matrix-free products: (x, y, z, 0) per
// bead, padded to a multiple of MF_CH beads; one block per chain.
__global__ void b_mf_rec_kernel(const double *__restrict__ crd, const BChainDev *__restrict__ ch,
                                const int *__restrict__ bo, double *__restrict__ rec)
{
    const int c = blockIdx.x;
    const int nb = ch[c].nb, nbp = (nb + MF_CH - 1) / MF_CH * MF_CH;
    const double *cr = crd + ch[c].row;
    double *r = rec + 4L * bo[c];
    for (int j = threadIdx.x; j < nbp; j += blockDim.x) {
        double x, y, z;
        if (j < nb) {
            x = cr[3 * j];
            y = cr[3 * j + 1];
            z = cr[3 * j + 2];
        } else {
            mf_far_point(j, nb, x, y, z);
        }
        r[4 * j] = x;
        r[4 * j + 1] = y;
        r[4 * j + 2] = z;
        r[4 * j + 3] = 0.0;
    }
}

// Y = D X for every running chain with D never stored (bd_matfree_dev.h):
// one workgroup per (chain, 64-bead row block) over the chain's whole column
// range, so every row is complete in its workgroup (no partial sums).  X and
// Y are the padded vectors (ld NP); the chain's pad rows of Y are zero.
constexpr int BMF_WAVES = 4;   // waves per workgroup: less LDS, so that the product shares CUs with the tests
template <typename T, int S>
__global__ __launch_bounds__(64 * BMF_WAVES) void b_mf_product_kernel(const double *__restrict__ rec,
                                                                     const int *__restrict__ bo,
                                                                     const int2 *__restrict__ items,
                                                                     const BChainDev *__restrict__ ch,
                                                                     const double *__restrict__ X,
                                                                     double *__restrict__ Y, long NP,
                                                                     const int *__restrict__ run)
{
    using Rec = MfRec<T, S>;
    __shared__ __align__(16) unsigned char smem[mf_smem_bytes<T, S, BMF_WAVES>()];
    const int2 it = items[blockIdx.x];
    const int c = it.x, rb = it.y;
    if (run && !run[c]) return;
    const int nb = ch[c].nb, npad = ch[c].npad, nch = (nb + MF_CH - 1) / MF_CH;
    const double *crec = rec + 4L * bo[c];
    const double *x = X + ch[c].row;
    double *y = Y + ch[c].row;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ic = min(rb * 64 + lane, nb - 1);
    const double xi = crec[4 * ic], yi = crec[4 * ic + 1], zi = crec[4 * ic + 2];
    double acc[3][S];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int k = 0; k < S; k++) acc[a][k] = 0.0;
    auto stage = [&](int, int j, Rec &rc) {
        rc.x = crec[4 * j];
        rc.y = crec[4 * j + 1];
        rc.z = crec[4 * j + 2];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int k = 0; k < S; k++) rc.v[a][k] = j < nb ? (T)x[(long)k * NP + 3 * j + a] : (T)0;
    };
    mf_rows_times_cols<T, S, BMF_WAVES>(xi, yi, zi, 0, nch, reinterpret_cast<Rec *>(smem), w, stage, acc);
    mf_wave_sums<S, BMF_WAVES>(acc, reinterpret_cast<double(*)[64][3 * S + 1]>(smem), w, lane,
                               [&](int l, int col, double v) {
        const int ib = rb * 64 + l;
        if (ib < nb) y[(long)(col % S) * NP + 3 * ib + col / S] = v;
    });
    if (rb == (nb - 1) / 64) {
        const int np = npad - 3 * nb;
        for (int e = threadIdx.x; e < S * np; e += blockDim.x) y[(long)(e / np) * NP + 3 * nb + e % np] = 0.0;
    }
}

// results of non-converged chains: y of the last iteration
__global__ void b_finish_kernel(const double *Y, double *out, long NP, int s, int jlast, const int *conv,
                                const int *rowblk)
{
    long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= NP * s) return;
    long row = g % NP, k = g / NP;
    int c = rowblk[8 * (row / TILE)];
    int j = conv[c] ? conv[c] : jlast;
    out[g] = Y[(long)(j % 2) * s * NP + k * NP + row];
}

// ------------------------------------------------------------ host
// y = D x for every chain (active: only chains with active[c] != 0), from
// the stored tiles or matrix-free; HIP events e0/e1 around the launch.
template <int S>
static int b_product(Batch &b, const double *x, double *y, const int *active, hipEvent_t e0, hipEvent_t e1)
{
    const long NP = b.NP;
    hipStream_t sm = b.stream;
    if (!b.mf) {
        if (b.mob32)
            return plan_product<float, S>(b.plan, b.mob32, x, y, (int)NP, (int)NP, &b.st->ticket, nullptr, active, sm,
                                          b.device, e0, e1);
        return plan_product<double, S>(b.plan, b.mob64, x, y, (int)NP, (int)NP, &b.st->ticket, nullptr, active, sm,
                                       b.device, e0, e1);
    }
    if (e0) BD_HIP(hipEventRecord(e0, sm));
    if (b.opt.precision == BD_FP32)
        hipLaunchKernelGGL((b_mf_product_kernel<float, S>), dim3(b.mf_nitems), dim3(64 * BMF_WAVES), 0, sm, b.mf_rec,
                           b.mf_bo, b.mf_items, b.chains, x, y, NP, active);
    else
        hipLaunchKernelGGL((b_mf_product_kernel<double, S>), dim3(b.mf_nitems), dim3(64 * BMF_WAVES), 0, sm,
                           b.mf_rec, b.mf_bo, b.mf_items, b.chains, x, y, NP, active);
    BD_HIP(hipGetLastError());
    if (e1) BD_HIP(hipEventRecord(e1, sm));
    return BD_OK;
}

// The mobility at the current coordinates: the tiles, or the bead records
// of the matrix-free products.
static int b_mobility(Batch &b, hipStream_t sm)
{
    if (b.mf) {
        hipLaunchKernelGGL(b_mf_rec_kernel, dim3(b.C), dim3(256), 0, sm, b.crd, b.chains, b.mf_bo, b.mf_rec);
        BD_HIP(hipGetLastError());
        return BD_OK;
    }
    if (b.mob32) return rpy_launch<float>(b.plan, b.crd, b.mob32, sm);
    return rpy_launch<double>(b.plan, b.crd, b.mob64, sm);
}

// capture: the iterations run to dimk_max without looking at the host
// (kernels of converged chains are no-ops), nothing waits on the host, and
// the iteration counts are tallied on the device (counters[2], [3]).
template <int S>
static int b_lanczos(Batch &b, int *iters_out, bool capture = false)
{
    hipStream_t sm = b.stream, ss = b.side;
    const int C = b.C, dmax = b.dimk_max;
    const long NP = b.NP;
    hipLaunchKernelGGL(b_lz_init_kernel, dim3((C + 255) / 256), dim3(256), 0, sm, b.active, b.run, b.failed, b.conv,
                       b.kb_hint, b.ready, C, b.counters);
    hipLaunchKernelGGL((b_qr0_kernel<S>), dim3(C), dim3(BT), 0, sm, b.xi, b.V, NP, b.chains, b.R0, b.counters);
    BD_HIP(hipGetLastError());
    int pending = -1, slot = 0, jlast = 0;
    bool stop = false;
    for (int j = 1; j <= dmax && !stop; j++) {
        double *Vj = b.V + (size_t)(j - 1) * S * NP;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (b.prof_on) {
            if (b.prof_used + 2 > b.prof_ev.size()) {
                hipEvent_t x, y;
                BD_HIP(hipEventCreateWithFlags(&x, kTimingEvent));
                BD_HIP(hipEventCreateWithFlags(&y, kTimingEvent));
                b.prof_ev.push_back(x);
                b.prof_ev.push_back(y);
            }
            e0 = b.prof_ev[b.prof_used];
            e1 = b.prof_ev[b.prof_used + 1];
            b.prof_used += 2;
        }
        BD_TRY(b_product<S>(b, Vj, b.W, b.run, e0, e1));
        hipLaunchKernelGGL((b_krylov_kernel<S>), dim3(C), dim3(BT), 0, sm, b.W, b.V, NP, dmax, j, b.chains, b.A,
                           b.B, b.run, b.failed);
        BD_HIP(hipGetLastError());
        jlast = j;
        if (j >= 2) {
            hipEvent_t ev = b.ev[j % LZ_NEV];
            BD_HIP(hipEventRecord(ev, sm));
            BD_HIP(hipStreamWaitEvent(ss, ev, 0));
            hipLaunchKernelGGL((b_small_kernel<S>), dim3(C), dim3(BT), 0, ss, b.A, b.B, b.R0, dmax, j, b.scr,
                               b.scr_stride, b.Cc, b.active, b.failed, b.kb_hint, b.ready, b.counters);
            hipLaunchKernelGGL((b_ycomb_kernel<S>), dim3(C, b.yrb), dim3(YR), 0, ss, b.Cc, dmax, j, b.V, NP, b.chains,
                               b.Y, b.ready, b.active, b.run, b.conv, b.ypart, b.yrb, b.ycnt, b.opt.lanc_tol,
                               b.counters);
            BD_HIP(hipGetLastError());
        }
        if (!capture && (j % 4 == 0 || j == dmax)) {
            BD_HIP(hipMemcpyAsync(b.host_counters + 2 * slot, b.counters, 2 * sizeof(int), hipMemcpyDeviceToHost, ss));
            BD_HIP(hipEventRecord(b.chunk_ev[slot], ss));
            if (pending >= 0) {
                BD_HIP(hipEventSynchronize(b.chunk_ev[pending]));
                if (b.host_counters[2 * pending] == 0 || b.host_counters[2 * pending + 1]) stop = true;
            }
            pending = slot;
            slot ^= 1;
        }
    }
    if (capture) {
        // the side stream rejoins; results of every chain; tallies
        BD_HIP(hipEventRecord(b.join_ev, ss));
        BD_HIP(hipStreamWaitEvent(sm, b.join_ev, 0));
        hipLaunchKernelGGL(b_finish_kernel, dim3((unsigned)((NP * S + 255) / 256)), dim3(256), 0, sm, b.Y, b.diff, NP,
                           S, dmax, b.conv, b.plan.rowblk);
        hipLaunchKernelGGL(b_tally_kernel, dim3(1), dim3(256), 0, sm, b.conv, b.failed, C, dmax, b.counters);
        BD_HIP(hipGetLastError());
        return BD_OK;
    }
    BD_HIP(hipStreamSynchronize(ss));
    BD_HIP(hipStreamSynchronize(sm));
    int cnt[2];
    BD_HIP(hipMemcpy(cnt, b.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt[1] == 1) { set_error("bond length too large (overstretched bond)"); return BD_ERR_BOND; }
    if (cnt[1] == 3) { set_error("spectral range of a Lanczos matrix beyond 1e8"); return BD_ERR_FACTOR; }
    if (cnt[1]) { set_error("batched Lanczos failed (matrix not SPD?)"); return BD_ERR_FACTOR; }
    hipLaunchKernelGGL(b_finish_kernel, dim3((unsigned)((NP * S + 255) / 256)), dim3(256), 0, sm, b.Y, b.diff, NP, S,
                       jlast, b.conv, b.plan.rowblk);
    BD_HIP(hipGetLastError());
    std::vector<int> conv(C), failed(C);
    BD_HIP(hipMemcpy(conv.data(), b.conv, C * sizeof(int), hipMemcpyDeviceToHost));
    BD_HIP(hipMemcpy(failed.data(), b.failed, C * sizeof(int), hipMemcpyDeviceToHost));
    for (int c = 0; c < C; c++)
        if (failed[c] && !conv[c]) {
            set_error("Krylov breakdown before convergence in chain " + std::to_string(c));
            return BD_ERR_FACTOR;
        }
    for (int c = 0; c < C; c++) {
        int it = conv[c] ? conv[c] : dmax;
        if (iters_out) iters_out[c] = it;
        b.total_iters += it;
        if (!conv[c]) b.n_not_conv++;
    }
    return BD_OK;
}

template <typename E>
static int batch_tune_t(Batch &b, const E *mob)
{
    switch (b.s) {
    case 1: return plan_tune<E, 1>(b.plan, b.nb, b.row, mob, b.V, b.W, (int)b.NP, &b.st->ticket, b.stream, b.device, b.tune_us);
    case 2: return plan_tune<E, 2>(b.plan, b.nb, b.row, mob, b.V, b.W, (int)b.NP, &b.st->ticket, b.stream, b.device, b.tune_us);
    case 3: return plan_tune<E, 3>(b.plan, b.nb, b.row, mob, b.V, b.W, (int)b.NP, &b.st->ticket, b.stream, b.device, b.tune_us);
    case 4: return plan_tune<E, 4>(b.plan, b.nb, b.row, mob, b.V, b.W, (int)b.NP, &b.st->ticket, b.stream, b.device, b.tune_us);
    default: set_error("batched path supports nts_mobsam 1..4"); return BD_ERR_ARG;
    }
}

static int batch_tune(Batch &b)
{
    if (b.mob32) return batch_tune_t<float>(b, b.mob32);
    return batch_tune_t<double>(b, b.mob64);
}

// Step k (1..s) of a mobility block after the Brownian block is ready:
// forces, drift D.F, the EM update (noise column k - 1 if noise) and the
// centre-of-mass removal.
static int b_step_tail(Batch &b, int k, bool noise)
{
    const long NP = b.NP;
    hipStream_t sm = b.stream;
    const int nbk = (b.NB + 255) / 256;
    hipLaunchKernelGGL(b_force_lds_kernel, dim3(b.bf_nitems), dim3(BF_ROWS), 0, sm, b.crd, b.bf_items, b.chains,
                       b.bond_style, b.dev_bpar, b.lvdw, b.dev_vpar, b.rcutoff * b.rcutoff, b.excl, b.frc,
                       b.counters);
    BD_TRY(b_product<1>(b, b.frc, b.dfr, nullptr, nullptr, nullptr));
    const double *dcol = noise ? b.diff + (size_t)(k - 1) * NP : nullptr;
    hipLaunchKernelGGL(b_update_kernel, dim3(nbk), dim3(256), 0, sm, b.crd, b.dfr, dcol, b.NB, b.bead_row,
                       b.opt.tim_stp, b.counters);
    hipLaunchKernelGGL(b_com_kernel, dim3(b.C), dim3(256), 0, sm, b.crd, b.chains, b.counters);
    BD_HIP(hipGetLastError());
    return BD_OK;
}

static int b_step(Batch &b, const double *xi_host, long long avail, long long *consumed)
{
    const int s = b.s;
    const long NP = b.NP;
    hipStream_t sm = b.stream;
    *consumed = 0;
    bool noise = true;
    if (b.cntr == 0) {
        BD_TRY(b_mobility(b, sm));
        long long need = 0;
        for (int c = 0; c < b.C; c++) need += 3LL * b.nb[c] * s;
        if (xi_host) {
            if (avail < need) { set_error("injected xi stream exhausted"); return BD_ERR_ARG; }
            // chain by chain (n_c x s column-major) into the padded layout
            b.stage.assign((size_t)NP * s, 0.0);
            long long p = 0;
            for (int c = 0; c < b.C; c++) {
                const int n = 3 * b.nb[c];
                for (int k = 0; k < s; k++)
                    for (int r = 0; r < n; r++) b.stage[(size_t)k * NP + b.row[c] + r] = xi_host[p++];
            }
            BD_HIP(hipMemcpyAsync(b.xi, b.stage.data(), (size_t)NP * s * sizeof(double), hipMemcpyHostToDevice, sm));
            *consumed = need;
        } else {
            hipLaunchKernelGGL(b_gauss_kernel, dim3((unsigned)((NP * s + 255) / 256)), dim3(256), 0, sm, b.xi, NP, s,
                               b.plan.rowblk, b.chains, b.seed, b.dev_rng);
            hipLaunchKernelGGL(b_rng_advance_kernel, dim3(1), dim3(1), 0, sm, b.dev_rng, (uint64_t)NP * s);
            BD_HIP(hipGetLastError());
        }
        int rc;
        switch (s) {
        case 1: rc = b_lanczos<1>(b, nullptr); break;
        case 2: rc = b_lanczos<2>(b, nullptr); break;
        case 3: rc = b_lanczos<3>(b, nullptr); break;
        case 4: rc = b_lanczos<4>(b, nullptr); break;
        default: set_error("batched path supports nts_mobsam 1..4"); return BD_ERR_ARG;
        }
        if (rc != BD_OK) return rc;
        b.n_updates++;
        hipLaunchKernelGGL(b_scale_copy_kernel, dim3((unsigned)((NP * s + 255) / 256)), dim3(256), 0, sm, b.diff,
                           b.diff, NP * s, std::sqrt(2.0 * b.opt.tim_stp));
    } else if (!b.opt.keep_reuse_noise) {
        noise = false;  // bd_solver_m.f90:364
    }
    b.cntr++;
    BD_TRY(b_step_tail(b, b.cntr, noise));
    if (b.cntr == s) b.cntr = 0;
    // a bond error of this step ends the run here, as in bds_run
    BD_HIP(hipMemcpyAsync(b.host_counters + 1, b.counters + 1, sizeof(int), hipMemcpyDeviceToHost, sm));
    BD_HIP(hipStreamSynchronize(sm));
    if (b.host_counters[1] == 1) { set_error("bond length too large (overstretched bond)"); return BD_ERR_BOND; }
    if (b.host_counters[1] != 0) { set_error("batched Lanczos failed (matrix not SPD?)"); return BD_ERR_FACTOR; }
    b.nts++;
    return BD_OK;
}

// One mobility block of s steps with device noise (b.cntr == 0 on entry),
// device work only -- the sequence the HIP graph holds.
static int b_block_device(Batch &b)
{
    const int s = b.s;
    const long NP = b.NP;
    hipStream_t sm = b.stream;
    BD_TRY(b_mobility(b, sm));
    hipLaunchKernelGGL(b_gauss_kernel, dim3((unsigned)((NP * s + 255) / 256)), dim3(256), 0, sm, b.xi, NP, s,
                       b.plan.rowblk, b.chains, b.seed, b.dev_rng);
    hipLaunchKernelGGL(b_rng_advance_kernel, dim3(1), dim3(1), 0, sm, b.dev_rng, (uint64_t)NP * s);
    BD_HIP(hipGetLastError());
    int rc;
    switch (s) {
    case 1: rc = b_lanczos<1>(b, nullptr, true); break;
    case 2: rc = b_lanczos<2>(b, nullptr, true); break;
    case 3: rc = b_lanczos<3>(b, nullptr, true); break;
    case 4: rc = b_lanczos<4>(b, nullptr, true); break;
    default: set_error("batched path supports nts_mobsam 1..4"); return BD_ERR_ARG;
    }
    if (rc != BD_OK) return rc;
    hipLaunchKernelGGL(b_scale_copy_kernel, dim3((unsigned)((NP * s + 255) / 256)), dim3(256), 0, sm, b.diff, b.diff,
                       NP * s, std::sqrt(2.0 * b.opt.tim_stp));
    for (int k = 1; k <= s; k++) BD_TRY(b_step_tail(b, k, k == 1 || b.opt.keep_reuse_noise));
    return BD_OK;
}

// Captures b_block_device into b.gexec (once per model / plan / seed).
static int b_graph_build(Batch &b)
{
    if (b.gexec) return BD_OK;
    hipGraph_t g = nullptr;
    BD_HIP(hipStreamBeginCapture(b.stream, hipStreamCaptureModeRelaxed));
    int rc = b_block_device(b);
    hipError_t e = hipStreamEndCapture(b.stream, &g);
    if (rc != BD_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    BD_HIP(e);
    e = hipGraphInstantiate(&b.gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BD_HIP(e);
    return BD_OK;
}

static void b_graph_drop(Batch &b)
{
    if (b.gexec) (void)hipGraphExecDestroy(b.gexec);
    b.gexec = nullptr;
}

}  // namespace bd

using namespace bd;

struct bd_batch {
    Batch b;
};

extern "C" {

bd_batch *bd_batch_create(int nchains, const int *nbeads, const bd_opts *opts)
{
    if (nchains <= 0 || !nbeads || !opts) { set_error("bad arguments"); return nullptr; }
    if (!opts->lhdia || opts->mob_fctr != BD_LANC || opts->nts_mobsam < 1 || opts->nts_mobsam > 4) {
        set_error("the batched path runs RPY with (block) Lanczos, nts_mobsam 1..4");
        return nullptr;
    }
    if (opts->integrator != BD_EM) { set_error("the batched path integrates with EM"); return nullptr; }
    bd_batch *h = new bd_batch();
    Batch &b = h->b;
    auto fail = [&](const char *m) -> bd_batch * {
        set_error(m);
        bd_batch_destroy(h);
        return nullptr;
    };
    b.C = nchains;
    b.opt = *opts;
    b.s = opts->nts_mobsam;
    b.dimk_max = (opts->lanc_mxitr > 2 ? opts->lanc_mxitr : 2) + 1;
    if (b.dimk_max * b.s > 512) return fail("lanc_mxitr * nts_mobsam too large for the batched path (order <= 512)");
    b.nb.assign(nbeads, nbeads + nchains);
    long rowp = 0;
    std::vector<BChainDev> ch(nchains);
    std::vector<int> bc, bl;
    std::vector<long> br;
    for (int c = 0; c < nchains; c++) {
        if (nbeads[c] < 2) return fail("chains need at least 2 beads");
        const int n = 3 * nbeads[c], np = ((n + TILE - 1) / TILE) * TILE;
        b.row.push_back(rowp);
        ch[c] = BChainDev{nbeads[c], n, np, 0, rowp};
        for (int i = 0; i < nbeads[c]; i++) {
            bc.push_back(c);
            bl.push_back(i);
            br.push_back(rowp + 3L * i);
        }
        rowp += np;
    }
    b.NP = rowp;
    b.NB = (int)bc.size();
    if (hipGetDevice(&b.device) != hipSuccess) return fail("no HIP device");
    if (hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&b.side, hipStreamNonBlocking) != hipSuccess)
        return fail("stream creation failed");
    if (plan_build(b.plan, b.nb, b.row) != BD_OK) return fail("plan construction failed");
    if (opts->mob_storage != BD_MOB_AUTO && opts->mob_storage != BD_MOB_TILES && opts->mob_storage != BD_MOB_MATFREE)
        return fail("mob_storage must be BD_MOB_AUTO, BD_MOB_TILES or BD_MOB_MATFREE");
    // automatic: matrix-free, measured faster than the stored tiles on the
    // 512-chain ensemble in both precisions, and without their n^2 storage
    b.mf = opts->mob_storage != BD_MOB_TILES;
    const size_t elems = b.mf ? 0 : (size_t)b.plan.ntiles * TILE * TILE;
    bool ok = true;
    auto dal = [&](double **p, size_t count) {
        if (ok) ok = hipMalloc(p, std::max<size_t>(count, 1) * sizeof(double)) == hipSuccess &&
                     hipMemset(*p, 0, std::max<size_t>(count, 1) * sizeof(double)) == hipSuccess;
    };
    auto ial = [&](int **p, size_t count) {
        if (ok) ok = hipMalloc(p, std::max<size_t>(count, 1) * sizeof(int)) == hipSuccess &&
                     hipMemset(*p, 0, std::max<size_t>(count, 1) * sizeof(int)) == hipSuccess;
    };
    {
        std::vector<int2> fi;
        for (int c = 0; c < nchains; c++)
            for (int rb = 0; rb < (nbeads[c] + BF_ROWS - 1) / BF_ROWS; rb++) fi.push_back(make_int2(c, rb));
        b.bf_nitems = (int)fi.size();
        ok = hipMalloc(&b.bf_items, fi.size() * sizeof(int2)) == hipSuccess &&
             hipMemcpy(b.bf_items, fi.data(), fi.size() * sizeof(int2), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
    } else if (!b.mf) {
        if (opts->precision == BD_FP32) ok = hipMalloc(&b.mob32, elems * sizeof(float)) == hipSuccess;
        else ok = hipMalloc(&b.mob64, elems * sizeof(double)) == hipSuccess;
    } else {
        std::vector<int> bo(nchains), order(nchains);
        long nrec = 0;
        for (int c = 0; c < nchains; c++) {
            bo[c] = (int)nrec;
            nrec += (nbeads[c] + MF_CH - 1) / MF_CH * MF_CH;
            order[c] = c;
        }
        std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return nbeads[u] > nbeads[v]; });
        std::vector<int2> items;
        for (int c : order)
            for (int rb = 0; rb < (nbeads[c] + 63) / 64; rb++) items.push_back(make_int2(c, rb));
        b.mf_nitems = (int)items.size();
        ok = nrec < (1L << 30) && hipMalloc(&b.mf_rec, (size_t)nrec * 4 * sizeof(double)) == hipSuccess &&
             hipMalloc(&b.mf_bo, nchains * sizeof(int)) == hipSuccess &&
             hipMalloc(&b.mf_items, items.size() * sizeof(int2)) == hipSuccess &&
             hipMemcpy(b.mf_bo, bo.data(), nchains * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
             hipMemcpy(b.mf_items, items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice) == hipSuccess;
    }
    const int s = b.s, dm = b.dimk_max;
    dal(&b.crd, b.NP); dal(&b.frc, b.NP); dal(&b.dfr, b.NP); dal(&b.diff, (size_t)b.NP * s);
    dal(&b.xi, (size_t)b.NP * s); dal(&b.V, (size_t)b.NP * s * dm); dal(&b.W, (size_t)b.NP * s);
    dal(&b.Y, (size_t)2 * b.NP * s);
    dal(&b.A, (size_t)nchains * dm * s * s); dal(&b.B, (size_t)nchains * dm * s * s);
    b.scr_stride = sp::scratch_doubles(s, dm, BNQP);
    dal(&b.R0, (size_t)nchains * s * s); dal(&b.scr, (size_t)nchains * b.scr_stride);
    dal(&b.Cc, (size_t)nchains * dm * s * s); dal(&b.dev_bpar, MXPARAM); dal(&b.dev_vpar, MXPARAM);
    ial(&b.active, nchains); ial(&b.run, nchains); ial(&b.failed, nchains); ial(&b.conv, nchains);
    ial(&b.kb_hint, nchains); ial(&b.ready, nchains); ial(&b.ycnt, nchains);
    {
        int mx = 0;
        for (int c = 0; c < nchains; c++) mx = std::max(mx, ((3 * nbeads[c] + TILE - 1) / TILE) * TILE);
        b.yrb = (mx + YRR - 1) / YRR;
    }
    dal(&b.ypart, (size_t)nchains * b.yrb * 2 * s);
    ial(&b.counters, 4); ial(&b.bead_chain, b.NB); ial(&b.bead_local, b.NB);
    dal(&b.crd_snap, b.NP);
    if (ok) ok = hipMalloc(&b.dev_rng, 2 * sizeof(uint64_t)) == hipSuccess &&
                 hipMemset(b.dev_rng, 0, 2 * sizeof(uint64_t)) == hipSuccess;
    if (!ok) return fail("device allocation failed");
    if (hipMalloc(&b.bead_row, b.NB * sizeof(long)) != hipSuccess ||
        hipMalloc(&b.chains, nchains * sizeof(BChainDev)) != hipSuccess ||
        hipMalloc(&b.st, sizeof(DevStatus)) != hipSuccess || hipMemset(b.st, 0, sizeof(DevStatus)) != hipSuccess ||
        hipHostMalloc(&b.host_counters, 4 * sizeof(int)) != hipSuccess)
        return fail("device allocation failed");
    if (hipMemcpy(b.chains, ch.data(), nchains * sizeof(BChainDev), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.bead_chain, bc.data(), b.NB * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.bead_local, bl.data(), b.NB * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b.bead_row, br.data(), b.NB * sizeof(long), hipMemcpyHostToDevice) != hipSuccess)
        return fail("upload failed");
    for (int k = 0; k < LZ_NEV; k++)
        if (hipEventCreateWithFlags(&b.ev[k], kStreamEvent) != hipSuccess) return fail("event failed");
    for (int k = 0; k < 2; k++)
        if (hipEventCreateWithFlags(&b.chunk_ev[k], hipEventDisableTiming) != hipSuccess) return fail("event failed");
    if (hipEventCreateWithFlags(&b.join_ev, kStreamEvent) != hipSuccess) return fail("event failed");
    // tile unit shape chosen by timing the Lanczos product on this plan
    if (!b.mf) {
        if (b.mob32) ok = hipMemset(b.mob32, 0, elems * sizeof(float)) == hipSuccess;
        else ok = hipMemset(b.mob64, 0, elems * sizeof(double)) == hipSuccess;
        if (!ok || batch_tune(b) != BD_OK) return fail("tile shape autotuning failed");
    }
    return h;
}

void bd_batch_destroy(bd_batch *h)
{
    if (!h) return;
    Batch &b = h->b;
    if (b.stream) (void)hipStreamSynchronize(b.stream);
    if (b.side) (void)hipStreamSynchronize(b.side);
    b_graph_drop(b);
    if (b.join_ev) (void)hipEventDestroy(b.join_ev);
    if (b.crd_snap) (void)hipFree(b.crd_snap);
    if (b.dev_rng) (void)hipFree(b.dev_rng);
    plan_free(b.plan);
    double **dp[] = {&b.crd, &b.frc, &b.dfr, &b.diff, &b.xi, &b.V, &b.W, &b.Y, &b.A, &b.B, &b.R0, &b.scr,
                     &b.Cc, &b.ypart, &b.dev_bpar, &b.dev_vpar, &b.mob64};
    for (auto p : dp)
        if (*p) (void)hipFree(*p);
    int **ip[] = {&b.active, &b.run, &b.failed, &b.conv, &b.kb_hint, &b.ready, &b.ycnt, &b.counters, &b.bead_chain,
                  &b.bead_local};
    for (auto p : ip)
        if (*p) (void)hipFree(*p);
    if (b.mob32) (void)hipFree(b.mob32);
    if (b.mf_rec) (void)hipFree(b.mf_rec);
    if (b.mf_bo) (void)hipFree(b.mf_bo);
    if (b.mf_items) (void)hipFree(b.mf_items);
    if (b.bf_items) (void)hipFree(b.bf_items);
    if (b.bead_row) (void)hipFree(b.bead_row);
    if (b.chains) (void)hipFree(b.chains);
    if (b.st) (void)hipFree(b.st);
    if (b.host_counters) (void)hipHostFree(b.host_counters);
    for (auto e : b.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : b.chunk_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto e : b.prof_ev) (void)hipEventDestroy(e);
    if (b.side) (void)hipStreamDestroy(b.side);
    if (b.stream) (void)hipStreamDestroy(b.stream);
    delete h;
}

int bd_batch_set_model(bd_batch *h, int bond_style, const double *bond_params, int lvdw, const double *vdw_params,
                       double rcutoff, int excluded_atoms)
{
    if (!h || !bond_params || (lvdw && !vdw_params) || (bond_style != 1 && bond_style != 2) || excluded_atoms < 0 ||
        excluded_atoms > 3) {
        set_error("bad model arguments (bond style 1 or 2, vdw style 1)");
        return BD_ERR_ARG;
    }
    Batch &b = h->b;
    b.bond_style = bond_style;
    for (int k = 0; k < MXPARAM; k++) b.bpar[k] = bond_params[k];
    if (bond_style == 2) b.bpar[3] = b.bpar[1] * b.bpar[1];  // bond_fene_set
    b.lvdw = lvdw ? 1 : 0;
    if (lvdw) {
        for (int k = 0; k < MXPARAM; k++) b.vpar[k] = vdw_params[k];
        double sir = b.vpar[1] / b.vpar[2], s2 = sir * sir, s6 = s2 * s2 * s2, s12 = s6 * s6;  // vdw_lj_set
        b.vpar[3] = 4 * b.vpar[0] * (s12 - s6);
        b.vpar[4] = -24 * b.vpar[0] * (2 * s12 - s6) / b.vpar[2];
    }
    b.rcutoff = rcutoff;
    b.excl = excluded_atoms;
    BD_HIP(hipMemcpy(b.dev_bpar, b.bpar, sizeof(b.bpar), hipMemcpyHostToDevice));
    BD_HIP(hipMemcpy(b.dev_vpar, b.vpar, sizeof(b.vpar), hipMemcpyHostToDevice));
    b.have_model = true;
    b_graph_drop(b);   // model scalars are kernel arguments of the graph
    return BD_OK;
}

int bd_batch_set_coordinates(bd_batch *h, const double *crd)
{
    if (!h || !crd) { set_error("null argument"); return BD_ERR_ARG; }
    Batch &b = h->b;
    std::vector<double> st((size_t)b.NP, 0.0);
    long p = 0;
    for (int c = 0; c < b.C; c++)
        for (int i = 0; i < 3 * b.nb[c]; i++) st[b.row[c] + i] = crd[p++];
    BD_HIP(hipMemcpy(b.crd, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice));
    return BD_OK;
}

int bd_batch_get_coordinates(bd_batch *h, double *crd)
{
    if (!h || !crd) { set_error("null argument"); return BD_ERR_ARG; }
    Batch &b = h->b;
    std::vector<double> st((size_t)b.NP);
    BD_HIP(hipStreamSynchronize(b.stream));
    BD_HIP(hipMemcpy(st.data(), b.crd, st.size() * sizeof(double), hipMemcpyDeviceToHost));
    long p = 0;
    for (int c = 0; c < b.C; c++)
        for (int i = 0; i < 3 * b.nb[c]; i++) crd[p++] = st[b.row[c] + i];
    return BD_OK;
}

int bd_batch_set_rng(bd_batch *h, uint64_t seed, uint64_t offset)
{
    if (!h) { set_error("null batch"); return BD_ERR_ARG; }
    Batch &b = h->b;
    BD_HIP(hipStreamSynchronize(b.stream));
    b.seed = seed;
    BD_HIP(hipMemcpy(b.dev_rng, &offset, sizeof(uint64_t), hipMemcpyHostToDevice));
    b_graph_drop(b);   // the seed is a kernel argument of the graph
    return BD_OK;
}

int bd_batch_run(bd_batch *h, int nsteps, const double *xi, long long nxi, long long *used)
{
    if (!h || nsteps < 0) { set_error("bad arguments"); return BD_ERR_ARG; }
    Batch &b = h->b;
    if (!b.have_model) { set_error("model not set"); return BD_ERR_STATE; }
    long long pos = 0;
    // a fresh call: no error pending (a failed call leaves its code until here)
    BD_HIP(hipMemsetAsync(b.counters + 1, 0, sizeof(int), b.stream));
    hipLaunchKernelGGL(b_com_kernel, dim3(b.C), dim3(256), 0, b.stream, b.crd, b.chains, b.counters);
    int rc = BD_OK;
    // whole mobility blocks with device noise run as one graph launch each;
    // the rest (injected xi, partial blocks, profiling) step by step
    const bool graph = !xi && !b.prof_on;
    for (int k = 0; k < nsteps && rc == BD_OK;) {
        if (graph && b.cntr == 0 && nsteps - k >= b.s) {
            rc = b_graph_build(b);
            if (rc != BD_OK) break;
            BD_HIP(hipMemcpyAsync(b.crd_snap, b.crd, (size_t)b.NP * sizeof(double), hipMemcpyDeviceToDevice,
                                  b.stream));
            BD_HIP(hipMemcpyAsync(b.dev_rng + 1, b.dev_rng, sizeof(uint64_t), hipMemcpyDeviceToDevice, b.stream));
            BD_HIP(hipGraphLaunch(b.gexec, b.stream));
            int cnt[4];
            BD_HIP(hipMemcpyAsync(b.host_counters, b.counters, 4 * sizeof(int), hipMemcpyDeviceToHost, b.stream));
            BD_HIP(hipStreamSynchronize(b.stream));
            for (int q = 0; q < 4; q++) cnt[q] = b.host_counters[q];
            if (cnt[1] != 0) {
                // an overstretched bond or a Lanczos failure: back to the
                // block's start, replayed step by step so that the error, the
                // step count and the state are the eager path's
                const int zero[4] = {0, 0, 0, 0};
                BD_HIP(hipMemcpy(b.crd, b.crd_snap, (size_t)b.NP * sizeof(double), hipMemcpyDeviceToDevice));
                BD_HIP(hipMemcpy(b.dev_rng, b.dev_rng + 1, sizeof(uint64_t), hipMemcpyDeviceToDevice));
                BD_HIP(hipMemcpy(b.counters + 1, zero + 1, 3 * sizeof(int), hipMemcpyHostToDevice));
                for (int q = 0; q < b.s && rc == BD_OK; q++, k++) {
                    long long cons = 0;
                    rc = b_step(b, nullptr, 0, &cons);
                }
                break;
            }
            const int zero2[2] = {0, 0};
            BD_HIP(hipMemcpy(b.counters + 2, zero2, 2 * sizeof(int), hipMemcpyHostToDevice));
            b.total_iters += cnt[2];
            b.n_not_conv += cnt[3];
            b.n_updates++;
            b.nts += b.s;
            k += b.s;
            continue;
        }
        long long cons = 0;
        rc = b_step(b, xi ? xi + pos : nullptr, xi ? nxi - pos : 0, &cons);
        pos += cons;
        k++;
    }
    if (used) *used = pos;
    BD_HIP(hipStreamSynchronize(b.stream));
    int cnt[2];
    BD_HIP(hipMemcpy(cnt, b.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (rc == BD_OK && cnt[1] == 1) { set_error("bond length too large (overstretched bond)"); rc = BD_ERR_BOND; }
    return rc;
}

int bd_batch_tile_shape(bd_batch *h, int *chr, int *wsc, double *times_us)
{
    if (!h) { set_error("null batch"); return BD_ERR_ARG; }
    Batch &b = h->b;
    if (b.mf) { set_error("the batch runs matrix-free products: it has no tile shape"); return BD_ERR_STATE; }
    if (chr) *chr = b.plan.chr;
    if (wsc) *wsc = b.plan.wsc;
    if (times_us)
        for (int k = 0; k < BD_TILE_SHAPES; k++) times_us[k] = b.tune_us[k];
    return BD_OK;
}

int bd_batch_set_tile_shape(bd_batch *h, int chr, int wsc)
{
    if (!h) { set_error("null batch"); return BD_ERR_ARG; }
    if (!tile_shape_supported(chr, wsc)) { set_error("unsupported tile unit shape"); return BD_ERR_ARG; }
    Batch &b = h->b;
    if (b.mf) { set_error("the batch runs matrix-free products: it has no tile shape"); return BD_ERR_STATE; }
    BD_HIP(hipStreamSynchronize(b.stream));
    b_graph_drop(b);
    plan_free(b.plan);
    return plan_build(b.plan, b.nb, b.row, chr, wsc);
}

int bd_batch_info(bd_batch *h, long long *nts, int *mob_updates, int *total_iters, int *n_not_converged)
{
    if (!h) { set_error("null batch"); return BD_ERR_ARG; }
    Batch &b = h->b;
    if (nts) *nts = b.nts;
    if (mob_updates) *mob_updates = b.n_updates;
    if (total_iters) *total_iters = b.total_iters;
    if (n_not_converged) *n_not_converged = b.n_not_conv;
    return BD_OK;
}

int bd_batch_prof(bd_batch *h, int on, double *total_ms, long long *launches, double *tile_bytes)
{
    if (!h) { set_error("null batch"); return BD_ERR_ARG; }
    Batch &b = h->b;
    BD_HIP(hipStreamSynchronize(b.stream));
    for (size_t k = 0; k + 1 < b.prof_used; k += 2) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, b.prof_ev[k], b.prof_ev[k + 1]) == hipSuccess) {
            b.prof_ms += ms;
            b.prof_count++;
        }
    }
    b.prof_used = 0;
    if (total_ms) *total_ms = b.prof_ms;
    if (launches) *launches = b.prof_count;
    if (tile_bytes) *tile_bytes = b.mf ? 0.0 : (double)b.plan.ntiles * TILE * TILE * (b.mob32 ? 4.0 : 8.0);
    if (on >= 0) {
        b.prof_on = on != 0;
        b.prof_ms = 0.0;
        b.prof_count = 0;
    }
    return BD_OK;
}

}  // extern "C"
