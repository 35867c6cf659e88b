// bd_internal.h -- shared state and helpers of libbdstep (gfx950 / CDNA4).
//
// Data layout in HBM
//   crd, frc, drift      n = 3N FP64, xyz interleaved (the reference's (3,N)).
//   mobility D           lower-triangle tiles of TILE x TILE (TILE = 64):
//                        tile (I, J), I >= J, at index I(I+1)/2 + J, each tile
//                        column-major; diagonal tiles hold both triangles;
//                        rows/columns beyond n are zero.  FP64 or FP32.
//   Krylov basis V       npad x (s * dimk_max) FP64 column-major, npad =
//                        ntile * TILE, padding rows zero.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bdstep.h"

namespace bd {

constexpr int TILE = 64;
constexpr int MXPARAM = 12;
constexpr int MAX_S = 8;          // largest Lanczos block size supported
// Events that only time kernels: no system-scope fence when they complete
// (no cache writeback and invalidation around the timed launch).
constexpr unsigned kTimingEvent = hipEventDisableSystemFence;
// Events that hand work from one stream of the device to another: the
// consumer is a kernel on the same GPU, so a device-scope release suffices
// (a system-scope one writes the L2s back to memory, which the next kernel
// on the producing stream then waits for).
constexpr unsigned kStreamEvent = hipEventDisableTiming | hipEventDisableSystemFence;
constexpr int NQUAD_MAX = 160;    // quadrature nodes for H^{-1/2}
constexpr int NBOUND = 48;        // Cholesky probes for the lower spectral bound
constexpr int LZ_NEV = 8;         // events handing Lanczos iterations to the side stream

// 1/sqrt(v) in FP64 for normal v > 0: v_rsq_f64 (relative error ~6e-8) and
// one third-order correction,
//   1/sqrt(v) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + ...),  e = 1 - v y^2,
// truncated after e^2 (~1e-22), so within about one ulp after rounding: four
// dependent operations after the estimate, against a square root and a
// division of about thirty.
__device__ inline double rsqrt_f64(double v)
{
    const double y = __builtin_amdgcn_rsq(v);
    const double e = fma(-(v * y), y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}

void set_error(const std::string &msg);

#define BD_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess) {                                                   \
            ::bd::set_error(std::string("HIP error ") + hipGetErrorString(e_) +    \
                            " at " + __FILE__ + ":" + std::to_string(__LINE__));  \
            return BD_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define BD_TRY(expr)                   \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != BD_OK) return rc_;  \
    } while (0)

// Device-side status words (one struct in device memory per context).
struct DevStatus {
    int ierr;          // 1: bond overstretched; 2: factorisation failed; 4: cell-list coordinate < 0;
                       // 5: rigid tether violated; 6: external wall crossed
    int converged_j;   // Lanczos: first j with err <= tol (0 = not yet)
    int last_j;        // last Lanczos iteration executed
    int rebuild;       // Verlet list rebuild flag for the current step
    int npairs;        // half-list pair count of the current table
    int overflow;      // pair list capacity exceeded
    int ticket;        // work counter for the mobility product
    int nq;            // quadrature nodes used by the last small problem
    int steps_done;    // EM steps completed since the last bd_run started
    int nrebuild;      // Verlet list rebuilds
    int ycnt;          // blocks of the Lanczos combine kernel done (last one finalises)
    int fail_j;        // Lanczos: smallest Krylov block whose CholeskyQR failed (0: none)
    int vmove;         // Verlet check: some block saw a bead move by skin / 2 (cleared by the last block)
    int vdone;         // Verlet check: blocks done (the last one publishes rebuild)
    double err;        // last Lanczos relative change
    double lam_lo, lam_hi;  // spectral bounds of the last small problem
};

// Tile layout and work decomposition of the mobility of one or more chains
// (bd_mobility.hip).  chain_i: nb, n, ntile, nsc per chain; chain_l: row
// offset of the chain in the vectors, first tile, first row-partial block.
namespace sp {
struct SplitState;
}

struct MobPlan {
    int nchains = 0;
    int chr = 2, wsc = 4;     // work unit: chr tile rows x wsc tile columns
    long ntiles = 0;
    int nunits = 0, nrowblk = 0;
    int *tiles = nullptr;     // 4 ints per stored tile: chain, I, J, 0
    int *units = nullptr;     // 16 ints per work unit (the record of plan_build)
    int *rowblk = nullptr;    // 8 ints per output block
    int *chain_i = nullptr;
    long *chain_l = nullptr;
    double *rowpart = nullptr, *colpart = nullptr;
};
// Lanczos epilogue of a product (see symm_reduce_lz_kernel).
struct LzEpi {
    const double *Vj, *Vp, *Bp;
    double *part;
};
// Lanczos prologue of a product: the second CholeskyQR pass of the block
// the product multiplies.  x holds Q; the product uses V = Q R2^{-1} with
// R2^T R2 the sum of the ng Gram partials of Q, writes V (same layout as
// x) and B = R2 R1 (R1 the first pass's factor); a failed factorisation
// sets st->ierr = 2.  gpart == nullptr: x is used as it is.
struct XPre {
    int blk = 0;   // index j of the block V_j being completed (for fail_j)
    const double *gpart = nullptr;
    int ng = 0;
    const double *R1 = nullptr;
    double *Bout = nullptr;
    double *V = nullptr;
    DevStatus *st = nullptr;
};
extern const int kTileShapes[BD_TILE_SHAPES][2];
bool tile_shape_supported(int chr, int wsc);
void plan_pick(int nb, int cus, int &chr, int &wsc, std::vector<int> &widths);
int plan_build(MobPlan &p, const std::vector<int> &nb, const std::vector<long> &row_off, int chr = 2, int wsc = 4,
               const std::vector<std::vector<int>> *widths = nullptr, int slots = 0);
// Times the supported unit shapes on this plan (block width S) and rebuilds
// it with the fastest; times_us (BD_TILE_SHAPES entries) receives the times.
template <typename E, int S>
int plan_tune(MobPlan &p, const std::vector<int> &nb, const std::vector<long> &row_off, const E *mob,
              const double *x, double *y, int ld, int *ticket, hipStream_t sm, int device, double *times_us);
void plan_free(MobPlan &p);
template <typename E>
int rpy_launch(const MobPlan &p, const double *crd, E *mob, hipStream_t sm);
template <typename E, int S>
int plan_product(const MobPlan &p, const E *mob, const double *x, double *y, int ldx, int ldy, int *ticket,
                 const DevStatus *stop, const int *active, hipStream_t sm, int device, hipEvent_t e0, hipEvent_t e1,
                 const LzEpi *epi = nullptr, const XPre *pre = nullptr);

struct Ctx {
    int nb = 0, n = 0, npad = 0, ntile = 0, ntiles_lower = 0;
    bd_opts opt{};
    int s = 1;               // Brownian block width used with HI
    int dimk_max = 0;
    hipStream_t stream = nullptr;
    int device = 0;

    int imcon = 0;
    double box[3] = {0, 0, 0};

    // coordinates and per-step vectors
    double *crd = nullptr, *frc = nullptr, *drift = nullptr, *diff = nullptr;
    double *crd_mob = nullptr;   // coordinates D was assembled at (unused by EM)
    double *pc_r0 = nullptr, *pc_d0 = nullptr;   // predictor-corrector: r and d(r) of the step

    // mobility
    double *mob64 = nullptr;
    float *mob32 = nullptr;
    bool mob_valid = false;
    double *chol = nullptr;      // dense n x n work for the Cholesky factor
    int chol_ld = 0;

    // matrix-free products (bd_matfree.hip): no stored D; records of the
    // coordinates D stands for (mf_nbp beads, 4 doubles each, padded) and of
    // the block being multiplied (mf_rows x s row-major, FP32 or FP64), and
    // the per-workgroup partial sums (mf_jg x nb x 3 MAX_S FP64)
    bool mf = false;
    int mf_nbp = 0, mf_rows = 0, mf_jg = 1;
    double *mf_crd = nullptr;
    void *mf_x = nullptr;
    double *mf_part = nullptr;

    // symmetric product plan
    MobPlan plan;
    MobPlan plan_wide;   // 2 x 4 units for blocks wider than 4 columns (built on first use)
    int nunits = 0;

    // Lanczos workspace
    double *V = nullptr;         // npad x (s*dimk_max)
    double *W = nullptr;         // npad x s
    double *Y = nullptr;         // 2 x npad x s (ping-pong B.xi iterates)
    double *Ablk = nullptr;      // s x s x dimk_max  diagonal blocks
    double *Bblk = nullptr;      // s x s x dimk_max  sub-diagonal blocks
    double *R0 = nullptr;        // s x s
    double *gpart = nullptr;     // Gram partials (nblk x s x s)
    int gblocks = 0;
    double *Cq = nullptr;        // quadrature partials NQUAD_MAX x (js x s)
    double *Ccoef = nullptr;     // js x s coefficients
    int *bound_ok = nullptr;     // NBOUND flags
    sp::SplitState *split = nullptr;   // state of the split small problem
    double *lzpart = nullptr;          // per-block partials of the product epilogue
    double *Qbuf = nullptr;            // next Krylov block before the second CholeskyQR pass (npad x s)
    double *npart = nullptr;     // norm partials
    DevStatus *st = nullptr;
    DevStatus *st_host = nullptr;  // pinned mirror, 2 slots
    hipStream_t side = nullptr;    // small-problem stream of the Lanczos solve
    hipEvent_t lz_ev[LZ_NEV] = {};
    hipEvent_t lz_chunk_ev[2] = {};
    // grouped tests: {iteration tested through, converged_j, ierr, nodes} written by
    // the combine kernel into pinned host memory (lz_flag: host view,
    // lz_flag_dev: the device's address of the same words)
    volatile int *lz_flag = nullptr;
    int *lz_flag_dev = nullptr;
    // iteration at which the last Lanczos solve converged (0: none yet); the
    // next solve tests the iterations around it one by one
    int lz_hint = 0;
    double *Rq = nullptr, *Rcoef = nullptr;  // reorthogonalisation scratch

    // topology
    bool have_topology = false;
    int nbonds = 0, natom_types = 0, nvdw_types = 0;
    int lvdw = 0, excluded_atoms = 0, mth_ptgen = 0;
    double rcutoff = 0, tskin = 0;
    int *bond_ij = nullptr;        // nbonds x 2 (0-based)
    int *bond_type = nullptr;      // nbonds (0-based)
    int *bead_bond_ptr = nullptr;  // CSR bead -> bonds (bond index ascending)
    int *bead_bond_idx = nullptr;
    int nbond_types = 0;
    int *bond_styles = nullptr;
    double *bond_params = nullptr; // 12 x ntypes (derived slots set)
    int *atom_types = nullptr;     // 0-based
    int *vdw_styles = nullptr;
    double *vdw_params = nullptr;
    int *ex_ptr = nullptr, *ex_col = nullptr;  // exclusions, 0-based
    int ex_total = 0;
    std::vector<double> h_bpar, h_vpar;       // host copies (12 x ntypes)
    std::vector<int> h_bsty, h_vsty;
    bool vdw_general = false;      // a vdw style other than 0 and 1 is set
    // angles, tethers, external fields, charges (bd_set_angles, ...)
    int nangles = 0, ntethers = 0, nexternals = 0;
    int *ang_ptr = nullptr, *ang_idx = nullptr;   // CSR bead -> 4 * angle + role
    int *ang_atoms = nullptr, *ang_type = nullptr, *ang_styles = nullptr;
    double *ang_params = nullptr;
    int *teth_ptr = nullptr, *teth_idx = nullptr; // CSR bead -> tether
    int *teth_type = nullptr, *teth_styles = nullptr;
    double *teth_pts = nullptr, *teth_params = nullptr;
    int *ext_styles = nullptr;
    double *ext_params = nullptr;
    double *charge = nullptr;
    // an isolated molecule without tethers: bds_run keeps its centre of mass
    // at the origin (limud, bd_solver_m.f90:153-158)
    bool limud() const { return imcon == 0 && ntethers == 0; }

    // pair lists
    int pair_cap = 0;            // capacity of the full list
    int *nbr_cnt = nullptr;      // N
    int *nbr_ptr = nullptr;      // N+1
    int *nbr_idx = nullptr;      // pair_cap (full list, both directions)
    double *crd_save = nullptr;  // Verlet reference positions
    bool verlet_first = true;
    // bds_run's pre-loop centring (bd_solver_m.f90:161) is still due: set by
    // bd_set_coordinates and bd_reset_run, cleared by the first bd_run
    bool origin_pending = true;
    // semi-implicit step (bd_se.hip): Krylov basis and work vectors, reduction
    // scalars (device and pinned host), NITSOL counts of the last step and
    // the run (nfe, njve, nrpre, nli, nni, nbt)
    double *se_buf = nullptr, *se_sc = nullptr, *se_hsc = nullptr;
    int se_last[6] = {0, 0, 0, 0, 0, 0};
    long long se_total[6] = {0, 0, 0, 0, 0, 0};
    int *cell_cnt = nullptr, *cell_ptr = nullptr, *cell_beads = nullptr, *bead_cell = nullptr;
    int *cell_fill = nullptr;   // beads of each cell in arrival order, before ranking into cell_beads
    int ncell[3] = {0, 0, 0};
    double cl_cs[3] = {0, 0, 0};   // CL cell sides L / nc
    int ncell_tot = 0;
    int *scan_tmp = nullptr;
    int last_rebuilt = 0;
    bool pairs_valid = false;

    // reductions
    double *red = nullptr;       // generic partial buffer
    int red_cap = 0;
    double *scal = nullptr;      // small device scalars
    double *host_scal = nullptr; // pinned

    // run state
    int cntr_mobsam = 0;
    bd_runinfo info{};
    uint64_t rng_seed = 0x5eed, rng_offset = 0;
    double *xi_dev = nullptr;    // staging for injected xi (n x s)

    // run statistics on the device (bd_stats.hip)
    bool stats_on = false;
    double *stats = nullptr;       // accumulators and the step's structure
    int *stats_mol = nullptr;      // (natm, first bead) per molecule
    double *stats_part = nullptr;  // per-molecule structure of the step
    int stats_nmol = 0, stats_bnsiz = 0, stats_bnsiz_stress = 0;

    // kernel timing of the tile product (bd_prof_*)
    bool prof_on = false;
    int prof_width = 0;                // 0: every launch, else launches of this block width
    std::vector<hipEvent_t> prof_ev;   // pairs (start, stop)
    size_t prof_used = 0;
    double prof_ms = 0.0;
    long long prof_count = 0;
};
void prof_flush(Ctx &c);
// Events bracketing the next product launch of block width w when the
// tile-product timing is on (both null otherwise).
int prof_pair(Ctx &c, int w, hipEvent_t *e0, hipEvent_t *e1);

// --- mobility (bd_mobility.hip)
int mob_alloc(Ctx &c);
void mob_free(Ctx &c);
int rpy_assemble(Ctx &c);
// y(npad, s) = D x(npad, s) on the context stream; x, y device, ld = npad.
int mob_product(Ctx &c, const double *x, double *y, int s, int ldx, int ldy);
// Same product inside a Lanczos solve: skipped once the solve has converged.
int mob_product_lz(Ctx &c, const double *x, double *y, int s, int ldx, int ldy, const LzEpi *epi = nullptr,
                   const XPre *pre = nullptr);
// y = alpha * (D x + add) with optional add vector (s = 1); used for the drift.
int mob_unpack(Ctx &c, double *dense, int ld);

// Output blocks of 64 rows of a product (the Krylov epilogue's partial count).
inline int mob_rowblocks(const Ctx &c) { return c.mf ? c.ntile : c.plan.nrowblk; }

// --- matrix-free mobility (bd_matfree.hip)
bool mob_use_matfree(const bd_opts &o, int nbeads);
int mf_alloc(Ctx &c);
void mf_free(Ctx &c);
int mf_assemble(Ctx &c, const double *crd);
int mf_pack(Ctx &c, const double *x, int s, int ldx);
int mf_product(Ctx &c, const double *x, double *y, int s, int ldx, int ldy, const DevStatus *stop);
// Lanczos product of X (packed already when packed) with the epilogue of
// symm_reduce_lz_kernel written into W; pre: the second CholeskyQR pass.
int mf_product_lz(Ctx &c, const double *X, double *W, int s, int ld, bool packed, const LzEpi &epi,
                  const XPre *pre);
int mf_dense(Ctx &c, double *dense, int ld);

// --- Brownian (bd_lanczos.hip)
int lanczos_alloc(Ctx &c);
void lanczos_free(Ctx &c);
// xi (npad x s, device) -> out (npad x s, device); returns ierr-style code.
int brownian_lanczos(Ctx &c, const double *xi, double *out, int s, int *lconv, int *iters);
int brownian_cholesky(Ctx &c, const double *xi, double *out, int s, int method, bool sync = true);

// --- forces (bd_forces.hip)
int forces_setup(Ctx &c, int nbonds, const int *bonds, int nbond_types,
                 const int *bond_styles, const double *bond_params, int natom_types,
                 const int *atom_types, int nvdw_types, const int *vdw_styles,
                 const double *vdw_params);
void forces_free(Ctx &c);
int set_angles(Ctx &c, int nangles, const int *angles, int ntypes, const int *styles, const double *params);
int set_tethers(Ctx &c, int ntethers, const int *tethers, const double *points, int ntypes, const int *styles,
                const double *params);
int set_externals(Ctx &c, int nexternals, const int *styles, const double *params);
int set_charges(Ctx &c, const double *charge);
// Computes forces into c.frc at c.crd.  Energies/stress reductions are
// computed when want_energy is set (written to c.scal[0..14]).
int compute_forces(Ctx &c, bool want_energy, bool sync);
int grow_pairs(Ctx &c);
int adj_force(Ctx &c, const double *u, double *out, double *tpart, int tie_rc);
int set_theta(Ctx &c, double eps, double sigma, double kb, int tie_rc);
int pair_table_host(Ctx &c, std::vector<int> &pairs);

// --- semi-implicit step (bd_se.hip)
int se_alloc(Ctx &c);
void se_free(Ctx &c);
int se_step(Ctx &c, const double *w);
int launch_drift(Ctx &c, const double *v);

// --- statistics (bd_stats.hip)
int stats_step(Ctx &c);
void stats_free(Ctx &c);

// --- step pieces (bd_api.hip)
int em_step(Ctx &c, const double *xi_host, long long avail, long long *consumed, const double *disp_dev,
            double *rec_dev);
int check_status(Ctx &c, bool clear);
int rpy_assemble_at(Ctx &c, const double *crd);
int launch_to_origin(Ctx &c, double *com_acc);
int launch_gauss(Ctx &c, double *out, long long count, uint64_t seed, uint64_t offset, int ld, int rows, int cols,
                 double scale = 1.0);

}  // namespace bd

struct bd_ctx {
    bd::Ctx c;
};
