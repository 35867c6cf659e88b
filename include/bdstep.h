/*
 * bdstep.h -- C ABI of the MI355X Brownian-dynamics step (libbdstep.so).
 *
 * Replaces, one for one, the call sites of BROWNPAK's EM step:
 *
 *   calc_rpy_tensor(coordinates, mob)        src/bd_solver_m.f90:377, 458-522
 *   brn_init / brn_finish                    src/brown_m.f90:73-118, 122-143
 *   brn_calc_bdw(mob, bdw, ierr, lconv, f)   src/brown_m.f90:147-181
 *     calc_bdw_cholesky                      src/brown_m.f90:516-554
 *     calc_bdw_lanc                          src/brown_m.f90:185-288
 *     calc_bdw_blanc                         src/brown_m.f90:292-512
 *   ia_setup / ia_calc_forces                src/interaction_m.f90:30-71, 84-147
 *   pt_init / pt_build                       src/pairtab_m.fpp:36-98, 129-148
 *   dsymv('L', n, 1, mob, n, F, 1, 1, drift) src/bd_solver_m.f90:446
 *   integrate_em / bds_run loop body         src/bd_solver_m.f90:167-238
 *
 * Conventions (those of the reference):
 *   - coordinates and forces: (3, N) FP64, Fortran column-major, i.e. xyz of
 *     bead 1, then bead 2, ...;  n = 3N is the order of the mobility matrix;
 *   - the Brownian block bdw / xi: (n, s) FP64 column-major;
 *   - bonds: (3, nb) int32 [type, i, j] with 1-based bead indices;
 *   - bond/vdw parameters: (12, ntypes) FP64 column-major, user slots as in
 *     the .cfg file (the derived slots are filled here, as ia_*_setup does);
 *   - atom types, pair types and styles are 1-based, as in the .cfg file.
 *
 * Every call is synchronous with respect to the host buffers it takes.  A
 * context owns all device memory and one HIP stream; it is not thread-safe
 * (use one context per host thread).  No exception crosses the ABI.
 *
 * Return codes mirror the reference's ierr:
 *   BD_OK 0, BD_ERR_BOND 1 (ia_calc_forces' ierr: bond overstretched,
 *   ia_bond_m.f90:213-217; rigid tether violated, ia_tether_m.f90:110-128;
 *   external wall crossed, ia_external_m.f90:93-101),
 *   BD_ERR_FACTOR 2 (Cholesky/eigen failure, brown_m.f90:540-544),
 *   BD_WARN_NOTCONV 3 (Lanczos not converged; results valid, bd_solver_m.f90:
 *   393-397), BD_ERR_ARG -1, BD_ERR_HIP -2, BD_ERR_STATE -3, BD_ERR_CELL -4
 *   (negative coordinate in cell-list mode, cell_list_m.f90:193-196),
 *   BD_ERR_CAPACITY -5, BD_ERR_SOLVE 4 (the semi-implicit step's nonlinear
 *   solve failed: integrate_se's nitsol iterm 1-6, bd_solver_m.f90:277-299).
 *   bd_last_error() returns a message for the calling
 *   thread's last failure.
 */
#ifndef BDSTEP_H
#define BDSTEP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bd_ctx bd_ctx;

enum {
    BD_OK = 0,
    BD_ERR_BOND = 1,
    BD_ERR_FACTOR = 2,
    BD_WARN_NOTCONV = 3,
    BD_ERR_SOLVE = 4,
    BD_ERR_ARG = -1,
    BD_ERR_HIP = -2,
    BD_ERR_STATE = -3,
    BD_ERR_CELL = -4,
    BD_ERR_CAPACITY = -5
};

/* mob_fctr (control key mob_fctr, bd_solver_m.f90:64) */
enum {
    BD_CHOL = 0,    /* 'CHOL': B.xi = U.xi with D = U^T U (reference behaviour) */
    BD_LANC = 1,    /* 'LANC': Lanczos (s = 1) / block Lanczos (s > 1)          */
    BD_CHOL_LT = 2  /* Cholesky with U^T.xi, whose covariance is D               */
};

/* precision of the mobility storage and of the D.z / D.V products */
enum {
    BD_FP64 = 0,    /* FP64 storage, FP64 arithmetic                            */
    BD_FP32 = 1     /* FP32 storage, FP64 accumulation; Lanczos small problem,
                       QR and recurrences stay FP64                             */
};

/* integrator (control key bdintg, bd_solver_m.f90:64, 169-174).  BD_PC is the
 * predictor-corrector (Heun) step with the mobility and the Brownian
 * displacement w of the step held fixed:
 *   r* = r + d(r) + w,   r' = r + (d(r) + d(r*)) / 2 + w,
 * d(r) = dt (U(r) + D F(r)) -- one more force and D.F evaluation per step.
 * BD_SE is integrate_se (bd_solver_m.f90:242-332): r' solves
 *   r' - d(r') - r - w = 0
 * (D of the step held fixed) by NITSOL's inexact Newton-GMRES method as
 * integrate_se configures it: finite-difference J*v, GMRES restarted every
 * se_kdmax vectors with the residual re-evaluated at restarts, at most
 * se_nlmxitr Newton steps, stopping at |F| <= se_ftol or a step shorter
 * than se_stptol (control keys se_nlmxitr, se_kdmax, se_tol). */
enum { BD_EM = 0, BD_PC = 1, BD_SE = 2 };

/* mob_storage: how the mobility products D.x (dsymv/dsymm('L'),
 * bd_solver_m.f90:446, brown_m.f90:232, 386) get D.  BD_MOB_TILES stores the
 * lower triangle of calc_rpy_tensor's matrix (bd_solver_m.f90:458-522) in
 * 64 x 64 tiles and streams it once per product; BD_MOB_MATFREE stores only
 * the coordinates and evaluates every 3 x 3 RPY block inside the product
 * (same formula, no n x n storage); BD_MOB_AUTO takes the one measured to be
 * faster for the bead count and precision (for a batch: matrix-free). */
enum { BD_MOB_AUTO = 0, BD_MOB_TILES = 1, BD_MOB_MATFREE = 2 };

/* mth_ptgen (pairtab_m.fpp:23-24) */
enum { BD_PT_DIR = 0, BD_PT_VER = 1, BD_PT_CL = 2 };

typedef struct {
    int lhdia;            /* hydrodynamic interaction (RPY) on/off            */
    int mob_fctr;         /* BD_CHOL / BD_LANC / BD_CHOL_LT                    */
    int nts_mobsam;       /* s: mobility reuse interval == Lanczos block size */
    int lanc_mxitr;       /* Krylov iteration cap (dimk_max = max(2,.)+1)     */
    double lanc_tol;      /* relative-change tolerance                        */
    double tim_stp;       /* dt                                               */
    int precision;        /* BD_FP64 / BD_FP32                                 */
    int keep_reuse_noise; /* 0: reference behaviour, the diffusion block is
                             zeroed on every step (bd_solver_m.f90:364), so
                             steps 2..s of a block get no noise; 1: keep it    */
    int flow_style;       /* 0 none, 1 shear, 2 planar, 3 uniaxial (420-441)   */
    double flow_rate;     /* flow_params(1)                                    */
    int lanc_reorth;      /* 1: full FP64 reorthogonalisation of every new
                             Krylov block (not in the reference, whose Lanczos
                             loses orthogonality; see brown_m.f90:402-435)     */
    int integrator;       /* BD_EM (integrate_em, bd_solver_m.f90:216-238),
                             BD_SE (integrate_se, 242-332) or BD_PC:
                             predictor-corrector, not in the reference         */
    int mob_storage;      /* BD_MOB_AUTO / BD_MOB_TILES / BD_MOB_MATFREE      */
    int se_nlmxitr;       /* BD_SE: Newton steps (0: NITSOL's 200)            */
    int se_kdmax;         /* BD_SE: Krylov vectors per cycle (0: 20)          */
    double se_ftol;       /* BD_SE: se_tol(1), stop at |F| <= se_ftol          */
    double se_stptol;     /* BD_SE: se_tol(2), stop at |step| <= se_stptol     */
} bd_opts;

/* Context.  Replaces bds_init/brn_init (bd_solver_m.f90:50-104,
 * brown_m.f90:73-118): allocates D, the diffusion block and the Krylov
 * workspaces on the current HIP device.  Returns NULL on failure. */
bd_ctx *bd_ctx_create(int nbeads, const bd_opts *opts);
/* bds_finish/brn_finish/ia_finish (bd_solver_m.f90:108-124). */
void bd_ctx_destroy(bd_ctx *ctx);
const char *bd_last_error(void);

/* Simulation box: imcon 0 unbounded, 1 periodic orthorhombic (simbox_m). */
int bd_set_box(bd_ctx *ctx, int imcon, const double box[3]);

/* ia_setup + pt_init (interaction_m.f90:30-71, pairtab_m.fpp:36-98):
 * topology, exclusions (connectivity_m.f90:43-317) and potential
 * parameters.  Bond styles 1-4 (ia_bond_m.f90) and vdw styles 1-6
 * (ia_vdw_m.f90: LJ, Gaussian, cosine, LJ + screened Coulomb, LJ +
 * Coulomb, DPD); styles 4 and 5 take the charges of bd_set_charges. */
int bd_set_topology(bd_ctx *ctx, int nbonds, const int *bonds, int nbond_types,
                    const int *bond_styles, const double *bond_params,
                    int natom_types, const int *atom_types, int nvdw_types,
                    const int *vdw_styles, const double *vdw_params, int lvdw,
                    int excluded_atoms, int mth_ptgen, double rcutoff,
                    double tskin);

/* The remaining terms of ia_calc_forces (interaction_m.f90:84-147), set
 * after bd_set_topology; each call replaces the previous set (0 removes it).
 *   bd_set_angles: angles (4, nangles) [type, i-1, i, i+1] 1-based, styles
 *     0 (none) or 1 (k (1 - cos theta), ia_angle_m.f90:46-79), params
 *     (12, ntypes) with k = params(1) (ia_add_angle_forces, 276-325).
 *   bd_set_tethers: tethers (2, ntethers) [type, atom], tether_points
 *     (3, ntethers), styles 1 (rigid: an error once |r - r0| >= eps) or 2
 *     (harmonic k, r0), ia_tether_m.f90:53-158; not imaged (375-418).  A
 *     tethered system is not re-centred after each step (limud,
 *     bd_solver_m.f90:153-158).
 *   bd_set_externals: styles 1 (constant x force params(2) on atom
 *     params(1)) or 2 (wall: an error once component params(1) of any atom
 *     passes params(2) on the side of sign params(3)),
 *     ia_external_m.f90:61-107.
 *   bd_set_charges: charge (N) for vdw styles 4 and 5 (NULL: all 0).
 * The dihedral styles of the reference are all empty (ia_dihedral_m.f90),
 * so dihedrals add nothing and need no call.
 * bd_energies returns the energies of the last evaluation that computed
 * them (bd_forces, or a step with statistics on) in stats order: bond,
 * angle, dihedral, vdw, tether, external, total. */
int bd_set_angles(bd_ctx *ctx, int nangles, const int *angles, int nangle_types,
                  const int *angle_styles, const double *angle_params);
int bd_set_tethers(bd_ctx *ctx, int ntethers, const int *tethers,
                   const double *tether_points, int ntether_types,
                   const int *tether_styles, const double *tether_params);
int bd_set_externals(bd_ctx *ctx, int nexternals, const int *external_styles,
                     const double *external_params);
int bd_set_charges(bd_ctx *ctx, const double *charge);
int bd_energies(bd_ctx *ctx, double *energies7);

/* Host <-> device coordinates, (3, N) FP64. */
int bd_set_coordinates(bd_ctx *ctx, const double *crd);
int bd_get_coordinates(bd_ctx *ctx, double *crd);

/* calc_rpy_tensor (bd_solver_m.f90:458-522) on the device coordinates. */
int bd_rpy_assemble(bd_ctx *ctx);
/* Copies the assembled mobility to the host as the full n x n column-major
 * matrix (unit diagonal), widening FP32 storage to FP64. */
int bd_get_mobility(bd_ctx *ctx, double *mob);

/* dsymv/dsymm('L') products (bd_solver_m.f90:446, brown_m.f90:386):
 * y(n,s) := D x(n,s) with the last assembled D.  New coordinates do not
 * replace D until the next bd_rpy_assemble, as the reference keeps mob over
 * nts_mobsam steps. */
int bd_mob_symv(bd_ctx *ctx, const double *x, int s, double *y);

/* brn_calc_dw + brn_calc_bdw (brown_m.f90:147-181, 558-573) on the
 * assembled D, with the injected Gaussian block xi (n, s) in place of MKL
 * VSL.  method: BD_CHOL, BD_LANC, BD_CHOL_LT.  out receives B.xi (n, s).
 * lconv/iters as the reference's lconv/f.  With BD_CHOL the assembled D is
 * kept intact (the factor lives in its own buffer). */
int bd_brownian(bd_ctx *ctx, const double *xi, int s, int method, double *out,
                int *lconv, int *iters);

/* ia_calc_forces (interaction_m.f90:84-147) at the device coordinates,
 * building the pair table with the method chosen in bd_set_topology.
 * Any output pointer may be NULL.  energies = [bond, vdw, total];
 * stress9 (3x3 column-major, divided by the volume if periodic);
 * bndlen3 = [mean, min, max] bond length. */
int bd_forces(bd_ctx *ctx, double *frc, double *energies, double *stress9,
              double *bndlen3);

/* The pair table used by the last bd_forces call, as a half list of
 * unordered pairs (i < j, 1-based), sorted by (i, j).  *npairs receives
 * the count; returns BD_ERR_CAPACITY if cap is too small (count still
 * reported).  rebuilt (may be NULL) reports whether the last call rebuilt
 * the Verlet list (always 1 for DIR/CL). */
int bd_pair_table(bd_ctx *ctx, int *pairs2, int cap, int *npairs, int *rebuilt);

/* nsteps iterations of the bds_run loop body (bd_solver_m.f90:167-189):
 * integrate_em, nts+1, centre-of-mass removal for an isolated molecule or
 * PBC wrap.  xi is the injected Gaussian stream consumed exactly as
 * brn_calc_dw would (n*s values per mobility update with HI, n per step
 * without); *used receives the number of values consumed.  With xi ==
 * NULL the device Philox generator (seeded by bd_set_rng) supplies it.
 * The first call after bd_set_coordinates or bd_reset_run also moves an
 * isolated molecule's centre of mass to the origin before its first step
 * (bd_solver_m.f90:161); later calls continue the same loop, so a run
 * split into several calls takes the same steps as one call (nsteps 0
 * does only that centring). */
int bd_run(bd_ctx *ctx, int nsteps, const double *xi, long long nxi,
           long long *used);
/* Seeds the device Gaussian generator (Philox-4x32-10 + Box-Muller).  The
 * generator's whole state is (seed, offset): offset counts the Gaussian
 * values drawn (rounded up to even per draw), so bd_get_rng after a run and
 * bd_set_rng with the same pair in a new context continue the same stream.
 * Replaces init_stream / save_seed (random_m.f90:57-86, 118-129), whose
 * MT19937 stream is not saved with the revive file (save_stream,
 * random_m.f90:135-147, is never called). */
int bd_set_rng(bd_ctx *ctx, uint64_t seed, uint64_t offset);
int bd_get_rng(bd_ctx *ctx, uint64_t *seed, uint64_t *offset);

/* Run statistics: nts, molecule_com[3] (accumulated COM shift),
 * Lanczos iterations of the last update, total iterations, mobility
 * updates, non-converged updates. */
typedef struct {
    long long nts;
    double molecule_com[3];
    int last_lanczos_iters;
    int total_lanczos_iters;
    int n_mob_updates;
    int n_not_converged;
    int n_pair_rebuilds;
} bd_runinfo;
int bd_get_runinfo(bd_ctx *ctx, bd_runinfo *info);
/* Resets the step counters and the reuse counter cntr_mobsam (a fresh
 * bds_run), keeping coordinates and topology. */
int bd_reset_run(bd_ctx *ctx);

/* Rg^2 of the device coordinates (compute_ic0, stats_m.f90:441-446). */
int bd_rgsq(bd_ctx *ctx, double *rgsq);

/* Continues a run from a revive file (read_dump, config_io_m.f90:15-172):
 * sets atc%molecule_com, the accumulated centre-of-mass shift of an
 * isolated molecule, and marks the coordinates just set as mid-run, so the
 * next bd_run does not centre them again.  The reference's revived bds_run
 * does centre again (bd_solver_m.f90:161), which only moves coordinates
 * that a step left centred by rounding; skipping it lets a revived run,
 * with bd_set_rng restoring the generator, repeat the uninterrupted run
 * bit for bit. */
int bd_resume_run(bd_ctx *ctx, const double *com3);

/* Run statistics on the device (stats_init / stats_accumulate,
 * stats_m.f90:53-105, 299-361).  After bd_stats_enable every step of bd_run
 * (and bd_run_record) accumulates, on the device, the running means of
 * stats_accumulate over bins of binsize steps (energies, rgsq, reedsq, span)
 * and binsize_stress steps (stress), from the energies and stress of the
 * step's force evaluation and the structure of the end-of-step coordinates:
 * compute_ic0 (one molecule, imcon 0) or compute_ic1 (imcon 1, molecules
 * (2, nmolecules) = (atom count, first atom 1-based), i.e. rows 2 and 3 of
 * atc%molecules).  Enabling again restarts the accumulators (stats_init
 * after a revive).  bd_stats_get returns, in the order stats_write prints
 * them (stats_m.f90:235-290):
 *   out[0..2]   bndlen, bndlen_min, bndlen_max of the last step
 *   out[3..9]   energy accumulators: bond, angle, dihedral, vdw, tether,
 *               external, total
 *   out[10..18] stress accumulator, 3 x 3 column-major
 *   out[19..20] rgsq, reedsq accumulators
 *   out[21..23] span accumulator
 *   out[24..26] reev of the last step
 *   out[27..29] molecule_com */
int bd_stats_enable(bd_ctx *ctx, int binsize, int binsize_stress,
                    int nmolecules, const int *molecules);
int bd_stats_get(bd_ctx *ctx, double *out30);

/* NITSOL counts of the last semi-implicit step and of all of them: nfe,
 * njve, nrpre (always 0), nli, nni, nbt (nitsol.f's info(1:6)).  Either
 * pointer may be NULL. */
int bd_se_info(bd_ctx *ctx, int *last6, long long *total6);

/* Synchronises the context's stream (timing helper). */
int bd_sync(bd_ctx *ctx);

/* Differentiable trajectory (no reference counterpart; BROWNPAK has no
 * gradients).  The Brownian displacement of every step is recorded once and
 * then replayed ("noise held fixed"), so a trajectory is a deterministic
 * function of the force-field parameters theta = (eps, sigma) of vdw type 1
 * and k of bond type 1.
 *   bd_run_record: like bd_run, and writes the displacement each step
 *     applied, disp_out (n, nsteps) column-major (zero columns on the steps
 *     the reference leaves without noise).
 *   bd_run_replay: nsteps EM steps with disp (n, nsteps) as the Brownian
 *     term (mobility, forces and drift as in bd_run).
 *   bd_set_theta: sets eps, sigma (and rc = 2^(1/6) sigma if tie_rc) of
 *     vdw type 1 and k of bond type 1, re-deriving the shifted-force terms.
 *   bd_adjoint_rgsq: replays nsteps from the device coordinates, returns the
 *     final Rg^2 (stats_m.f90:441-446) in *loss and d loss / d(eps, sigma, k)
 *     in grad3 from a hand-written adjoint sweep with checkpoints every
 *     ckpt steps (a multiple of nts_mobsam); the coordinates end at the final
 *     state.  Isolated chains (imcon 0) only. */
int bd_set_theta(bd_ctx *ctx, double eps, double sigma, double kbond, int tie_rc);
int bd_run_record(bd_ctx *ctx, int nsteps, const double *xi, long long nxi,
                  long long *used, double *disp_out);
int bd_run_replay(bd_ctx *ctx, int nsteps, const double *disp);
int bd_adjoint_rgsq(bd_ctx *ctx, int nsteps, const double *disp, int ckpt,
                    int tie_rc, double *loss, double *grad3);

/* Measurement helpers (no reference counterpart; the reference's only timer
 * is the wall clock around bds_run, setup_m.f90:116-124).  With profiling
 * on, every launch of the mobility tile-product kernel is bracketed by HIP
 * events on the context's stream; bd_prof_get returns the summed kernel
 * time (ms), the launch count and the bytes of stored mobility tiles one
 * launch streams.  bd_prof_enable(ctx, 1) brackets every launch; on = 1 + w
 * (w = 1..8) only the launches whose block has w columns (w = nts_mobsam:
 * the Lanczos products, w = 1: the drift products); any on != 0 clears the
 * totals. */
int bd_prof_enable(bd_ctx *ctx, int on);
int bd_prof_get(bd_ctx *ctx, double *total_ms, long long *launches,
                double *tile_bytes);

/* Batched ensemble: many independent chains of uneven length advanced
 * together on one device, each exactly one bds_run (bd_solver_m.f90:128-238)
 * of a linear chain with HI (RPY, calc_rpy_tensor bd_solver_m.f90:458-522)
 * and (block) Lanczos noise (calc_bdw_lanc / calc_bdw_blanc, brown_m.f90:
 * 185-512), unbounded domain (imcon 0, centre of mass to the origin after
 * every step as smbx_to_origin, simbox_m.f90:182-202).  The reference runs
 * one chain per process; this is the same step for every chain, sharing the
 * launches.
 *   bd_batch_create: nbeads[c] beads of chain c (>= 2), bonds i-(i+1);
 *     opts: lhdia 1, mob_fctr BD_LANC, nts_mobsam 1..4, lanc_mxitr,
 *     lanc_tol, tim_stp, precision (tile storage), keep_reuse_noise.
 *   bd_batch_set_model: one bond type (style 1 harmonic or 2 FENE, the .cfg
 *     slots of ia_bond_m.f90) for every bond, LJ style 1 (ia_vdw_m.f90:
 *     107-180) when lvdw, pair cutoff rcutoff (r^2 < rc^2 as pt_build), and
 *     the 1..excluded_atoms ring exclusions of a linear chain (|i-j| <= k).
 *   bd_batch_set_coordinates / get: chains back to back, each (3, nb[c]).
 *   bd_batch_run: nsteps EM steps of every chain.  xi (optional) is the
 *     injected Gaussian stream: at every mobility update it supplies, chain
 *     by chain, an (n_c, s) column-major block (n_c = 3 nb[c]); *used returns
 *     the values consumed.  Without xi a device Philox generator is used,
 *     and every whole mobility block (nts_mobsam steps: assembly, noise,
 *     block Lanczos to lanc_mxitr with converged chains idle, the steps) is
 *     one launch of a HIP graph captured at the first such block; the
 *     result is bit-identical to stepping launch by launch (which the call
 *     does for injected xi, partial blocks and while profiling).  A Lanczos
 *     failure inside a graph block restores the block's start and replays
 *     it step by step, so errors and state are those of the step path.
 *     Errors: BD_ERR_BOND if any bond overstretched during the call.
 *   bd_batch_info: steps taken, mobility updates, summed Lanczos iterations
 *     over chains and updates, and chain-updates that hit lanc_mxitr.
 *   bd_batch_prof: as bd_prof_enable + bd_prof_get for the batch (on < 0
 *     reads without changing the state); tile_bytes is 0 for matrix-free
 *     products.
 *   Mobility products (opts->mob_storage): BD_MOB_MATFREE and BD_MOB_AUTO
 *     evaluate every chain's RPY blocks inside the product (one workgroup
 *     per chain and 64-bead row block, the chain's whole column range, so
 *     each chain's arithmetic is independent of the others); BD_MOB_TILES
 *     stores every chain's lower triangle in tiles, as below.  The tile
 *     shape calls return BD_ERR_STATE for a matrix-free batch.
 *   Tile shapes: the mobility product works in units of chr x wsc tiles of
 *     64 x 64; bd_batch_create times the BD_TILE_SHAPES supported shapes
 *     (in this order: 2x4, 1x4, 4x4, 2x2, 2x5, 2x6, 2x8) on the batch's own
 *     tile layout and keeps the fastest (the choice is remembered for the
 *     process, per chain-length list).  bd_batch_tile_shape returns the
 *     shape in use and the BD_TILE_SHAPES timings (us per product, 0 if not
 *     timed); bd_batch_set_tile_shape forces a shape.  Different shapes sum
 *     the partial products in different groupings, so results agree to
 *     rounding, not bit for bit.  (A single-chain context picks its shape
 *     from the chain length and the GPU's CU count, without timing.) */
#define BD_TILE_SHAPES 7
typedef struct bd_batch bd_batch;
bd_batch *bd_batch_create(int nchains, const int *nbeads, const bd_opts *opts);
void bd_batch_destroy(bd_batch *batch);
int bd_batch_set_model(bd_batch *batch, int bond_style, const double *bond_params,
                       int lvdw, const double *vdw_params, double rcutoff,
                       int excluded_atoms);
int bd_batch_set_coordinates(bd_batch *batch, const double *crd);
int bd_batch_get_coordinates(bd_batch *batch, double *crd);
int bd_batch_set_rng(bd_batch *batch, uint64_t seed, uint64_t offset);
int bd_batch_run(bd_batch *batch, int nsteps, const double *xi, long long nxi,
                 long long *used);
int bd_batch_info(bd_batch *batch, long long *nts, int *mob_updates,
                  int *total_iters, int *n_not_converged);
int bd_batch_tile_shape(bd_batch *batch, int *chr, int *wsc, double *times_us);
int bd_batch_set_tile_shape(bd_batch *batch, int chr, int wsc);
int bd_batch_prof(bd_batch *batch, int on, double *total_ms,
                  long long *launches, double *tile_bytes);

/* Matrix-free RPY product on device buffers, without a context:
 *   y(n, s) := D(crd) x(n, s),  n = 3 nbeads,
 * with D the RPY mobility of calc_rpy_tensor (bd_solver_m.f90:458-522,
 * unit diagonal) applied as the dsymm('L') of brown_m.f90:386 (or the
 * dsymv of bd_solver_m.f90:446 for s = 1) would apply the stored matrix,
 * but evaluated block by block inside the product (D is never formed).
 * crd (3, N), x (ldx, s) and y (ldy, s) are FP64 device pointers
 * (column-major, ldx and ldy >= n); s in 1..8; precision BD_FP64 or BD_FP32
 * (FP32 arithmetic in the product, FP64 sums); stream is a hipStream_t
 * (NULL: the default stream).  The work is ordered on stream; the scratch
 * (coordinate records, the packed block, partial sums) is allocated and
 * freed by the call, which therefore returns after the stream has drained. */
int bd_rpy_apply(int nbeads, const double *crd, const double *x, int ldx, int s, double *y, int ldy,
                 int precision, void *stream);

/* bd_rpy_apply without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and the same results, bit for bit; the scratch is a static
 * device array of the library (5.75 MiB on every GPU the library runs on,
 * held until the process ends), so nbeads is limited to 8192 (BD_ERR_ARG
 * beyond).  The call only enqueues work on stream and returns.  Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU (other
 * streams or threads) race and must be serialised by the caller. */
int bd_rpy_apply_noalloc(int nbeads, const double *crd, const double *x, int ldx, int s, double *y, int ldy,
                         int precision, void *stream);

/* Work decomposition of the mobility product (no counterpart in the
 * reference: introspection for tests and tuning; host only, no device call).
 * The single-chain plan a context of nbeads beads uses on a GPU with cus
 * CUs: the unit template chr x wsc (tile rows x tile columns), the unit
 * records in execution order (16 ints each: chain, super-column, first and
 * past-last tile row, tile rows of the chain, super-columns of the chain,
 * stored-tile mask (bit ii * wsc + jj), first tile column, tile base (2
 * ints), x row offset (2 ints), row-partial offset (2 ints), canonical
 * index, super-column width) and the output-block records (8 ints each:
 * chain, tile row K, super-columns, first and past-last canonical unit of
 * K's super-column, row-partial offset, y row offset, K's super-column |
 * K's offset in it << 16).  The arrays are filled when their capacity (in
 * ints) suffices; the counts always. */
int bd_plan_layout(int nbeads, int cus, int *chr, int *wsc, int *nunits, int *units, int units_cap,
                   int *nrowblk, int *rowblk, int rowblk_cap);

#ifdef __cplusplus
}
#endif
#endif /* BDSTEP_H */
