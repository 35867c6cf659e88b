/*
 * bdoracle.h -- CPU FP64 restatement of BROWNPAK's Brownian-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path in
 * brownpak_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Conventions follow the reference exactly:
 *   - coordinates/forces are (3,N) column-major (xyz interleaved), FP64;
 *   - the mobility matrix is n x n column-major, n = 3N;
 *   - bonds are (3,nb) int32 [type, i, j], 1-based;
 *   - bond/vdw parameter blocks are (12, ntypes) column-major (mxparam = 12);
 *   - the Brownian block bdw is (n, s) column-major.
 * Every function cites the reference routine it restates (file:line).
 */
#ifndef BDORACLE_H
#define BDORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MXPARAM 12

/* ---- mobility -------------------------------------------------------- */
void orc_rpy(int nbeads, const double *crd, double *mob);
/* Host threads of orc_rpy and orc_symv_lower (1: the serial code the tests
 * use; more: OpenMP, for timing on all host cores). */
void orc_set_threads(int threads);
void orc_symv_lower(int n, double alpha, const double *a, const double *x,
                    double beta, double *y);

/* ---- Brownian displacement B.xi ------------------------------------- */
int orc_chol_bdw(int n, int s, double *mob, double *bdw);
int orc_lanczos(int n, const double *mob, double *bdw, int lanc_mxitr,
                double lanc_tol, int *lconv, int *iters);
int orc_block_lanczos(int n, int s, const double *mob, double *bdw,
                      int lanc_mxitr, double lanc_tol, int *lconv, int *iters);
int orc_brn_calc_bdw(int method, int n, int s, double *mob, double *bdw,
                     int lanc_mxitr, double lanc_tol, int *lconv, int *iters);
/* _ex variants: do_reorth = 1 adds full FP64 reorthogonalisation of each new
 * Krylov block (not in the reference). */
int orc_lanczos_ex(int n, const double *mob, double *bdw, int lanc_mxitr,
                   double lanc_tol, int do_reorth, int *lconv, int *iters);
int orc_block_lanczos_ex(int n, int s, const double *mob, double *bdw,
                         int lanc_mxitr, double lanc_tol, int do_reorth,
                         int *lconv, int *iters);
int orc_brn_calc_bdw_ex(int method, int n, int s, double *mob, double *bdw,
                        int lanc_mxitr, double lanc_tol, int do_reorth,
                        int *lconv, int *iters);

/* LAPACK-semantics helpers, exported so tests can pin them separately. */
void orc_geqp3(int m, int n, double *a, int *jpvt, double *tau);
void orc_orgqr(int m, int n, double *a, const double *tau);
int orc_syev(int n, double *a, double *w); /* a: n x n col-major in, eigvecs out */

/* ---- topology / forces ---------------------------------------------- */
typedef struct {
    int nbeads;
    int imcon;            /* 0 unbounded, 1 periodic orthorhombic */
    double box[3];        /* box edge lengths (diagonal basis)     */
    int nbonds;
    const int *bonds;     /* (3, nbonds) [type, i, j], 1-based      */
    int nbond_types;
    const int *bond_styles;
    double *bond_params;  /* (12, nbond_types); derived slots filled by orc_setup */
    int natom_types;
    const int *atom_types; /* (nbeads) 1-based atom type          */
    int nvdw_types;
    const int *vdw_styles;
    double *vdw_params;   /* (12, nvdw_types); derived slots filled by orc_setup */
    int lvdw;
    int excluded_atoms;   /* 0..3 rings excluded                    */
    int mth_ptgen;        /* 0 DIR, 1 VER, 2 CL                     */
    double rcutoff, tskin;
    /* the remaining ia_calc_forces terms (interaction_m.f90:84-147) */
    const double *charge; /* (nbeads) or NULL (all 0)               */
    int nangles;
    const int *angles;    /* (4, nangles) [type, i-1, i, i+1]        */
    int nangle_types;
    const int *angle_styles;
    const double *angle_params;   /* (12, nangle_types)             */
    int ntethers;
    const int *tethers;   /* (2, ntethers) [type, atom]              */
    const double *tether_points;  /* (3, ntethers)                  */
    int ntether_types;
    const int *tether_styles;
    const double *tether_params;  /* (12, ntether_types)            */
    int nexternals;
    const int *external_styles;
    const double *external_params; /* (12, nexternals)              */
} orc_system;

typedef struct {
    double energy_bond, energy_vdw, energy_tot;
    double stress[9];     /* 3x3 column-major */
    double bndlen, bndlen_min, bndlen_max;
    double energy_angle, energy_dihedral, energy_tether, energy_external;
} orc_energies;

void orc_setup_params(orc_system *sys);

/* Exclusion table (connectivity_m.f90:43-317): CSR over beads, 1-based
 * values, rows sorted ascending, self included when excluded_atoms > 0. */
int orc_exclusions(int nbeads, int nbonds, const int *bonds, int excluded_atoms,
                   int *row_ptr, int *col, int cap);

/* Pair table builders (pairtab_m.fpp).  Output: half list as CSR over rows
 * (row_ptr has nbeads+1 entries, 0-based offsets; col holds 1-based bead
 * indices).  Returns the number of pairs, or -1 if cap is exceeded.
 * For VER, save/dr buffers hold the Verlet state; *rebuilt reports whether
 * the list was rebuilt on this call (first_call -> always). */
typedef struct {
    int first_call;
    double *crd_save;     /* (3, nbeads) */
    double rskin_sq, tskin_sq;
} orc_verlet;

int orc_pairs_n2(const orc_system *sys, const int *ex_ptr, const int *ex_col,
                 const double *crd, double rc_sq, int *row_ptr, int *col, int cap);
int orc_pairs_verlet(const orc_system *sys, const int *ex_ptr, const int *ex_col,
                     const double *crd, orc_verlet *vs, int *row_ptr, int *col,
                     int cap, int *rebuilt);
int orc_pairs_cell_list(const orc_system *sys, const int *ex_ptr,
                        const int *ex_col, const double *crd, int *row_ptr,
                        int *col, int cap);

/* ia_calc_forces (interaction_m.f90:84-147) over a given pair table. */
int orc_forces(const orc_system *sys, const double *crd, const int *pt_ptr,
               const int *pt_col, double *frc, orc_energies *en);

/* ---- simulation box helpers (simbox_m.f90:106-202) ------------------ */
void orc_get_image(const double *box, double *r);
void orc_wrap_all(const double *box, int nbeads, double *crd);
void orc_to_origin(int nbeads, double *crd, double *com);
double orc_rgsq(int nbeads, const double *crd);

/* ---- the BD time loop (bd_solver_m.f90:128-238, 355-454) ------------ */
typedef struct {
    int lhdia;            /* hydrodynamic interaction on/off            */
    int mob_fctr;         /* 0 CHOL (reference U.xi), 1 LANC, 2 CHOL_LT (U^T.xi) */
    int nts_mobsam;       /* s                                          */
    int lanc_mxitr;
    double lanc_tol;
    double tim_stp;
    int keep_reuse_noise; /* 0: reference behaviour (columns 2..s zeroed), 1: keep them */
    int flow_style;       /* 0 none, 1 shear, 2 planar, 3 uniaxial      */
    double flow_rate;
    int lanc_reorth;      /* 1: FP64 full reorthogonalisation (not in the reference) */
    int integrator;       /* 0 EM (integrate_em); 1 predictor-corrector (Heun drift,
                             mobility and Brownian term held over the step; not in
                             the reference); 2 SE (integrate_se, NITSOL) */
    int se_nlmxitr, se_kdmax;     /* control keys se_nlmxitr, se_kdmax        */
    double se_ftol, se_stptol;    /* control key se_tol(1:2)                  */
} orc_bdopts;

typedef struct {
    long long nts;
    double molecule_com[3];
    int last_lanczos_iters;
    int n_not_converged;
    int ierr;
    int total_lanczos_iters;
    int n_mob_updates;
    /* semi-implicit steps: NITSOL totals and the last termination (iterm) */
    int se_nfe, se_nli, se_nni, se_nbt, se_iterm;
} orc_runinfo;

/* Runs nsteps EM steps in place on crd.  xi is the injected Gaussian stream:
 * each mobility update consumes n*s values (HI), each step n values (no HI).
 * Returns the number of xi values consumed, or a negative error. */
long long orc_bd_run(const orc_system *sys, const orc_bdopts *opt, double *crd,
                     int nsteps, const double *xi, long long nxi,
                     orc_runinfo *info);
/* NITSOL v0.3 (src/nitsol) restated in bdnitsol.c: GMRES, finite-difference
 * J*v, backtracking.  fn returns 0 or nonzero on failure.  input/info as
 * nitsol's; returns iterm. */
typedef int (*orc_fn)(void *ctx, int n, const double *x, double *f);
int orc_nitsol(int n, double *x, orc_fn fn, void *ctx, double ftol, double stptol, const int *input, int *info);
/* LAPACK dlaic1 restated from its definition (see bdnitsol.c). */
void orc_dlaic1(int job, int j, const double *x, double sest, const double *w, double gamma, double *sestpr,
                double *s, double *c);

long long orc_bd_run_ex(const orc_system *sys, const orc_bdopts *opt, double *crd,
                        int nsteps, const double *xi, long long nxi,
                        const double *disp, orc_runinfo *info);

#ifdef __cplusplus
}
#endif
#endif
