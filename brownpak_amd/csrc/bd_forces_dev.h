// bd_forces_dev.h -- scalar force laws shared by the single-system and the
// batched force kernels: minimum image, the reference's r^2 expression,
// ia_get_bond_force (ia_bond_m.f90:85-362) and vdw_lj (ia_vdw_m.f90:155-180).
#pragma once

#include "bd_internal.h"

namespace bd {

__device__ inline double img(double r, double L)
{
    // r - L*nint(r/L), nint = round half away from zero, no contraction
    return __dsub_rn(r, __dmul_rn(L, round(__ddiv_rn(r, L))));
}

__device__ inline double rsq3(double x, double y, double z)
{
    return __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

struct BoxArg {
    int imcon;
    double L[3];
};

__device__ inline void pair_vec(const double *crd, int i, int j, const BoxArg &bx, double &x, double &y, double &z)
{
    x = __dsub_rn(crd[3 * j], crd[3 * i]);
    y = __dsub_rn(crd[3 * j + 1], crd[3 * i + 1]);
    z = __dsub_rn(crd[3 * j + 2], crd[3 * i + 2]);
    if (bx.imcon) {
        x = img(x, bx.L[0]);
        y = img(y, bx.L[1]);
        z = img(z, bx.L[2]);
    }
}

// How the force kernels find the partners of a bead whose neighbour list
// overflowed: every bead j is scanned and taken as many times as the list
// would hold it.  DIR and VER: once when within rc2 at the list's reference
// coordinates.  CL (bead_cell set): once per cell of the bead's 27-cell
// stencil that holds j, with no distance test, as the cell-list table of
// link_cell_m.f90 has none; with two cells along an axis the neighbour cell
// is reached from both sides and j counts twice along it.
struct PairScan {
    const double *crd_ref;
    double rc2;
    const int *bead_cell;
    int nc[3];
};

__device__ inline int scan_mult(const PairScan &ps, const BoxArg &bx, int i, int j)
{
    if (ps.bead_cell) {
        int ci = ps.bead_cell[i], cj = ps.bead_cell[j], m = 1;
        for (int d = 0; d < 3; d++) {
            const int nc = ps.nc[d], a = ci % nc, b = cj % nc;
            ci /= nc;
            cj /= nc;
            int md = 0;
            for (int dd = -1; dd <= 1; dd++) md += ((a + dd + nc) % nc == b) ? 1 : 0;
            m *= md;
        }
        return m;
    }
    double x, y, z;
    pair_vec(ps.crd_ref, i, j, bx, x, y, z);
    return rsq3(x, y, z) < ps.rc2 ? 1 : 0;
}

// ia_get_bond_force (ia_bond_m.f90:85-362); returns 1 on overstretch.
__device__ inline int bond_force(double r, int sty, const double *p, double &enrg, double &frc)
{
    enrg = 0.0;
    frc = 0.0;
    switch (sty) {
    case 1:
        enrg = 0.5 * p[0] * (r - p[1]) * (r - p[1]);
        frc = p[0] * (r - p[1]);
        return 0;
    case 2: {
        double k = p[0], rmax = p[1], r0 = p[2], rmaxsq = p[3];
        double extn = r - r0, extnsq = extn * extn;
        if (r >= rmax + r0) return 1;
        enrg = -0.5 * k * rmaxsq * log(1.0 - extnsq / rmaxsq);
        frc = k * extn / (1.0 - extnsq / rmaxsq);
        return 0;
    }
    case 3: {
        double k = p[0], rmax = p[1], eps = p[2], sigma = p[3], rmaxsq = p[4], rcut = p[5];
        double rsq = r * r;
        if (r >= rmax) return 1;
        if (r >= rcut) {
            enrg = -0.5 * k * rmaxsq * log(1.0 - rsq / rmaxsq);
            frc = k * r / (1.0 - rsq / rmaxsq);
        } else {
            double sir = sigma / r, sir2 = sir * sir, sir6 = sir2 * sir2 * sir2, sir12 = sir6 * sir6;
            enrg = -0.5 * k * rmaxsq * log(1.0 - rsq / rmaxsq) + 4 * eps * (sir12 - sir6) + eps;
            frc = k * r / (1.0 - rsq / rmaxsq) - 24 * eps * (2 * sir12 - sir6) / r;
        }
        return 0;
    }
    case 4: {
        double lp = p[0], rmax = p[1];
        if (r >= rmax) return 1;
        double rt = r / rmax, rrt = 1.0 / (1.0 - rt);
        enrg = (-0.5 * rt * rt + 0.25 * rrt * rrt + 0.25 * rt) * (rmax / lp);
        frc = (-rt - 0.25 * rrt * rrt + 0.25) / lp;
        return 0;
    }
    default:
        return 0;
    }
}

// vdw_lj (ia_vdw_m.f90:155-180), truncated and force-shifted 12-6.
__device__ inline void lj_force(double r, const double *p, double &enrg, double &frc)
{
    enrg = 0.0;
    frc = 0.0;
    if (r < p[2]) {
        double sir = p[1] / r, sir2 = sir * sir, sir6 = sir2 * sir2 * sir2, sir12 = sir6 * sir6;
        enrg = 4 * p[0] * (sir12 - sir6) - p[3] - (r - p[2]) * p[4];
        frc = -24 * p[0] * (2 * sir12 - sir6) / r - p[4];
    }
}

// ia_get_vdw_force (ia_vdw_m.f90:51-81) for every style; qq = qi*qj.
// Styles: 1 LJ (force-shifted, 155-180), 2 Gaussian (210-229), 3 cosine
// (249-266), 4 LJ + screened Coulomb (310-352), 5 LJ + Coulomb (391-432),
// 6 DPD (454-470).  The derived slots are filled by vdw_derived.
__device__ inline void vdw_force(double r, double qq, int sty, const double *p, double &enrg, double &frc)
{
    enrg = 0.0;
    frc = 0.0;
    switch (sty) {
    case 1:
        lj_force(r, p, enrg, frc);
        return;
    case 2:
        if (r < p[2]) {
            const double ex = exp(-p[1] * r * r);
            enrg = p[0] * ex - p[3] - (r - p[2]) * p[4];
            frc = -2 * p[0] * p[1] * r * ex - p[4];
        }
        return;
    case 3: {
        const double pf = M_PI / p[1], pr = pf * r;
        if (r < p[1]) {
            enrg = p[0] * (1.0 + cos(pr));
            frc = -p[0] * pf * sin(pr);
        }
        return;
    }
    case 4:
    case 5: {
        // 4: eps sigma rcut rcut_coul C kappa | pot_rcut pot_rcut_coul (7, 8)
        // 5: eps sigma rcut rcut_coul C       | pot_rcut pot_rcut_coul (6, 7)
        const double rcc = p[3], kap = sty == 4 ? p[5] : 0.0;
        const double prc = sty == 4 ? p[6] : p[5], prcc = sty == 4 ? p[7] : p[6];
        const double ekr = sty == 4 ? p[4] * qq * exp(-kap * r) / r : p[4] * qq / r;
        const double fc = sty == 4 ? -ekr * (1 + kap * r) / r : -ekr / r;
        if (rcc > 0.0) {
            if (r < rcc) {
                enrg = ekr - qq * prcc;
                frc = fc;
            }
        } else {
            enrg = ekr;
            frc = fc;
        }
        if (r < p[2]) {
            const double sir = p[1] / r, sir2 = sir * sir, sir6 = sir2 * sir2 * sir2, sir12 = sir6 * sir6;
            enrg = enrg + 4 * p[0] * (sir12 - sir6) - prc;
            frc = frc - 24 * p[0] * (2 * sir12 - sir6) / r;
        }
        return;
    }
    case 6: {
        const double ir = 1.0 - r / p[1];
        if (r < p[1]) {
            enrg = 0.5 * p[0] * p[1] * ir * ir;
            frc = -p[0] * ir;
        }
        return;
    }
    default:
        return;
    }
}

// ia_get_angle_force (ia_angle_m.f90:46-79), style 1: k (1 - cos theta),
// theta the angle between q1 = r_i - r_(i-1) and q2 = r_(i+1) - r_i.
// Returns the forces on i-1 (f0) and i+1 (f2); the force on i is -(f0 + f2).
__device__ inline void angle_force(const double q1[3], const double q2[3], int sty, const double *p, double &enrg,
                                   double f0[3], double f2[3])
{
    enrg = 0.0;
    for (int d = 0; d < 3; d++) f0[d] = f2[d] = 0.0;
    if (sty != 1) return;
    const double m1 = sqrt(q1[0] * q1[0] + q1[1] * q1[1] + q1[2] * q1[2]);
    const double m2 = sqrt(q2[0] * q2[0] + q2[1] * q2[1] + q2[2] * q2[2]);
    double h1[3], h2[3];
    for (int d = 0; d < 3; d++) {
        h1[d] = q1[d] / m1;
        h2[d] = q2[d] / m2;
    }
    double ct = h1[0] * h2[0] + h1[1] * h2[1] + h1[2] * h2[2];
    ct = fmin(fmax(ct, -1.0), 1.0);
    const double k = p[0];
    enrg = k * (1.0 - ct);
    for (int d = 0; d < 3; d++) {
        f0[d] = k * (-h2[d] + ct * h1[d]) / m1;
        f2[d] = k * (h1[d] - ct * h2[d]) / m2;
    }
}

// ia_get_tether_force (ia_tether_m.f90:53-80): style 1 rigid (no force; an
// error once |r - r0| >= eps, 110-128), style 2 harmonic (146-158).
// Returns 1 on a rigid-tether violation.
__device__ inline int tether_force(double r, int sty, const double *p, double &enrg, double &frc)
{
    enrg = 0.0;
    frc = 0.0;
    if (sty == 1) return fabs(r - p[0]) < p[1] ? 0 : 1;
    if (sty == 2) {
        enrg = 0.5 * p[0] * (r - p[1]) * (r - p[1]);
        frc = p[0] * (r - p[1]);
    }
    return 0;
}

}  // namespace bd
