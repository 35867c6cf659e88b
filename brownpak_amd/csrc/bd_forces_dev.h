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

}  // namespace bd
