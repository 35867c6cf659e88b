"""Pins the CPU oracle (oracle/bdoracle.c) against what the mathematics and
the reference's own library calls fix.  No GPU needed.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a transposed operand, a wrong sign or index) fails it:
  * RPY: an independent vectorised closed form, continuity of both branches
    at r = 2, symmetry and positive definiteness;
  * QR / eigen: LAPACK dgeqp3/dorgqr and numpy eigh;
  * Cholesky / Lanczos: LAPACK dpotrf, dstev, dsbevd used exactly as the
    reference uses them (tests/lapack_restatement.py), and the limit
    D^(1/2) xi from a dense eigendecomposition;
  * forces: central differences of the energy and sum(F) = 0;
  * pair tables / exclusions: brute force over all pairs and graph rings.
"""
import itertools
import os

import numpy as np
import pytest

import lapack_restatement as LR
from helpers import fene_wca_chain, oracle_system, read_cfg, rel
from brownpak_amd.configs import chain_bonds, periodic_melt, random_walk_chain

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rpy_closed_form(crd):
    """Vectorised RPY for bead radius 1 (independent of the oracle's loops)."""
    N = len(crd)
    r = crd[None, :, :] - crd[:, None, :]
    d = np.linalg.norm(r, axis=2)
    np.fill_diagonal(d, 1.0)
    rr = r[:, :, :, None] * r[:, :, None, :] / (d ** 2)[:, :, None, None]
    far = d >= 2.0
    c1 = np.where(far, 1 + 2 / (3 * d ** 2), 1 - 9 * d / 32)
    c2 = np.where(far, 1 - 2 / d ** 2, 3 * d / 32)
    cons = np.where(far, 0.75 / d, 1.0)
    I3 = np.eye(3)[None, None]
    M = cons[:, :, None, None] * (c1[:, :, None, None] * I3 + c2[:, :, None, None] * rr)
    for i in range(N):
        M[i, i] = np.eye(3)
    return M.transpose(0, 2, 1, 3).reshape(3 * N, 3 * N)


def sqrt_exact(D, xi):
    w, V = np.linalg.eigh(D)
    x2 = xi.reshape(D.shape[0], -1)
    return (V @ (np.sqrt(w)[:, None] * (V.T @ x2))).reshape(xi.shape)


# ------------------------------------------------------------------ RPY
@pytest.mark.parametrize("nb,seed", [(2, 0), (30, 1), (100, 2)])
def test_rpy_closed_form(oracle, nb, seed):
    crd = random_walk_chain(nb, seed)
    D = oracle.rpy(crd)
    assert np.abs(D - rpy_closed_form(crd)).max() < 1e-14
    assert np.array_equal(D, D.T)
    assert np.all(np.diag(D) == 1.0)
    assert np.linalg.eigvalsh(D).min() > 0


def test_rpy_overlap_branch_continuity(oracle):
    # both branches agree at r = 2: 0.625 along r, 0.4375 transverse
    for r in (2.0 - 1e-12, 2.0, 2.0 + 1e-12):
        crd = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
        D = oracle.rpy(crd)
        assert abs(D[0, 3] - 0.625) < 1e-11       # (C1 + C2) * cons along x
        assert abs(D[1, 4] - 0.4375) < 1e-11
        assert abs(D[2, 5] - 0.4375) < 1e-11
    # overlap branch value at r = 1 along x: 1 - 9/32 + 3/32 = 0.8125
    D = oracle.rpy(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]]))
    assert abs(D[0, 3] - 0.8125) < 1e-15 and abs(D[1, 4] - (1 - 9 / 32)) < 1e-15


def test_symv_lower_uses_lower_triangle_only(oracle):
    rng = np.random.default_rng(3)
    crd = random_walk_chain(40, 3)
    D = oracle.rpy(crd)
    x = rng.standard_normal(D.shape[0])
    Dcorrupt = D.copy()
    Dcorrupt[np.triu_indices_from(D, 1)] = 1e9   # dsymv('L') must not look here
    y = oracle.symv_lower(Dcorrupt, x, alpha=1.0, beta=0.0)
    assert rel(y, D @ x) < 1e-14
    y2 = oracle.symv_lower(D, x, alpha=0.5, beta=2.0, y=x)
    assert rel(y2, 0.5 * (D @ x) + 2.0 * x) < 1e-14


# ------------------------------------------------------------ QR / eigen
def test_threaded_rpy_and_symv_match_serial(oracle):
    """The OpenMP paths (host timing on all cores only): the same RPY
    entries bit for bit, and a symv equal up to the order of its sums; the
    serial default is restored."""
    crd = random_walk_chain(300, 11)
    x = np.random.default_rng(12).standard_normal(900)
    D1 = oracle.rpy(crd)
    y1 = oracle.symv_lower(D1, x)
    oracle.set_threads(4)
    try:
        D4 = oracle.rpy(crd)
        y4 = oracle.symv_lower(D1, x)
    finally:
        oracle.set_threads(1)
    assert np.array_equal(D1, D4)
    assert rel(y4, y1) < 1e-14
    assert np.array_equal(oracle.symv_lower(D1, x), y1)


def test_geqp3_orgqr_match_lapack(oracle):
    import scipy.linalg.lapack as lp
    rng = np.random.default_rng(4)
    for m, n in [(50, 4), (300, 3), (20, 1)]:
        a = rng.standard_normal((m, n)) * np.array([1, 10, 0.1, 3])[:n]
        qr, jp, tau = oracle.geqp3(a)
        qr2, jp2, tau2, _, info = lp.dgeqp3(np.asfortranarray(a))
        assert info == 0
        assert np.array_equal(jp, jp2)
        assert np.abs(qr - qr2).max() < 1e-13 * np.abs(a).max()
        q = oracle.orgqr(qr, tau)
        q2, _, _ = lp.dorgqr(qr2, tau2)
        assert np.abs(q - q2[:, :n]).max() < 1e-13


def test_syev_matches_eigh(oracle):
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 40):
        a = rng.standard_normal((n, n))
        a = a + a.T
        w, v, info = oracle.syev(a)
        assert info == 0
        assert np.abs(np.sort(w) - np.linalg.eigvalsh(a)).max() < 1e-12 * max(1, np.abs(w).max())
        assert np.abs(v.T @ v - np.eye(n)).max() < 1e-12
        assert np.abs(v @ np.diag(w) @ v.T - a).max() < 1e-12 * max(1, np.abs(a).max())


# ------------------------------------------------------------ B.xi
def test_cholesky_is_U_xi(oracle):
    crd = random_walk_chain(50, 6)
    D = oracle.rpy(crd)
    xi = np.random.default_rng(6).standard_normal((150, 2))
    b, _, _, ierr = oracle.brn_calc_bdw(oracle.CHOL, D, xi)
    assert ierr == 0
    assert rel(b, LR.cholesky_u_xi(D, xi)) < 1e-13
    bl, _, _, _ = oracle.brn_calc_bdw(oracle.CHOL_LT, D, xi)
    L = np.linalg.cholesky(D)
    assert rel(bl, L @ xi) < 1e-13
    # the reference's U.xi does not have covariance D; L.xi does
    U = L.T
    assert np.abs(U @ U.T - D).max() > 1e-3
    assert np.abs(L @ L.T - D).max() < 1e-13


def test_cholesky_rejects_indefinite(oracle):
    A = np.eye(6)
    A[0, 1] = A[1, 0] = 2.0  # indefinite
    _, _, _, ierr = oracle.brn_calc_bdw(oracle.CHOL, A, np.ones((6, 1)))
    assert ierr == 1


# Lanczos without reorthogonalisation (the reference's algorithm) loses
# orthogonality after ~20 iterations; from then on a 1e-16 perturbation of D
# moves the converged B.xi by ~tol (see test_reference_lanczos_sensitivity).
# Rounding-level parity is therefore pinned (a) before orthogonality is lost
# (fixed iteration count, tol = 0) and (b) with full reorthogonalisation;
# converged runs of the reference algorithm agree to the tolerance level.
@pytest.mark.parametrize("nb,s,seed", [(30, 1, 7), (100, 1, 8), (100, 4, 11), (60, 3, 13), (100, 2, 12)])
def test_lanczos_fixed_iterations_match_lapack(oracle, nb, s, seed):
    D = oracle.rpy(random_walk_chain(nb, seed))
    xi = np.random.default_rng(seed).standard_normal((3 * nb, s))
    mx = 7 if s > 1 else 12
    b, lc, it, ierr = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=mx, lanc_tol=0.0)
    ref = LR.lanczos_s1(D, xi[:, 0], mx, 0.0) if s == 1 else LR.block_lanczos(D, xi, mx, 0.0)
    assert ierr == 0 and not lc and it == mx + 1 and ref[2] == mx + 1
    assert rel(b.reshape(ref[0].shape), ref[0]) < 1e-12


@pytest.mark.parametrize("nb,s,seed,tol", [(100, 1, 8, 1e-6), (100, 1, 9, 1e-10), (100, 4, 11, 1e-6),
                                           (100, 2, 12, 1e-8), (60, 3, 13, 1e-6)])
def test_lanczos_reorth_converged_matches_lapack(oracle, nb, s, seed, tol):
    D = oracle.rpy(random_walk_chain(nb, seed))
    xi = np.random.default_rng(seed).standard_normal((3 * nb, s))
    b, lc, it, ierr = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=60, lanc_tol=tol, reorth=True)
    ref = (LR.lanczos_s1(D, xi[:, 0], 60, tol, reorth=True) if s == 1
           else LR.block_lanczos(D, xi, 60, tol, reorth=True))
    assert ierr == 0 and lc and ref[1] and it == ref[2]
    assert rel(b.reshape(ref[0].shape), ref[0]) < 1e-11
    assert rel(b, sqrt_exact(D, xi)) < 100 * tol


@pytest.mark.parametrize("nb,s,seed,tol", [(100, 1, 8, 1e-6), (100, 4, 11, 1e-6), (60, 3, 13, 1e-6)])
def test_lanczos_reference_converged_agreement(oracle, nb, s, seed, tol):
    D = oracle.rpy(random_walk_chain(nb, seed))
    xi = np.random.default_rng(seed).standard_normal((3 * nb, s))
    b, lc, it, ierr = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=60, lanc_tol=tol)
    ref = LR.lanczos_s1(D, xi[:, 0], 60, tol) if s == 1 else LR.block_lanczos(D, xi, 60, tol)
    assert ierr == 0 and lc and ref[1] and abs(it - ref[2]) <= 1
    assert rel(b.reshape(ref[0].shape), ref[0]) < 10 * tol
    assert rel(b, sqrt_exact(D, xi)) < 100 * tol


def test_reference_lanczos_sensitivity():
    """The reference algorithm itself: a 1e-16 symmetric perturbation of D
    changes its converged B.xi far above rounding (LAPACK restatement)."""
    from oracle import oracle as O
    D = O.rpy(random_walk_chain(100, 8))
    xi = np.random.default_rng(8).standard_normal(300)
    b0, _, _ = LR.lanczos_s1(D, xi, 60, 1e-6)
    E = np.random.default_rng(1).standard_normal(D.shape)
    b1, _, _ = LR.lanczos_s1(D + 0.5e-16 * (E + E.T), xi, 60, 1e-6)
    d = rel(b1, b0)
    print(f"perturbation 1e-16 -> converged B.xi moves by {d:.2e}")
    # far above rounding, of the order of the tolerance
    assert 1e-8 < d < 1e-5


def test_block_lanczos_full_krylov_space_is_exact(oracle):
    # n = 12, s = 4: after 3 iterations the Krylov space is all of R^n
    D = oracle.rpy(random_walk_chain(4, 14))
    xi = np.random.default_rng(14).standard_normal((12, 4))
    b, lc, it, _ = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=3, lanc_tol=0.0)
    assert rel(b, sqrt_exact(D, xi)) < 1e-10


def test_lanczos_nonconvergence_flag(oracle):
    D = oracle.rpy(random_walk_chain(100, 15))
    xi = np.random.default_rng(15).standard_normal((300, 4))
    _, lc, it, ierr = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=1, lanc_tol=1e-12)
    assert ierr == 0 and not lc and it == 3   # dimk_max = max(2, 1) + 1


# ------------------------------------------------------------ forces
def _energy(oracle, sysobj, crd, method="DIR"):
    tab = oracle.pairs(sysobj, crd, method)
    _, en, ierr = oracle.forces(sysobj, crd, tab)
    assert ierr == 0
    return en.energy_tot


@pytest.mark.parametrize("style,params", [(1, [25.0, 2.0]), (2, [7.5, 3.0, 0.2]),
                                          (3, [7.5, 3.0, 1.0, 2.0])])
def test_bond_forces_are_minus_energy_gradient(oracle, style, params):
    crd = random_walk_chain(8, 16, bond=1.9)
    bonds = np.array([[1, i, i + 1] for i in range(1, 8)])
    sysobj = oracle.System(8, bonds, [style], [params], lvdw=False)
    f, en, ierr = oracle.forces(sysobj, crd, None)
    assert ierr == 0
    h = 1e-6
    g = np.zeros_like(crd)
    for i in range(8):
        for d in range(3):
            cp, cm = crd.copy(), crd.copy()
            cp[i, d] += h
            cm[i, d] -= h
            ep = oracle.forces(sysobj, cp, None)[1].energy_tot
            em = oracle.forces(sysobj, cm, None)[1].energy_tot
            g[i, d] = (ep - em) / (2 * h)
    # frc = dU/dr applied as +frc*rij/r on bead i (rij = rj - ri): F = -grad E
    assert rel(f, -g) < 1e-6
    assert np.abs(f.sum(0)).max() < 1e-10


def test_marko_siggia_bond_as_written(oracle):
    """Style 4 (ia_bond_m.f90:334-362) is restated as written: its force is
    not -dU/dr of its energy, and the restatement must not 'fix' that."""
    lp_, rmax = 10.0, 5.0
    sysobj = oracle.System(2, [[1, 1, 2]], [4], [[lp_, rmax]], lvdw=False)
    for r in (1.0, 2.5, 4.0):
        crd = np.array([[0.0, 0, 0], [r, 0, 0]])
        f, en, ierr = oracle.forces(sysobj, crd, None)
        x = r / rmax
        frc = (-x - 0.25 / (1 - x) ** 2 + 0.25) / lp_
        assert ierr == 0 and abs(f[0, 0] - frc) < 1e-14 and abs(f[1, 0] + frc) < 1e-14
        h = 1e-6
        ep = oracle.forces(sysobj, np.array([[0.0, 0, 0], [r + h, 0, 0]]), None)[1].energy_tot
        em = oracle.forces(sysobj, np.array([[0.0, 0, 0], [r - h, 0, 0]]), None)[1].energy_tot
        assert abs((ep - em) / (2 * h) - frc) > 1e-3


def test_lj_force_shifted_and_gradient(oracle):
    crd = random_walk_chain(12, 17, bond=1.9)
    eps, sig, rc = 1.3, 2.0, 5.0
    sysobj = oracle.System(12, [], [], None, vdw_styles=[1], vdw_params=[[eps, sig, rc]], lvdw=True,
                           rcutoff=rc)
    f, en, _ = oracle.forces(sysobj, crd, oracle.pairs(sysobj, crd))
    h = 1e-6
    g = np.zeros_like(crd)
    for i in range(12):
        for d in range(3):
            cp, cm = crd.copy(), crd.copy()
            cp[i, d] += h
            cm[i, d] -= h
            g[i, d] = (_energy(oracle, sysobj, cp) - _energy(oracle, sysobj, cm)) / (2 * h)
    assert rel(f, -g) < 1e-6
    assert np.abs(f.sum(0)).max() < 1e-10
    # force-shifted: U(rc) = 0 and U'(rc) = 0 for a two-bead system
    two = oracle.System(2, [], [], None, vdw_styles=[1], vdw_params=[[eps, sig, rc]], lvdw=True,
                        rcutoff=rc + 1.0)
    for r in (rc - 1e-7,):
        c2 = np.array([[0.0, 0, 0], [r, 0, 0]])
        ff, ee, _ = oracle.forces(two, c2, oracle.pairs(two, c2))
        assert abs(ee.energy_vdw) < 1e-12 and abs(ff[0, 0]) < 1e-6


def test_fene_overstretch_sets_ierr(oracle):
    crd = np.array([[0.0, 0, 0], [3.01, 0, 0]])
    sysobj = oracle.System(2, [[1, 1, 2]], [2], [[7.5, 3.0, 0.0]], lvdw=False)
    _, _, ierr = oracle.forces(sysobj, crd, None)
    assert ierr == 1


# ------------------------------------------------------------ pair tables
def _brute_pairs(crd, rc, excl, box=None):
    N = len(crd)
    out = set()
    for i in range(N):
        for j in range(i + 1, N):
            if (i + 1, j + 1) in excl:
                continue
            r = crd[j] - crd[i]
            if box is not None:
                r = r - box * np.sign(r / box) * np.floor(np.abs(r / box) + 0.5)
            if float(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) < rc * rc:
                out.add((i + 1, j + 1))
    return out


def _ring_exclusions(nb, bonds, k):
    adj = {i: set() for i in range(1, nb + 1)}
    for _, i, j in bonds:
        adj[i].add(j)
        adj[j].add(i)
    ex = set()
    for i in range(1, nb + 1):
        if k == 0:
            continue
        frontier, seen = {i}, {i}
        for _ in range(k):
            frontier = set().union(*[adj[a] for a in frontier]) if frontier else set()
            seen |= frontier
        for j in seen:
            ex.add((min(i, j), max(i, j)))
    return ex


@pytest.mark.parametrize("excl", [0, 1, 2, 3])
def test_exclusions_are_graph_rings(oracle, excl):
    nb = 20
    bonds = [[1, i, i + 1] for i in range(1, nb)] + [[1, 5, 15]]
    sysobj = oracle.System(nb, bonds, [1], [[1.0, 1.0]], excluded_atoms=excl)
    ptr, col = oracle.exclusions(sysobj)
    got = set()
    for i in range(nb):
        row = col[ptr[i]:ptr[i + 1]]
        assert np.all(np.diff(row) > 0)
        for j in row:
            got.add((min(i + 1, j), max(i + 1, j)))
    assert got == _ring_exclusions(nb, bonds, excl)


def test_dir_pairs_match_brute_force(oracle):
    m = fene_wca_chain(120, 18)
    sysobj = oracle_system(oracle, m)
    tab = oracle.pairs(sysobj, m["crd"], "DIR")
    assert oracle.pair_set(*tab) == _brute_pairs(m["crd"], m["rcutoff"], set())


def test_chns10_pairs_forces_and_cell_list(oracle):
    m = read_cfg(os.path.join(GOLDEN, "chns-10.cfg"))
    nb = len(m["crd"])
    box = np.array(m["box"])
    base = dict(atom_types=m["atom_types"], vdw_styles=m["vdw_styles"], vdw_params=m["vdw_params"],
                natom_types=m["natom_types"], excluded_atoms=1, rcutoff=m["vdw_params"][0][2],
                imcon=1, box=m["box"])
    sdir = oracle.System(nb, m["bonds"], m["bond_styles"], m["bond_params"], mth_ptgen=0, **base)
    excl = _ring_exclusions(nb, m["bonds"], 1)
    tab = oracle.pairs(sdir, m["crd"], "DIR")
    assert oracle.pair_set(*tab) == _brute_pairs(m["crd"], base["rcutoff"], excl, box)
    # CL: every pair within rc is in the (unfiltered) cell-list set, no pair
    # appears twice with >= 3 cells per side, and forces agree with DIR
    scl = oracle.System(nb, m["bonds"], m["bond_styles"], m["bond_params"], mth_ptgen=2, **base)
    ptr, col = oracle.pairs(scl, m["crd"], "CL")
    cl_pairs = [(min(i + 1, int(j)), max(i + 1, int(j))) for i in range(nb) for j in col[ptr[i]:ptr[i + 1]]]
    assert len(cl_pairs) == len(set(cl_pairs))
    assert oracle.pair_set(*tab) <= set(cl_pairs)
    f1, e1, _ = oracle.forces(sdir, m["crd"], tab)
    f2, e2, _ = oracle.forces(scl, m["crd"], (ptr, col))
    assert rel(f2, f1) < 1e-13 and abs(e1.energy_tot - e2.energy_tot) <= 1e-13 * abs(e1.energy_tot)
    assert np.abs(f1.sum(0)).max() < 1e-13 * np.abs(f1).max()


def _cl_multiset_brute(crd, L, rc, excl):
    """Cell-list pair multiplicities enumerated pair by pair (not cell by
    cell, as the oracle does): same cell -> 1; otherwise the number of the
    13 forward offsets d with cell(i) + d = cell(j) (mod nc) plus those with
    cell(j) + d = cell(i)."""
    from collections import Counter
    dst = [(1, 0, 0), (1, 1, 0), (-1, 1, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 1), (1, 0, 1), (-1, -1, 1),
           (0, -1, 1), (1, -1, 1), (-1, 1, 1), (0, 1, 1), (1, 1, 1)]
    nc = np.floor(np.asarray(L) / rc).astype(int)
    cs = np.asarray(L) / nc
    cell = np.minimum((crd / cs).astype(int), nc - 1)
    out = Counter()
    nb = len(crd)
    for i in range(nb):
        for j in range(i + 1, nb):
            if (i + 1, j + 1) in excl:
                continue
            ci, cj = cell[i], cell[j]
            if (ci == cj).all():
                m = 1
            else:
                m = sum(1 for d in dst if (((ci + d) % nc) == cj).all()) + \
                    sum(1 for d in dst if (((cj + d) % nc) == ci).all())
            if m:
                out[(i + 1, j + 1)] += m
    return out


def test_cell_list_repeats_pairs_on_two_cell_grids(oracle):
    """A box of 2 cells per side (cell_list_m.f90:134-136): the reference's
    cell-list table repeats pairs; the oracle's multiplicities equal a
    pair-by-pair enumeration, and its forces count a repeated pair twice."""
    from collections import Counter
    L = 11.0
    crd = np.random.default_rng(45).uniform(0.0, L, (60, 3))
    nb = len(crd)
    bonds = np.concatenate([chain_bonds(20, offset=c * 20) for c in range(3)])
    so = oracle.System(nb, bonds, [1], [[25.0, 2.0]], vdw_styles=[1], vdw_params=[[1.0, 2.0, 5.0]],
                       excluded_atoms=1, mth_ptgen=2, rcutoff=5.0, imcon=1, box=(L, L, L))
    excl = _ring_exclusions(nb, bonds, 1)
    ptr, col = oracle.pairs(so, crd, "CL", cap=nb * nb * 16)
    got = Counter((min(i + 1, int(j)), max(i + 1, int(j))) for i in range(nb) for j in col[ptr[i]:ptr[i + 1]])
    want = _cl_multiset_brute(crd, (L, L, L), 5.0, excl)
    assert got == want
    assert max(want.values()) > 1
    # forces: each listed pair once, repeats included
    f, _, _ = oracle.forces(so, crd, (ptr, col))
    sd = oracle.System(nb, bonds, [1], [[25.0, 2.0]], vdw_styles=[1], vdw_params=[[1.0, 2.0, 5.0]],
                       excluded_atoms=1, mth_ptgen=0, rcutoff=5.0, imcon=1, box=(L, L, L))
    fd, _, _ = oracle.forces(sd, crd, oracle.pairs(sd, crd, "DIR"))
    assert np.abs(f - fd).max() > 1.0   # the repeats change the forces


def test_verlet_schedule(oracle):
    crd, L = periodic_melt(4, 20, 0.3, 19)
    nb = len(crd)
    bonds = np.concatenate([np.array([[1, c * 20 + i, c * 20 + i + 1] for i in range(1, 20)]) for c in range(4)])
    sysobj = oracle.System(nb, bonds, [1], [[25.0, 2.0]], vdw_styles=[1], vdw_params=[[1.0, 2.0, 5.0]],
                           excluded_atoms=1, mth_ptgen=1, rcutoff=5.0, tskin=0.6, imcon=1, box=(L, L, L))
    st = oracle.VerletState(sysobj)
    t0 = oracle.pairs(sysobj, crd, "VER", verlet_state=st)
    excl = _ring_exclusions(nb, bonds, 1)
    assert oracle.pair_set(*t0) == _brute_pairs(crd, 5.6, excl, np.array([L, L, L]))
    # a displacement below skin/2 keeps the list, one above rebuilds it
    small = crd.copy()
    small[0, 0] += 0.29
    assert oracle.pairs(sysobj, small, "VER", verlet_state=st) is None
    big = crd.copy()
    big[0, 0] += 0.31
    assert oracle.pairs(sysobj, big, "VER", verlet_state=st) is not None


# ------------------------------------------------------------ the step
def test_free_draining_step_closed_form(oracle):
    m = fene_wca_chain(20, 20)
    sysobj = oracle_system(oracle, m)
    rng = np.random.default_rng(20)
    xi = rng.standard_normal(60)
    dt = 1e-3
    crd0 = m["crd"].copy()
    crd0 -= crd0.mean(0)
    out, used, info = oracle.bd_run(sysobj, m["crd"], 1, xi, lhdia=False, tim_stp=dt)
    f, _, _ = oracle.forces(sysobj, crd0, oracle.pairs(sysobj, crd0))
    expect = crd0 + dt * f + np.sqrt(2 * dt) * xi.reshape(20, 3)
    com = expect.mean(0)
    assert used == 60 and info.nts == 1
    assert rel(out, expect - com) < 1e-13
    assert np.allclose(info.molecule_com, com, atol=1e-14)


def test_hi_step_closed_form_and_reuse_quirk(oracle):
    m = fene_wca_chain(15, 21)
    sysobj = oracle_system(oracle, m)
    n = 45
    s = 2
    rng = np.random.default_rng(21)
    xi = rng.standard_normal(n * s)
    dt = 1e-3
    c0 = m["crd"] - m["crd"].mean(0)
    D = oracle.rpy(c0)
    bdw = np.linalg.cholesky(D).T @ xi.reshape(s, n).T     # reference U.xi, (n, s)
    f0, _, _ = oracle.forces(sysobj, c0, oracle.pairs(sysobj, c0))
    c1 = c0 + dt * (D @ f0.ravel()).reshape(-1, 3) + np.sqrt(2 * dt) * bdw[:, 0].reshape(-1, 3)
    c1 -= c1.mean(0)
    f1, _, _ = oracle.forces(sysobj, c1, oracle.pairs(sysobj, c1))
    # second step reuses D (assembled at c0); the reference has zeroed the noise
    c2_ref = c1 + dt * (D @ f1.ravel()).reshape(-1, 3)
    c2_ref -= c2_ref.mean(0)
    c2_keep = c1 + dt * (D @ f1.ravel()).reshape(-1, 3) + np.sqrt(2 * dt) * bdw[:, 1].reshape(-1, 3)
    c2_keep -= c2_keep.mean(0)
    out, used, info = oracle.bd_run(sysobj, m["crd"], 2, xi, mob_fctr=oracle.CHOL, nts_mobsam=s)
    assert used == n * s and info.n_mob_updates == 1
    assert rel(out, c2_ref) < 1e-12
    out2, _, _ = oracle.bd_run(sysobj, m["crd"], 2, xi, mob_fctr=oracle.CHOL, nts_mobsam=s,
                               keep_reuse_noise=True)
    assert rel(out2, c2_keep) < 1e-12


# ------------------------------------------------------------ predictor-corrector
# Not in the reference: pinned by what Heun's method fixes for linear drifts
# (its amplification factor 1 + h a + (h a)^2 / 2), by its order, and by the
# Brownian term entering the predictor once and the update once.
def test_pc_linear_spring_heun_factor(oracle):
    """Harmonic dimer (r0 = 0), free draining, no noise: the separation obeys
    d' = -2k d, so one step multiplies it by 1 - 2kh + 2(kh)^2 (EM: 1 - 2kh)."""
    k, h, nst = 5.0, 0.01, 7
    sysobj = oracle.System(2, [[1, 1, 2]], [1], [[k, 0.0]], vdw_styles=(), lvdw=False)
    crd = np.array([[0.3, -0.2, 0.5], [1.1, 0.4, -0.7]])
    d0 = crd[1] - crd[0]
    xi = np.zeros(6 * nst)
    for integ, g in ((oracle.EM, 1 - 2 * k * h), (oracle.PC, 1 - 2 * k * h + 2 * (k * h) ** 2)):
        out, used, info = oracle.bd_run(sysobj, crd, nst, xi, lhdia=False, tim_stp=h, integrator=integ)
        assert used == 6 * nst and info.nts == nst
        assert np.allclose(out[1] - out[0], d0 * g ** nst, rtol=1e-13, atol=0)
        assert np.allclose(out.mean(0), 0.0, atol=1e-15)


def test_pc_planar_flow_with_noise_closed_form(oracle):
    """Planar elongation (x' = e x, y' = -e y), no forces, one step with
    noise w: x1 = x (1 + eh + (eh)^2/2) + w (1 + eh/2), the extra w eh/2
    coming from the predictor's drift (EM: x (1 + eh) + w)."""
    e, h, nb = 3.0, 0.02, 4
    sysobj = oracle.System(nb, [], [], None, vdw_styles=(), lvdw=False)
    crd = np.random.default_rng(40).standard_normal((nb, 3))
    xi = np.random.default_rng(41).standard_normal(3 * nb)
    w = np.sqrt(2 * h) * xi.reshape(nb, 3)
    c0 = crd - crd.mean(0)
    a = np.array([e, -e, 0.0]) * h
    for integ in (oracle.EM, oracle.PC):
        out, _, _ = oracle.bd_run(sysobj, crd, 1, xi, lhdia=False, tim_stp=h, integrator=integ,
                                  flow_style=2, flow_rate=e)
        if integ == oracle.EM:
            expect = c0 * (1 + a) + w
        else:
            expect = c0 * (1 + a + a * a / 2) + w * (1 + a / 2)
        expect -= expect.mean(0)
        assert rel(out, expect) < 1e-13, integ


def test_pc_second_order_fene_chain(oracle):
    """Deterministic FENE chain (smooth forces), free draining: halving dt
    divides the PC error at a fixed time by ~4 (EM: ~2).  With RPY the
    mobility is held over the step (as over nts_mobsam steps), so the error
    from D varying along the step stays first order, as in EM."""
    m = fene_wca_chain(8, 42)
    sysobj = oracle.System(8, m["bonds"], m["bond_styles"], m["bond_params"], vdw_styles=(), lvdw=False)
    T = 0.08
    def run(integ, h):
        nst = int(round(T / h))
        n = 3 * len(m["crd"])
        out, _, _ = oracle.bd_run(sysobj, m["crd"], nst, np.zeros(n * nst), lhdia=False,
                                  tim_stp=h, integrator=integ)
        return out
    ref = run(oracle.PC, T / 6400)
    for integ, order in ((oracle.PC, 2), (oracle.EM, 1)):
        e1 = np.linalg.norm(run(integ, T / 80) - ref)
        e2 = np.linalg.norm(run(integ, T / 160) - ref)
        ratio = e1 / e2
        assert abs(np.log2(ratio) - order) < 0.25, (integ, ratio)


def test_periodic_wrap(oracle):
    crd, L = periodic_melt(2, 10, 0.3, 22)
    nb = len(crd)
    bonds = np.concatenate([np.array([[1, c * 10 + i, c * 10 + i + 1] for i in range(1, 10)]) for c in range(2)])
    sysobj = oracle.System(nb, bonds, [1], [[25.0, 2.0]], vdw_styles=[1], vdw_params=[[1.0, 2.0, 5.0]],
                           excluded_atoms=1, mth_ptgen=0, rcutoff=5.0, imcon=1, box=(L, L, L))
    xi = np.random.default_rng(22).standard_normal(3 * nb * 5) * 30
    out, _, _ = oracle.bd_run(sysobj, crd, 5, xi, lhdia=False)
    assert np.all(out >= 0) and np.all(out < L)


def test_rgsq_closed_form(oracle):
    crd = random_walk_chain(50, 23)
    assert abs(oracle.rgsq(crd) - np.mean(np.sum((crd - crd.mean(0)) ** 2, 1))) < 1e-12


# ------------------------------------------------------------ closed forms of the force laws
def _dimer(oracle, style_kw, r):
    sysobj = oracle.System(2, *style_kw)
    f, en, ierr = oracle.forces(sysobj, np.array([[0.0, 0, 0], [r, 0, 0]]), None)
    return f, en, ierr


def test_force_values_closed_forms(oracle):
    """Force and energy values at points where the laws have textbook
    values (a wrong prefactor or parameter slot fails here even when the
    force is the energy's gradient):
      FENE (k, Q0, r0 = 0) at r = Q0/2: f = (2/3) k Q0, U = -k Q0^2 ln(3/4) / 2;
      FENE with r0: the same at r = r0 + Q0/2;
      harmonic (k, r0) at r = r0 + 1/2: f = k/2, U = k/8;
      WCA (rc = 2^(1/6) sigma) at r = sigma: U = eps, dU/dr = -24 eps / sigma;
      at r = 2^(1/6) sigma: U = 0, dU/dr = 0;
      force-shifted LJ (rc = 2.5 sigma): dU/dr(sigma) - dU/dr(2^(1/6) sigma)
      = -24 eps / sigma (the shift cancels)."""
    k, q0 = 7.5, 3.0
    for r0 in (0.0, 0.8):
        f, en, ierr = _dimer(oracle, ([[1, 1, 2]], [2], [[k, q0, r0]]), r0 + q0 / 2)
        assert ierr == 0
        assert abs(f[0, 0] - 2.0 / 3.0 * k * q0) < 1e-13 and abs(f[1, 0] + 2.0 / 3.0 * k * q0) < 1e-13
        assert abs(en.energy_bond - (-0.5 * k * q0 ** 2 * np.log(0.75))) < 1e-13
    f, en, _ = _dimer(oracle, ([[1, 1, 2]], [1], [[25.0, 2.0]]), 2.5)
    assert abs(f[0, 0] - 12.5) < 1e-13 and abs(en.energy_bond - 25.0 / 8) < 1e-13
    eps, sig = 1.3, 2.0
    rm = 2.0 ** (1.0 / 6.0) * sig
    wca = ([], [], None)
    kw = dict(vdw_styles=[1], vdw_params=[[eps, sig, rm]], lvdw=True, rcutoff=rm)

    def pair(r, params, rc):
        sysobj = oracle.System(2, [], [], None, vdw_styles=[1], vdw_params=[params], lvdw=True, rcutoff=rc)
        crd = np.array([[0.0, 0, 0], [r, 0, 0]])
        f, en, ierr = oracle.forces(sysobj, crd, oracle.pairs(sysobj, crd))
        assert ierr == 0
        return f[0, 0], en.energy_vdw

    f, u = pair(sig, [eps, sig, rm], rm)
    assert abs(u - eps) < 1e-13 and abs(f - (-24 * eps / sig)) < 1e-12
    f, u = pair(rm * (1 - 1e-12), [eps, sig, rm], rm)
    assert abs(u) < 1e-10 and abs(f) < 1e-9
    rc = 2.5 * sig
    f1, _ = pair(sig, [eps, sig, rc], rc)
    f2, _ = pair(rm, [eps, sig, rc], rc)
    assert abs((f1 - f2) - (-24 * eps / sig)) < 1e-12


# ------------------------------------------------------------ replay branch of the oracle
@pytest.mark.parametrize("case", ["chol_s1", "lanczos_s4"])
def test_oracle_replay_reproduces_the_run(oracle, case):
    """orc_bd_run_ex with recorded displacements (the noise-held-fixed
    trajectory of the differentiable path) against the ordinary run: the
    displacement of a mobility update is sqrt(2 dt) B.xi at that update's
    coordinates, computed independently with brn_calc_bdw, and zero on the
    later steps of a block (the zeroed diffusion block, bd_solver_m.f90:364);
    replaying those must give the run's coordinates to rounding."""
    m = fene_wca_chain(40, 41)
    so = oracle_system(oracle, m)
    n, dt = 120, 1e-3
    if case == "chol_s1":
        s, nblk, kw = 1, 5, dict(mob_fctr=oracle.CHOL, nts_mobsam=1)
        bkw = dict()
        meth = oracle.CHOL
    else:
        s, nblk, kw = 4, 2, dict(mob_fctr=oracle.LANC, nts_mobsam=4, lanc_reorth=True)
        bkw = dict(reorth=True)
        meth = oracle.LANC
    rng = np.random.default_rng(5)
    crd = m["crd"].copy()
    start = crd.copy()
    disp = []
    for _ in range(nblk):
        xi = rng.standard_normal(n * s)
        c0 = crd - crd.mean(axis=0)     # bds_run moves the COM to the origin first (bd_solver_m.f90:161)
        bxi, _, _, ierr = oracle.brn_calc_bdw(meth, oracle.rpy(c0), xi.reshape(s, n).T, **bkw)
        assert ierr == 0
        disp.append(np.sqrt(2 * dt) * bxi[:, 0].reshape(-1, 3))
        disp.extend([np.zeros((40, 3))] * (s - 1))
        crd, used, _ = oracle.bd_run(so, crd, s, xi, **kw)
        assert used == n * s
    rep = oracle.bd_run_replay(so, start, np.array(disp), nts_mobsam=s)
    assert rel(rep, crd) < 1e-13
