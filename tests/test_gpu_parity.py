"""Parity of the HIP path (through the C ABI) with the CPU oracle on the
same positions and the same injected Gaussian vectors.

Bars:
  * pair sets: bit-exact;
  * mobility, D.V, Cholesky B.xi, forces, energies, stress, steps: 1e-10
    relative in FP64 (measured norm-wise);
  * Lanczos: 1e-10 at a fixed iteration count and with FP64
    reorthogonalisation; converged runs of the reference algorithm (no
    reorthogonalisation) within 10 * lanc_tol and +-1 iteration, because that
    algorithm's own result moves by ~tol under 1e-16 perturbations
    (tests/test_oracle.py::test_reference_lanczos_sensitivity);
  * FP32-storage mixed precision: <= 1e-6 relative against FP64.
"""
import os

import numpy as np
import pytest

from helpers import fene_wca_chain, gpu_system, oracle_system, read_cfg, rel
from brownpak_amd.configs import chain_bonds, periodic_melt, random_walk_chain

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TOL = 1e-10
# mobility products from stored tiles and matrix-free (evaluated in place)
STORAGES = ["tiles", "matfree"]


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def chain100():
    return fene_wca_chain(100, 101)


@pytest.fixture(scope="module")
def chain2000():
    return fene_wca_chain(2000, 2000)


# ------------------------------------------------------------- mobility
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rpy_assembly_n100(oracle, chain100, prec):
    g = gpu_system(chain100, mob_fctr="CHOL", precision=prec)
    g.calc_rpy_tensor()
    D = g.get_mobility()
    Dref = oracle.rpy(chain100["crd"])
    bound = 1e-15 if prec == "fp64" else 6e-8
    assert np.abs(D - Dref).max() <= bound * np.abs(Dref).max()
    assert np.array_equal(D, D.T)


@pytest.mark.parametrize("storage", STORAGES)
def test_rpy_assembly_and_product_n2000(oracle, chain2000, storage):
    _say(f"case: rpy n2000 {storage}")
    g = gpu_system(chain2000, mob_fctr="LANC", nts_mobsam=4, mob_storage=storage)
    g.calc_rpy_tensor()
    D = g.get_mobility()
    Dref = oracle.rpy(chain2000["crd"])
    assert np.abs(D - Dref).max() <= 1e-15
    x = np.random.default_rng(7).standard_normal((6000, 4))
    y = g.mob_symv(x)
    assert rel(y, Dref @ x) < 1e-13
    y1 = g.mob_symv(x[:, 0])
    assert rel(y1, Dref @ x[:, 0]) < 1e-13


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("s", [1, 2, 3, 4, 5, 8])
def test_mob_symv_block_sizes(oracle, chain100, s, storage):
    g = gpu_system(chain100, mob_fctr="CHOL", nts_mobsam=s, mob_storage=storage)
    g.calc_rpy_tensor()
    Dref = oracle.rpy(chain100["crd"])
    x = np.random.default_rng(s).standard_normal((300, s))
    assert rel(g.mob_symv(x), Dref @ x) < 1e-13


# ------------------------------------------------------------- Cholesky
@pytest.mark.parametrize("method", ["CHOL", "CHOL_LT"])
def test_cholesky_bdw(oracle, chain100, method):
    g = gpu_system(chain100, mob_fctr=method, nts_mobsam=2)
    g.calc_rpy_tensor()
    xi = np.random.default_rng(11).standard_normal((300, 2))
    out, lc, it = g.brn_calc_bdw(xi, method=method)
    ref, _, _, ierr = oracle.brn_calc_bdw(oracle.CHOL if method == "CHOL" else oracle.CHOL_LT,
                                          oracle.rpy(chain100["crd"]), xi)
    assert ierr == 0 and rel(out, ref) < TOL


@pytest.mark.parametrize("nb", [160, 161, 400])
def test_cholesky_bdw_sizes(oracle, nb):
    """Both factorisations: one workgroup up to n = 480 (N = 160), 32-column
    panels over many workgroups above."""
    m = fene_wca_chain(nb, 8)
    n = 3 * nb
    xi = np.random.default_rng(12).standard_normal((n, 1))
    for method in ("CHOL", "CHOL_LT"):
        g = gpu_system(m, mob_fctr=method, nts_mobsam=1)
        g.calc_rpy_tensor()
        out, _, _ = g.brn_calc_bdw(xi, method=method)
        ref, _, _, ierr = oracle.brn_calc_bdw(oracle.CHOL if method == "CHOL" else oracle.CHOL_LT,
                                              oracle.rpy(m["crd"]), xi)
        assert ierr == 0 and rel(out, ref) < TOL, (nb, method)


@pytest.mark.parametrize("nb,k", [(10, 3), (100, 50), (170, 120)])
def test_cholesky_reports_a_failed_pivot(oracle, nb, k):
    """Two coincident beads give RPY blocks with no direction (0/0), so a
    pivot is not positive: the oracle's factorisation fails and so does the
    device's, in the first panel (N = 10), in a later panel of the
    one-workgroup factorisation (N = 100) and of the panelled one (N = 170)."""
    from brownpak_amd import BdError
    m = fene_wca_chain(nb, 8)
    m["crd"] = m["crd"].copy()
    m["crd"][k + 1] = m["crd"][k]
    xi = np.random.default_rng(13).standard_normal((3 * nb, 1))
    _, _, _, ierr = oracle.brn_calc_bdw(oracle.CHOL, oracle.rpy(m["crd"]), xi)
    assert ierr != 0
    for method in ("CHOL", "CHOL_LT"):
        g = gpu_system(m, mob_fctr=method, nts_mobsam=1)
        g.calc_rpy_tensor()
        with pytest.raises(BdError) as ei:
            g.brn_calc_bdw(xi, method=method)
        assert ei.value.code == 2 and "positive definite" in str(ei.value)


# ------------------------------------------------------------- Lanczos
@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("nb,s,mx", [(100, 1, 12), (100, 4, 7), (100, 3, 7), (100, 6, 5), (100, 8, 4),
                                     (2000, 4, 7)])
def test_lanczos_fixed_iterations(oracle, chain100, chain2000, nb, s, mx, storage):
    _say(f"case: lanczos fixed nb={nb} s={s} {storage}")
    m = chain100 if nb == 100 else chain2000
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=s, lanc_mxitr=mx, lanc_tol=0.0, mob_storage=storage)
    g.calc_rpy_tensor()
    xi = np.random.default_rng(nb + s).standard_normal((3 * nb, s))
    out, lc, it = g.brn_calc_bdw(xi)
    ref, lc2, it2, ierr = oracle.brn_calc_bdw(oracle.LANC, oracle.rpy(m["crd"]), xi, lanc_mxitr=mx,
                                              lanc_tol=0.0)
    assert not lc and not lc2 and it == it2 == mx + 1
    assert rel(out, ref) < TOL


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("nb,s", [(100, 1), (100, 4), (100, 8), (2000, 4)])
def test_lanczos_reorth_converged(oracle, chain100, chain2000, nb, s, storage):
    _say(f"case: lanczos reorth nb={nb} s={s} {storage}")
    m = chain100 if nb == 100 else chain2000
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=s, lanc_mxitr=50, lanc_tol=1e-6, lanc_reorth=True,
                   mob_storage=storage)
    g.calc_rpy_tensor()
    xi = np.random.default_rng(nb * s).standard_normal((3 * nb, s))
    out, lc, it = g.brn_calc_bdw(xi)
    ref, lc2, it2, _ = oracle.brn_calc_bdw(oracle.LANC, oracle.rpy(m["crd"]), xi, lanc_mxitr=50,
                                           lanc_tol=1e-6, reorth=True)
    _say(f"  iters gpu {it} oracle {it2} rel {rel(out, ref):.2e}")
    assert lc and lc2 and it == it2
    assert rel(out, ref) < TOL


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("nb,s", [(100, 1), (100, 4), (2000, 4)])
def test_lanczos_reference_converged(oracle, chain100, chain2000, nb, s, storage):
    _say(f"case: lanczos reference nb={nb} s={s} {storage}")
    m = chain100 if nb == 100 else chain2000
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=s, lanc_mxitr=50, lanc_tol=1e-6, mob_storage=storage)
    g.calc_rpy_tensor()
    xi = np.random.default_rng(nb + 7 * s).standard_normal((3 * nb, s))
    out, lc, it = g.brn_calc_bdw(xi)
    D = oracle.rpy(m["crd"])
    ref, lc2, it2, _ = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=50, lanc_tol=1e-6)
    _say(f"  iters gpu {it} oracle {it2} rel {rel(out, ref):.2e}")
    assert lc and lc2 and abs(it - it2) <= 1
    assert rel(out, ref) < 1e-5
    if nb == 2000 and storage == "tiles":
        # the oracle's own spread under 1e-16 symmetric perturbations of D:
        # the device result lies inside that envelope
        spread = 0.0
        for q in range(3):
            E = np.random.default_rng(100 + q).standard_normal(D.shape)
            rp, _, itp, _ = oracle.brn_calc_bdw(oracle.LANC, D + 0.5e-16 * (E + E.T), xi, lanc_mxitr=50,
                                                lanc_tol=1e-6)
            spread = max(spread, rel(rp, ref))
            _say(f"  perturbed oracle: iters {itp} moved by {rel(rp, ref):.2e}")
        _say(f"  gpu vs oracle {rel(out, ref):.2e}; oracle perturbation spread {spread:.2e}")
        assert rel(out, ref) <= 20 * spread


@pytest.mark.parametrize("nb", [100, 2000])
def test_lanczos_repeated_solves(oracle, chain100, chain2000, nb):
    """Several solves in one context: from the second on, the tests are
    grouped around the previous solve's converged iteration.  Solves that
    converge at, well before (a block of near-rigid translations) and after
    that iteration must each match the oracle's iteration and result."""
    _say(f"case: lanczos repeated solves nb={nb}")
    m = chain100 if nb == 100 else chain2000
    s = 4
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=s, lanc_mxitr=50, lanc_tol=1e-6)
    g.calc_rpy_tensor()
    D = oracle.rpy(m["crd"])
    rng = np.random.default_rng(nb + 11)
    smooth = np.zeros((3 * nb, s))
    for k in range(s):
        smooth[k % 3::3, k] = 1.0 + 0.01 * rng.standard_normal(nb)
    cases = [rng.standard_normal((3 * nb, s)), rng.standard_normal((3 * nb, s)), smooth,
             rng.standard_normal((3 * nb, s))]
    iters = []
    for xi in cases:
        out, lc, it = g.brn_calc_bdw(xi)
        ref, lc2, it2, _ = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=50, lanc_tol=1e-6)
        _say(f"  iters gpu {it} oracle {it2} rel {rel(out, ref):.2e}")
        assert lc and lc2 and abs(it - it2) <= 1
        assert rel(out, ref) < 1e-5
        iters.append(it)
    # the smooth block converges well before the random ones
    assert iters[2] < iters[1] - 1


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("nb", [16, 20])
def test_lanczos_converged_before_krylov_exhaustion(oracle, nb, storage):
    """Short chains at s = 4: the Krylov space is complete after ceil(3N/4)
    blocks, and the recurrence issued ahead of the tests tries to form
    blocks beyond it.  A CholeskyQR failure there belongs to an iteration
    after convergence and must not void the converged result (the reference
    stops at convergence, brown_m.f90:497-506).  Several solves in one
    context, against the oracle as the converged-reference bar."""
    _say(f"case: exhaustion nb={nb} {storage}")
    m = fene_wca_chain(nb, nb)
    D = oracle.rpy(m["crd"])
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=50, lanc_tol=1e-6, mob_storage=storage)
    g.calc_rpy_tensor()
    rng = np.random.default_rng(nb)
    for _ in range(4):
        xi = rng.standard_normal((3 * nb, 4))
        out, lc, it = g.brn_calc_bdw(xi)
        ref, lc2, it2, ierr = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=50, lanc_tol=1e-6)
        _say(f"  iters gpu {it} oracle {it2} rel {rel(out, ref):.2e}")
        assert lc and lc2 and ierr == 0 and abs(it - it2) <= 1
        assert rel(out, ref) < 1e-5


def test_lanczos_exhausted_before_convergence_is_an_error(oracle):
    """N = 10, s = 4: the reference converges only at iteration 9, one past
    the complete Krylov space (8 blocks), through a pivoted Householder QR of
    a rank-deficient block; CholeskyQR cannot form that block, so the device
    solve reports BD_ERR_FACTOR (such chains take mob_fctr CHOL)."""
    from brownpak_amd import BdError
    m = fene_wca_chain(10, 10)
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=50, lanc_tol=1e-6)
    g.calc_rpy_tensor()
    xi = np.random.default_rng(10).standard_normal((30, 4))
    _, lc2, it2, _ = oracle.brn_calc_bdw(oracle.LANC, oracle.rpy(m["crd"]), xi, lanc_mxitr=50, lanc_tol=1e-6)
    assert lc2 and it2 == 9
    with pytest.raises(BdError) as ei:
        g.brn_calc_bdw(xi)
    assert ei.value.code == 2


def test_lanczos_not_converged_flag(chain100):
    g = gpu_system(chain100, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=1, lanc_tol=1e-12)
    g.calc_rpy_tensor()
    out, lc, it = g.brn_calc_bdw(np.random.default_rng(3).standard_normal((300, 4)))
    assert not lc and it == 3 and np.all(np.isfinite(out))


@pytest.mark.parametrize("storage", STORAGES)
def test_mixed_precision_error_n2000(chain2000, storage):
    _say(f"case: mixed precision n2000 {storage}")
    xi = np.random.default_rng(5).standard_normal((6000, 4))
    outs = {}
    for prec in ("fp64", "fp32"):
        g = gpu_system(chain2000, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=50, lanc_tol=1e-6,
                       precision=prec, lanc_reorth=True, mob_storage=storage)
        g.calc_rpy_tensor()
        outs[prec] = g.brn_calc_bdw(xi)
        f = np.random.default_rng(6).standard_normal(6000)
        outs[prec + "_df"] = g.mob_symv(f)
    e_b = rel(outs["fp32"][0], outs["fp64"][0])
    e_df = rel(outs["fp32_df"], outs["fp64_df"])
    _say(f"  mixed rel err: B.xi {e_b:.2e}  D.F {e_df:.2e}")
    assert e_b < 1e-6 and e_df < 1e-6


@pytest.mark.parametrize("storage", ["tiles", "matfree"])
def test_mixed_headline_configuration(oracle, chain2000, storage):
    """The configuration the throughput figure is quoted on: N = 2000,
    block Lanczos s = 4, tol 1e-6, no reorthogonalisation, FP32 mobility
    products (stored FP32 tiles with FP64 arithmetic, or FP32 matrix-free),
    several solves in one context so that the tests are grouped around the
    previous solve's converged iteration.  Against FP64 (same storage) and
    the oracle (FP64 reference algorithm).  Bounds:
      * tiles: stopping iteration within 1 of FP64's, B.xi within 2e-6 of
        FP64, one-step displacement within 2e-6;
      * matfree (FP32 arithmetic in the product, ~8e-8 per product, about
        four times the tiles'; measured to stop 1-3 iterations after FP64,
        the reason the automatic choice keeps stored tiles for FP32):
        within 3 iterations, 4e-6 and 4e-6;
      * against the oracle one more iteration (the FP64 device path itself
        stops within 1 of it, test_lanczos_reference_converged) and 1e-5
        (the converged-reference bar)."""
    _say(f"case: mixed headline {storage}")
    itol, btol = (1, 2e-6) if storage == "tiles" else (3, 4e-6)
    g32 = gpu_system(chain2000, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=50, lanc_tol=1e-6, precision="fp32",
                     mob_storage=storage)
    g64 = gpu_system(chain2000, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=50, lanc_tol=1e-6, precision="fp64",
                     mob_storage=storage)
    g32.calc_rpy_tensor()
    g64.calc_rpy_tensor()
    D = oracle.rpy(chain2000["crd"])
    rng = np.random.default_rng(2001)
    for q in range(3):
        xi = rng.standard_normal((6000, 4))
        o32, lc32, it32 = g32.brn_calc_bdw(xi)
        o64, lc64, it64 = g64.brn_calc_bdw(xi)
        ref, lco, ito, _ = oracle.brn_calc_bdw(oracle.LANC, D, xi, lanc_mxitr=50, lanc_tol=1e-6)
        _say(f"  solve {q}: iters fp32 {it32} fp64 {it64} oracle {ito}; fp32 vs fp64 {rel(o32, o64):.2e}, "
             f"vs oracle {rel(o32, ref):.2e}")
        assert lc32 and lc64 and lco
        assert abs(it32 - it64) <= itol and abs(it32 - ito) <= itol + 1
        assert rel(o32, o64) < btol and rel(o32, ref) < 1e-5
    # one EM step from the same state and noise
    xi = rng.standard_normal((6000, 4))
    disp = {}
    for prec in ("fp32", "fp64"):
        g = gpu_system(chain2000, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=50, lanc_tol=1e-6, precision=prec,
                       mob_storage=storage)
        c0 = g.get_coordinates()
        c0 = c0 - c0.mean(axis=0)
        g.bds_run(1, xi.T.ravel())
        disp[prec] = g.get_coordinates() - c0
    _say(f"  displacement fp32 vs fp64 {rel(disp['fp32'], disp['fp64']):.2e}")
    assert rel(disp["fp32"], disp["fp64"]) < btol


# ------------------------------------------------------------- forces
def test_forces_chain(oracle, chain100):
    g = gpu_system(chain100, mob_fctr="CHOL")
    f, e, st, bl = g.ia_calc_forces()
    so = oracle_system(oracle, chain100)
    tab = oracle.pairs(so, chain100["crd"], "DIR")
    fr, en, ierr = oracle.forces(so, chain100["crd"], tab)
    assert ierr == 0
    assert rel(f, fr) < TOL
    assert abs(e[0] - en.energy_bond) <= TOL * abs(en.energy_bond)
    assert abs(e[1] - en.energy_vdw) <= TOL * max(abs(en.energy_vdw), 1e-300)
    assert rel(st.ravel(order="F"), np.array(en.stress)) < TOL
    assert abs(bl[0] - en.bndlen) < 1e-12 and bl[1] == en.bndlen_min and bl[2] == en.bndlen_max
    pairs, _ = g.pair_table()
    assert set(map(tuple, pairs.tolist())) == oracle.pair_set(*tab)


def test_forces_and_pairs_chns10(oracle):
    m = read_cfg(os.path.join(GOLDEN, "chns-10.cfg"))
    m.update(natom_types=1, excluded_atoms=1, rcutoff=m["vdw_params"][0][2], tskin=0.0, mth_ptgen="DIR")
    g = gpu_system(m, lhdia=False)
    f, e, st, bl = g.ia_calc_forces()
    so = oracle_system(oracle, m)
    tab = oracle.pairs(so, m["crd"], "DIR")
    fr, en, _ = oracle.forces(so, m["crd"], tab)
    assert rel(f, fr) < TOL
    assert abs(e[2] - en.energy_tot) <= TOL * abs(en.energy_tot)
    assert rel(st.ravel(order="F"), np.array(en.stress)) < TOL
    pairs, _ = g.pair_table()
    assert set(map(tuple, pairs.tolist())) == oracle.pair_set(*tab)


def _melt(nchains, nper, seed):
    crd, L = periodic_melt(nchains, nper, 0.3, seed)
    bonds = np.concatenate([chain_bonds(nper, offset=c * nper) for c in range(nchains)])
    return dict(crd=crd, bonds=bonds, bond_styles=[1], bond_params=[[25.0, 2.0]], vdw_styles=[1],
                vdw_params=[[1.0, 2.0, 5.0]], natom_types=1, excluded_atoms=1, rcutoff=5.0, tskin=0.6,
                mth_ptgen="VER", imcon=1, box=(L, L, L))


@pytest.mark.parametrize("nchains,nper", [(20, 100), (200, 100)])
def test_verlet_pairs_bit_exact_over_steps(oracle, nchains, nper):
    _say(f"case: verlet melt {nchains}x{nper}")
    m = _melt(nchains, nper, 31)
    nb = nchains * nper
    g = gpu_system(m, lhdia=False, tim_stp=1e-3)
    so = oracle_system(oracle, m, "VER")
    vs = oracle.VerletState(so)
    rng = np.random.default_rng(31)
    oracle_tab = None
    rebuilds = 0
    for step in range(6):
        crd = g.get_coordinates()
        f, _, _, _ = g.ia_calc_forces()
        t = oracle.pairs(so, crd, "VER", verlet_state=vs)
        if t is not None:
            oracle_tab = t
            rebuilds += 1
        pairs, rebuilt = g.pair_table()
        assert rebuilt == (t is not None)
        assert set(map(tuple, pairs.tolist())) == oracle.pair_set(*oracle_tab)
        fr, _, _ = oracle.forces(so, crd, oracle_tab)
        assert rel(f, fr) < TOL
        # a big random kick forces rebuilds on some steps
        xi = rng.standard_normal(3 * nb) * (6.0 if step % 2 else 0.5)
        g.bds_run(1, xi)
    assert rebuilds >= 2


# ------------------------------------------------------------- full steps
def _pair_multiset(pairs):
    from collections import Counter
    return Counter(map(tuple, np.asarray(pairs).reshape(-1, 2).tolist()))


def _oracle_multiset(ptr, col):
    from collections import Counter
    return Counter((min(i + 1, int(j)), max(i + 1, int(j))) for i in range(len(ptr) - 1)
                   for j in col[ptr[i]:ptr[i + 1]])


@pytest.mark.parametrize("case", ["melt_7cells", "box_2cells"])
def test_cell_list_pairs_and_forces(oracle, case):
    """mth_ptgen CL: the table is the reference's (no distance test, pairs
    repeated on grids of fewer than 3 cells per side), as a multiset; forces
    over it to 1e-10; then EM steps against the oracle."""
    _say(f"case: cell list {case}")
    if case == "melt_7cells":
        m = _melt(20, 100, 46)
    else:
        L = 11.0
        crd = np.random.default_rng(47).uniform(0.0, L, (60, 3))
        # spread the beads so that the forces stay moderate
        crd = np.stack(np.meshgrid(*[np.arange(4) * 2.7 + 0.3] * 3, indexing="ij"), -1).reshape(-1, 3)[:60]
        bonds = np.concatenate([chain_bonds(20, offset=c * 20) for c in range(3)])
        m = dict(crd=crd, bonds=bonds, bond_styles=[1], bond_params=[[25.0, 2.0]], vdw_styles=[1],
                 vdw_params=[[1.0, 2.0, 5.0]], natom_types=1, excluded_atoms=1, rcutoff=5.0, tskin=0.0,
                 imcon=1, box=(L, L, L))
    m = dict(m, mth_ptgen="CL", tskin=0.0)
    nb = len(m["crd"])
    g = gpu_system(m, lhdia=False)
    so = oracle_system(oracle, m, "CL")
    f, e, _, _ = g.ia_calc_forces()
    tab = oracle.pairs(so, m["crd"], "CL", cap=nb * nb * 16)
    pairs, rebuilt = g.pair_table()
    assert _pair_multiset(pairs) == _oracle_multiset(*tab)
    fr, en, _ = oracle.forces(so, m["crd"], tab)
    assert rel(f, fr) < TOL
    assert abs(e[2] - en.energy_tot) <= TOL * abs(en.energy_tot)
    nsteps = 5
    xi = np.random.default_rng(48).standard_normal(3 * nb * nsteps)
    g.bds_run(nsteps, xi)
    ref, _, _ = oracle.bd_run(so, m["crd"], nsteps, xi, lhdia=False)
    assert rel(g.get_coordinates(), ref) < TOL


@pytest.mark.parametrize("case", ["melt_7cells", "box_2cells"])
def test_cell_list_overflow_steps_match_oracle(oracle, case):
    """A fresh context stepped at once: its neighbour list starts at 64
    entries per bead, fewer than these tables hold, and the step loop does not
    stop to grow it, so every bead takes the all-bead fallback, which counts
    each partner as often as the cell-list table would (twice along an axis
    of two cells)."""
    _say(f"case: cell list overflow {case}")
    if case == "melt_7cells":
        m = _melt(20, 100, 46)
    else:
        L = 11.0
        crd = np.stack(np.meshgrid(*[np.arange(4) * 2.7 + 0.3] * 3, indexing="ij"), -1).reshape(-1, 3)[:60]
        bonds = np.concatenate([chain_bonds(20, offset=c * 20) for c in range(3)])
        m = dict(crd=crd, bonds=bonds, bond_styles=[1], bond_params=[[25.0, 2.0]], vdw_styles=[1],
                 vdw_params=[[1.0, 2.0, 5.0]], natom_types=1, excluded_atoms=1, rcutoff=5.0, tskin=0.0,
                 imcon=1, box=(L, L, L))
    m = dict(m, mth_ptgen="CL", tskin=0.0)
    nb = len(m["crd"])
    so = oracle_system(oracle, m, "CL")
    ptr, col = oracle.pairs(so, m["crd"], "CL", cap=nb * nb * 16)
    count = np.zeros(nb, dtype=int)   # entries per bead, both directions (partners are 1-based)
    for i in range(nb):
        count[i] += ptr[i + 1] - ptr[i]
        np.add.at(count, np.asarray(col[ptr[i]:ptr[i + 1]], dtype=int) - 1, 1)
    assert count.max() > 64
    nsteps = 3
    xi = np.random.default_rng(49).standard_normal(3 * nb * nsteps)
    g = gpu_system(m, lhdia=False)
    g.bds_run(nsteps, xi)
    ref, _, _ = oracle.bd_run(so, m["crd"], nsteps, xi, lhdia=False)
    assert rel(g.get_coordinates(), ref) < TOL


def test_verlet_list_growth_forces_rebuild(oracle):
    """Verlet lists of a fresh context hold 64 entries per bead, fewer than
    this cutoff (rc = 8, skin 0.6) gives; the first run steps over the
    all-bead fallback and grows the list when it ends.  The grown list holds
    nothing yet, so the second run (small kicks: no displacement trigger)
    must rebuild before its forces.  Both runs together against the oracle's
    six steps."""
    _say("case: verlet list growth")
    m = _melt(8, 100, 53)
    m = dict(m, vdw_params=[[1.0, 2.0, 8.0]], rcutoff=8.0)
    crd0 = m["crd"].copy()
    nb = len(crd0)
    L = m["box"][0]
    d = crd0[:, None, :] - crd0[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d * d).sum(-1)
    near = (r2 < (8.0 + 0.6) ** 2).sum(1) - 1
    assert near.max() > 64
    xi1 = np.random.default_rng(54).standard_normal(3 * nb * 3)
    xi2 = np.random.default_rng(55).standard_normal(3 * nb * 3) * 1e-3
    g = gpu_system(m, lhdia=False, tim_stp=1e-3)
    so = oracle_system(oracle, m, "VER")
    g.bds_run(3, xi1)
    g.bds_run(3, xi2)
    ref, _, _ = oracle.bd_run(so, crd0, 6, np.concatenate([xi1, xi2]), lhdia=False)
    assert rel(g.get_coordinates(), ref) < TOL


def test_cell_list_rejects_negative_coordinates_and_small_boxes():
    from brownpak_amd import BdError
    m = dict(_melt(2, 50, 49), mth_ptgen="CL", tskin=0.0)
    crd = m["crd"].copy()
    crd[3, 0] = -0.6 * m["box"][0]          # x / cs <= -1 (cs = L / 2)
    g = gpu_system(dict(m, crd=crd), lhdia=False)
    with pytest.raises(BdError) as ei:
        g.ia_calc_forces()
    assert ei.value.code == -4
    with pytest.raises(BdError):
        gpu_system(dict(m, box=(9.0, 9.0, 9.0)), lhdia=False)      # one cell of rc = 5 per side
    with pytest.raises(BdError):
        gpu_system(dict(m, imcon=0), lhdia=False)


def test_bd_run_cholesky_c1(oracle, chain100):
    _say("case: bd_run chol c1")
    nsteps = 20
    xi = np.random.default_rng(41).standard_normal(300 * nsteps)
    g = gpu_system(chain100, mob_fctr="CHOL", nts_mobsam=1)
    used = g.bds_run(nsteps, xi)
    out = g.get_coordinates()
    so = oracle_system(oracle, chain100)
    ref, used2, info = oracle.bd_run(so, chain100["crd"], nsteps, xi, mob_fctr=oracle.CHOL, nts_mobsam=1)
    assert used == used2 == 300 * nsteps
    assert rel(out, ref) < TOL
    ri = g.runinfo()
    assert ri["nts"] == nsteps
    assert rel(np.array(ri["molecule_com"]), np.array(info.molecule_com)) < 1e-8


@pytest.mark.parametrize("storage", STORAGES)
@pytest.mark.parametrize("keep", [False, True])
def test_bd_run_block_lanczos_reorth(oracle, chain100, keep, storage):
    nsteps, s = 8, 4
    xi = np.random.default_rng(42).standard_normal(300 * s * 2)
    g = gpu_system(chain100, mob_fctr="LANC", nts_mobsam=s, lanc_reorth=True, keep_reuse_noise=keep,
                   mob_storage=storage)
    g.bds_run(nsteps, xi)
    out = g.get_coordinates()
    so = oracle_system(oracle, chain100)
    ref, used, info = oracle.bd_run(so, chain100["crd"], nsteps, xi, mob_fctr=oracle.LANC, nts_mobsam=s,
                                    lanc_reorth=True, keep_reuse_noise=keep)
    assert rel(out, ref) < 1e-9
    assert g.runinfo()["n_mob_updates"] == info.n_mob_updates == 2


def test_bd_run_free_draining_periodic(oracle):
    m = _melt(10, 50, 43)
    nsteps = 10
    xi = np.random.default_rng(43).standard_normal(1500 * nsteps)
    g = gpu_system(m, lhdia=False)
    g.bds_run(nsteps, xi)
    out = g.get_coordinates()
    so = oracle_system(oracle, m, "VER")
    ref, _, _ = oracle.bd_run(so, m["crd"], nsteps, xi, lhdia=False)
    assert rel(out, ref) < TOL


@pytest.mark.parametrize("case", ["chol_c1", "lanczos_s4", "periodic_ver"])
def test_bd_run_predictor_corrector(oracle, chain100, case):
    """Predictor-corrector step (not in the reference) against the oracle's
    restatement of it: same injected xi, FP64, 1e-10 relative."""
    _say(f"case: bd_run predictor-corrector {case}")
    if case == "periodic_ver":
        m = _melt(10, 50, 44)
        nsteps, nx = 10, 1500 * 10
        kw, okw = dict(lhdia=False), dict(lhdia=False)
        so = oracle_system(oracle, m, "VER")
    else:
        m = chain100
        so = oracle_system(oracle, m)
        if case == "chol_c1":
            nsteps, nx = 12, 300 * 12
            kw = dict(mob_fctr="CHOL", nts_mobsam=1)
            okw = dict(mob_fctr=oracle.CHOL, nts_mobsam=1)
        else:
            nsteps, nx = 8, 300 * 4 * 2
            kw = dict(mob_fctr="LANC", nts_mobsam=4, lanc_reorth=True, keep_reuse_noise=True)
            okw = dict(mob_fctr=oracle.LANC, nts_mobsam=4, lanc_reorth=True, keep_reuse_noise=True)
    xi = np.random.default_rng(44).standard_normal(nx)
    g = gpu_system(m, integrator="PC", **kw)
    used = g.bds_run(nsteps, xi)
    out = g.get_coordinates()
    ref, used2, _ = oracle.bd_run(so, m["crd"], nsteps, xi, integrator=oracle.PC, **okw)
    assert used == used2
    assert rel(out, ref) < (1e-9 if case == "lanczos_s4" else TOL)
    # and it is not the EM trajectory
    em, _, _ = oracle.bd_run(so, m["crd"], nsteps, xi, **okw)
    assert rel(em, ref) > 1e3 * rel(out, ref)


@pytest.mark.parametrize("flow", [1, 2, 3])
def test_bd_run_flow_styles(oracle, chain100, flow):
    """Flow fields of calc_drift (shear, planar, uniaxial; bd_solver_m.f90:
    420-441) in EM steps with Cholesky noise, against the oracle."""
    nsteps = 6
    xi = np.random.default_rng(60 + flow).standard_normal(300 * nsteps)
    g = gpu_system(chain100, mob_fctr="CHOL", nts_mobsam=1, flow_style=flow, flow_rate=5.0)
    g.bds_run(nsteps, xi)
    so = oracle_system(oracle, chain100)
    ref, _, _ = oracle.bd_run(so, chain100["crd"], nsteps, xi, mob_fctr=oracle.CHOL, nts_mobsam=1,
                              flow_style=flow, flow_rate=5.0)
    still, _, _ = oracle.bd_run(so, chain100["crd"], nsteps, xi, mob_fctr=oracle.CHOL, nts_mobsam=1)
    assert rel(g.get_coordinates(), ref) < TOL
    assert rel(still, ref) > 1e-6   # the flow moved the chain


def test_marko_siggia_chain(oracle):
    """Bond style 4 (Marko-Siggia, ia_bond_m.f90:334-362, restated as
    written) on a chain: forces and energies, then EM steps with Lanczos
    noise, against the oracle."""
    m = fene_wca_chain(60, 61)
    m = dict(m, bond_styles=[4], bond_params=[[10.0, 5.0]])
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=2, lanc_reorth=True)
    f, e, _, _ = g.ia_calc_forces()
    so = oracle_system(oracle, m)
    fr, en, ierr = oracle.forces(so, m["crd"], oracle.pairs(so, m["crd"], "DIR"))
    assert ierr == 0 and rel(f, fr) < TOL
    assert abs(e[0] - en.energy_bond) <= TOL * abs(en.energy_bond)
    nsteps = 4
    xi = np.random.default_rng(62).standard_normal(180 * 2 * 2)
    g.bds_run(nsteps, xi)
    ref, _, _ = oracle.bd_run(so, m["crd"], nsteps, xi, mob_fctr=oracle.LANC, nts_mobsam=2, lanc_reorth=True)
    assert rel(g.get_coordinates(), ref) < 1e-9


def test_device_gaussian_stream(chain100):
    """The device generator (bd_set_rng) against its CPU restatement
    (tests/philox_ref.py, pinned by the Philox-4x32-10 known-answer
    vectors): the recorded Brownian displacements of a free-draining run
    with dt = 1/2 are the stream itself, sqrt(2 dt) = 1, step by step
    (n values per step, the position kept even)."""
    from philox_ref import gaussian_stream
    from brownpak_amd import BdStep
    nb, nsteps = 7, 3
    n = 3 * nb
    for seed, off in ((12345, 0), (2 ** 40 + 17, 1000)):
        h = BdStep(nb, lhdia=False, tim_stp=0.5)
        h.set_box(0, (0, 0, 0))
        h.ia_setup([], [], None, lvdw=False)
        h.set_coordinates(np.zeros((nb, 3)))
        h.set_rng(seed, off)
        rec = h.run_record(nsteps)
        h.close()
        for q in range(nsteps):
            want = gaussian_stream(seed, off + q * (n + (n & 1)), n)
            assert np.allclose(rec[q].ravel(), want, rtol=1e-13, atol=1e-13), (seed, q)


def test_bond_overstretch_reports_ierr(chain100):
    from brownpak_amd import BdError
    m = dict(chain100)
    crd = chain100["crd"].copy()
    crd[50:] += np.array([5.0, 0.0, 0.0])   # bond 50-51 beyond rmax = 3
    m["crd"] = crd
    g = gpu_system(m, mob_fctr="CHOL")
    with pytest.raises(BdError) as ei:
        g.ia_calc_forces()
    assert ei.value.code == 1
    with pytest.raises(BdError) as ei:
        g.bds_run(3, np.zeros(900))
    assert ei.value.code == 1
    assert g.runinfo()["nts"] == 0


def test_rgsq(oracle, chain100):
    g = gpu_system(chain100, mob_fctr="CHOL")
    assert abs(g.rgsq() - oracle.rgsq(chain100["crd"])) < 1e-12 * oracle.rgsq(chain100["crd"])


def test_device_rng_statistics(chain100):
    g = gpu_system(chain100, lhdia=False, tim_stp=1e-3)
    g.set_rng(1234)
    # pure noise: zero forces by using no bonds and no vdw
    from brownpak_amd import BdStep
    h = BdStep(20000, lhdia=False, tim_stp=0.5)
    h.set_box(0, (0, 0, 0))
    h.ia_setup([], [], None, lvdw=False)
    h.set_coordinates(np.zeros((20000, 3)))
    h.set_rng(99)
    h.bds_run(1)
    d = h.get_coordinates().ravel()   # sqrt(2 dt) xi = xi for dt = 0.5, then COM removed
    assert abs(d.mean()) < 1e-12
    assert abs(d.std() - 1.0) < 0.01
    assert abs(np.mean(d ** 4) / np.mean(d ** 2) ** 2 - 3.0) < 0.1
