"""The semi-implicit Euler step (integrate_se + NITSOL, bd_solver_m.f90:
242-332) on the device against the oracle's restatement
(oracle/bdnitsol.c).  Both solve x - dt (U(x) + D F(x)) - x0 - w = 0 to
|F| <= se_tol(1); their Newton paths differ only by rounding, so the new
coordinates agree to the solve's tolerance: 1e-10 relative with
se_tol(1) = 1e-12 (the Jacobian is close to the identity)."""
import numpy as np
import pytest

from helpers import fene_wca_chain, gpu_system, oracle_system, rel
from brownpak_amd.configs import chain_bonds, periodic_melt
from oracle import bdstats as S

pytestmark = pytest.mark.gpu
TOL = 1e-10
SE = dict(integrator="SE", se_nlmxitr=50, se_kdmax=20, se_tol=(1e-12, 1e-15))
ORC_SE = dict(se_nlmxitr=50, se_kdmax=20, se_tol=(1e-12, 1e-15))


@pytest.mark.parametrize("case", ["free", "chol", "lanc_s4_matfree", "lanc_s4_tiles"])
def test_se_steps_match_oracle(oracle, case):
    m = fene_wca_chain(100, 41)
    nsteps = 5
    rng = np.random.default_rng(17)
    if case == "free":
        kw, okw, nxi = dict(lhdia=False), dict(lhdia=False), 300 * nsteps
    elif case == "chol":
        kw, okw, nxi = dict(mob_fctr="CHOL"), dict(mob_fctr=oracle.CHOL), 300 * nsteps
    else:
        st = case.split("_")[-1]
        kw = dict(mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=40, lanc_tol=1e-6, lanc_reorth=True, mob_storage=st)
        okw = dict(mob_fctr=oracle.LANC, nts_mobsam=4, lanc_mxitr=40, lanc_tol=1e-6, lanc_reorth=True)
        nxi = 300 * 4 * 2
    xi = rng.standard_normal(nxi)
    g = gpu_system(m, tim_stp=2e-3, **kw, **SE)
    used = g.bds_run(nsteps, xi)
    so = oracle_system(oracle, m)
    ref, oused, info = oracle.bd_run(so, m["crd"], nsteps, xi, tim_stp=2e-3, integrator=oracle.SE, **okw,
                                     **ORC_SE)
    assert info.se_iterm == 0 and used == oused
    out = g.get_coordinates()
    print(f"  {case}: rel {rel(out, ref):.2e}  nni gpu/oracle {g.se_info()[1][4]}/{info.se_nni}", flush=True)
    assert rel(out, ref) < TOL
    last, tot = g.se_info()
    assert tot[4] >= nsteps and abs(int(tot[4]) - info.se_nni) <= nsteps
    assert np.allclose(g.runinfo()["molecule_com"], info.molecule_com, atol=1e-12)


def test_se_periodic_melt(oracle):
    nch, nper = 8, 50
    crd, L = periodic_melt(nch, nper, 0.3, 23)
    bonds = np.concatenate([chain_bonds(nper, offset=c * nper) for c in range(nch)])
    m = dict(crd=crd, bonds=bonds, bond_styles=[1], bond_params=[[25.0, 2.0]], vdw_styles=[1],
             vdw_params=[[1.0, 2.0, 5.0]], natom_types=1, excluded_atoms=1, rcutoff=5.0, tskin=0.0,
             mth_ptgen="DIR", imcon=1, box=(L, L, L))
    xi = np.random.default_rng(3).standard_normal(3 * nch * nper * 4)
    g = gpu_system(m, lhdia=False, tim_stp=1e-3, **SE)
    g.bds_run(4, xi)
    so = oracle_system(oracle, m)
    ref, _, info = oracle.bd_run(so, m["crd"], 4, xi, lhdia=False, tim_stp=1e-3, integrator=oracle.SE, **ORC_SE)
    assert info.se_iterm == 0
    assert rel(g.get_coordinates(), ref) < TOL


def test_se_stats_take_the_solution_energies(oracle):
    """The last force evaluation of the solve is at the solution, so the
    statistics record the energies of the new coordinates."""
    m = fene_wca_chain(60, 2)
    g = gpu_system(m, lhdia=False, tim_stp=2e-3, **SE)
    g.stats_init((1, 1))
    g.bds_run(1, np.random.default_rng(8).standard_normal(180))
    new = g.get_coordinates()
    so = oracle_system(oracle, m)
    _, en, _ = oracle.forces(so, new, oracle.pairs(so, new, "DIR"))
    got = g.stats()
    assert abs(got["energy_tot"] - en.energy_tot) <= 1e-9 * abs(en.energy_tot)
    assert np.isclose(got["rgsq"], S.compute_ic0(new)["rgsq"], rtol=1e-12)


def test_se_reports_nonlinear_failure():
    from brownpak_amd import BdError
    m = fene_wca_chain(60, 2)
    g = gpu_system(m, lhdia=False, tim_stp=2e-3, integrator="SE", se_nlmxitr=1, se_kdmax=2,
                   se_tol=(1e-300, 0.0))
    with pytest.raises(BdError) as ei:
        g.bds_run(1, np.random.default_rng(8).standard_normal(180))
    assert ei.value.code == 4 and "iterm 1" in str(ei.value)
    assert g.runinfo()["nts"] == 0


def test_se_solve_is_deterministic():
    m = fene_wca_chain(100, 5)
    xi = np.random.default_rng(6).standard_normal(300 * 3)
    outs = []
    for _ in range(2):
        g = gpu_system(m, mob_fctr="CHOL", tim_stp=2e-3, **SE)
        g.bds_run(3, xi)
        outs.append((g.get_coordinates(), g.se_info()[1]))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
