"""The remaining ia_calc_forces terms on the device against the oracle:
angles (ia_angle_m.f90), tethers (ia_tether_m.f90), external fields
(ia_external_m.f90), vdw styles 2-6 with charges (ia_vdw_m.f90), their
energies and stress, the reference's chns-10 model with its angles, runs
with them (tethered runs are not re-centred), and the error paths.
Tolerance: 1e-10 relative (FP64; another summation order)."""
import os

import numpy as np
import pytest

from helpers import fene_wca_chain, gpu_system, oracle_system, read_cfg, rel
from oracle import bdstats as S

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "chns-10.cfg")
TOL = 1e-10


def _terms(nb, rng):
    """Angles along the chain, two harmonic tethers, an x force, charges."""
    angles = [[1, i + 1, i + 2, i + 3] for i in range(nb - 2)]
    tethers = [[1, 1], [1, nb]]
    return dict(angles=angles, angle_styles=[1], angle_params=[[2.0]],
                tethers=tethers, tether_styles=[2], tether_params=[[0.5, 0.2]],
                external_styles=[1], external_params=[[nb // 2, 0.3]],
                charge=rng.uniform(-1.0, 1.0, nb))


def _apply(g, m, t, tpts):
    g.set_angles(t["angles"], t["angle_styles"], t["angle_params"])
    g.set_tethers(t["tethers"], tpts, t["tether_styles"], t["tether_params"])
    g.set_externals(t["external_styles"], t["external_params"])
    g.set_charges(t["charge"])


def _oracle(O, m, t, tpts):
    so = oracle_system(O, m)
    return O.System(len(m["crd"]), m["bonds"], m["bond_styles"], m["bond_params"],
                    atom_types=m.get("atom_types"), vdw_styles=m["vdw_styles"], vdw_params=m["vdw_params"],
                    natom_types=m["natom_types"], lvdw=True, excluded_atoms=m["excluded_atoms"],
                    mth_ptgen=so.s.mth_ptgen, rcutoff=m["rcutoff"], tskin=m["tskin"], imcon=m["imcon"],
                    box=m["box"], charge=t["charge"], angles=t["angles"], angle_styles=t["angle_styles"],
                    angle_params=t["angle_params"], tethers=t["tethers"], tether_points=tpts,
                    tether_styles=t["tether_styles"], tether_params=t["tether_params"],
                    external_styles=t["external_styles"], external_params=t["external_params"])


def _compare_forces(O, g, so, crd):
    f, e, st, bl = g.ia_calc_forces()
    tab = O.pairs(so, crd, "DIR")
    fr, en, ierr = O.forces(so, crd, tab)
    assert ierr == 0
    assert rel(f, fr) < TOL
    ge = g.energies()
    for k in ("energy_bond", "energy_angle", "energy_vdw", "energy_tether", "energy_external", "energy_tot"):
        want = getattr(en, k)
        assert abs(ge[k] - want) <= TOL * max(abs(en.energy_tot), 1.0), (k, ge[k], want)
    assert rel(st.ravel(order="F"), np.array(en.stress)) < TOL
    return en


@pytest.mark.parametrize("vdw", [1, 4, 5])
def test_chain_with_every_term(oracle, vdw):
    rng = np.random.default_rng(7)
    m = fene_wca_chain(120, 21)
    if vdw != 1:   # LJ + (screened) Coulomb with a Coulomb cut-off below rcutoff
        p = [1.0, 2.0, 2.0 * 2 ** (1 / 6), 2.2, 0.6] + ([0.9] if vdw == 4 else [])
        m = dict(m, vdw_styles=[vdw], vdw_params=[p], rcutoff=3.0)
    t = _terms(120, rng)
    tpts = m["crd"][[0, 119]] + [[0.5, 0.0, 0.0], [0.0, -0.5, 0.0]]
    g = gpu_system(m, lhdia=False)
    _apply(g, m, t, tpts)
    so = _oracle(oracle, m, t, tpts)
    en = _compare_forces(oracle, g, so, m["crd"])
    assert en.energy_angle > 0 and en.energy_tether > 0 and en.energy_external != 0


@pytest.mark.parametrize("style", [2, 3, 6])
def test_soft_vdw_styles_periodic(oracle, style):
    from brownpak_amd.configs import chain_bonds, periodic_melt
    crd, L = periodic_melt(6, 40, 0.3, 5)
    bonds = np.concatenate([chain_bonds(40, offset=c * 40) for c in range(6)])
    params = {2: [3.0, 0.4, 2.5], 3: [1.2, 2.5], 6: [20.0, 2.5]}[style]
    m = dict(crd=crd, bonds=bonds, bond_styles=[1], bond_params=[[25.0, 2.0]], vdw_styles=[style],
             vdw_params=[params], natom_types=1, excluded_atoms=1, rcutoff=2.5, tskin=0.0, mth_ptgen="DIR",
             imcon=1, box=(L, L, L))
    g = gpu_system(m, lhdia=False)
    so = oracle_system(oracle, m)
    f, e, st, bl = g.ia_calc_forces()
    fr, en, _ = oracle.forces(so, crd, oracle.pairs(so, crd, "DIR"))
    assert rel(f, fr) < TOL
    assert abs(e[1] - en.energy_vdw) <= TOL * abs(en.energy_vdw)
    assert rel(st.ravel(order="F"), np.array(en.stress)) < TOL


def test_chns10_with_angles(oracle):
    """The reference's chns-10 model (tests/golden/chns-10.cfg) with its 80
    KG angles, periodic box, DIR pairs."""
    m = read_cfg(GOLDEN)
    m.update(natom_types=1, excluded_atoms=1, rcutoff=m["vdw_params"][0][2], tskin=0.0, mth_ptgen="DIR")
    cfg = S.read_config(GOLDEN)
    angles = cfg["angles"].T.copy()
    astyles, apar = cfg["angle_styles"], cfg["angle_params"].T.copy()
    g = gpu_system(m, lhdia=False)
    g.set_angles(angles, astyles, apar)
    so = oracle.System(100, m["bonds"], m["bond_styles"], m["bond_params"], vdw_styles=m["vdw_styles"],
                       vdw_params=m["vdw_params"], natom_types=1, lvdw=True, excluded_atoms=1, mth_ptgen=0,
                       rcutoff=m["rcutoff"], imcon=1, box=m["box"], angles=angles, angle_styles=astyles,
                       angle_params=apar)
    f, e, st, bl = g.ia_calc_forces()
    fr, en, _ = oracle.forces(so, m["crd"], oracle.pairs(so, m["crd"], "DIR"))
    assert rel(f, fr) < TOL
    assert abs(g.energies()["energy_angle"] - en.energy_angle) <= TOL * abs(en.energy_tot)
    assert rel(st.ravel(order="F"), np.array(en.stress)) < TOL


@pytest.mark.parametrize("lhdia", [False, True])
def test_run_with_every_term(oracle, lhdia):
    """bds_run with angles, tethers, an x force and Debye charges: the
    tethered chain is not re-centred (limud false)."""
    rng = np.random.default_rng(9)
    nb, nsteps = 100, 6
    m = fene_wca_chain(nb, 13)
    p = [1.0, 2.0, 2.0 * 2 ** (1 / 6), 0.0, 0.4, 0.8]
    m = dict(m, vdw_styles=[4], vdw_params=[p], rcutoff=3.0)
    t = _terms(nb, rng)
    tpts = m["crd"][[0, nb - 1]] + [[0.3, 0.0, 0.0], [0.0, 0.3, 0.0]]
    xi = rng.standard_normal(3 * nb * nsteps)
    g = gpu_system(m, lhdia=lhdia, mob_fctr="CHOL", tim_stp=1e-3)
    _apply(g, m, t, tpts)
    g.bds_run(nsteps, xi)
    so = _oracle(oracle, m, t, tpts)
    ref, _, _ = oracle.bd_run(so, m["crd"], nsteps, xi, lhdia=lhdia, tim_stp=1e-3)
    out = g.get_coordinates()
    assert rel(out, ref) < TOL
    assert np.linalg.norm(out.mean(axis=0)) > 1e-3


def test_term_errors():
    from brownpak_amd import BdError
    m = fene_wca_chain(30, 2)
    g = gpu_system(m, lhdia=False)
    # a rigid tether the bead already violates
    g.set_tethers([[1, 3]], [m["crd"][2] + [1.0, 0, 0]], [1], [[0.0, 0.5]])
    with pytest.raises(BdError) as ei:
        g.ia_calc_forces()
    assert ei.value.code == 1 and "tether" in str(ei.value)
    g.set_tethers([], [], [], None)
    g.ia_calc_forces()
    # a wall below the top bead (z <= v required, sign -1); the run first
    # centres the molecule, so the wall is placed for centred coordinates
    zc = m["crd"][:, 2] - m["crd"][:, 2].mean()
    g.set_externals([2], [[3, zc.max() - 1.0, -1.0]])
    with pytest.raises(BdError) as ei:
        g.bds_run(1, np.zeros(90))
    assert ei.value.code == 1 and "wall" in str(ei.value)
    with pytest.raises(BdError):
        g.set_externals([1], [[31, 1.0]])   # atom index out of range
    with pytest.raises(BdError):
        g.set_angles([[1, 1, 2, 40]], [1], [[1.0]])


def test_stats_energies_of_every_term(oracle):
    """The statistics take the angle, tether and external energies."""
    rng = np.random.default_rng(4)
    nb = 60
    m = fene_wca_chain(nb, 8)
    t = _terms(nb, rng)
    tpts = m["crd"][[0, nb - 1]] + [[0.3, 0.0, 0.0], [0.0, 0.3, 0.0]]
    g = gpu_system(m, lhdia=False, tim_stp=1e-3)
    _apply(g, m, t, tpts)
    g.stats_init((5, 5))
    so = _oracle(oracle, m, t, tpts)
    st = S.Stats(5, 5)
    crd = m["crd"].copy()
    for step in range(3):
        _, en, _ = oracle.forces(so, crd, oracle.pairs(so, crd, "DIR"))
        g.bds_run(1, rng.standard_normal(3 * nb))
        crd = g.get_coordinates()
        st.accumulate({k: getattr(en, k) for k in S.ENERGIES}, np.array(en.stress).reshape(3, 3, order="F"),
                      (en.bndlen, en.bndlen_min, en.bndlen_max), S.compute_ic0(crd))
    got = g.stats()
    for k in S.ENERGIES:
        assert abs(got[k] - st.acc[k]) <= TOL * max(abs(st.acc["energy_tot"]), 1.0), k
