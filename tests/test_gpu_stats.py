"""Device run statistics (bd_stats_enable / bd_stats_get) against the
oracle's restatement of stats_accumulate, compute_ic0 and compute_ic1
(oracle/bdstats.py), step by step, for an isolated chain with
hydrodynamic interactions and for periodic chains.
Tolerance: 1e-10 relative on every accumulator (FP64; the device sums in
another order than the reference's loops)."""
import numpy as np
import pytest

from helpers import fene_wca_chain, gpu_system, oracle_system
from brownpak_amd.configs import chain_bonds, periodic_melt
from oracle import bdstats as S

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _close(a, b, scale=None):
    a, b = np.asarray(a, float), np.asarray(b, float)
    s = np.max(np.abs(b)) if scale is None else scale
    return np.max(np.abs(a - b)) <= TOL * max(s, 1e-300)


def _check(got, st):
    assert _close(got["bndlen"], st.bndlen[0]) and _close(got["bndlen_min"], st.bndlen[1])
    assert _close(got["bndlen_max"], st.bndlen[2])
    for k in S.ENERGIES:
        assert _close(got[k], st.acc[k], scale=max(abs(st.acc["energy_tot"]), 1.0)), k
    assert _close(got["stress"], st.stress)
    assert _close(got["rgsq"], st.rgsq) and _close(got["reedsq"], st.reedsq)
    assert _close(got["span"], st.span)
    assert _close(got["reev"], st.reev)


def _run_and_compare(oracle, m, g, st, nsteps, n_noise, structure, seed):
    so = oracle_system(oracle, m)
    rng = np.random.default_rng(seed)
    crd = g.get_coordinates()
    if m["imcon"] == 0:
        crd = crd - crd.mean(axis=0)   # the run's first call centres the molecule
    for step in range(nsteps):
        tab = oracle.pairs(so, crd, "DIR")
        _, en, ierr = oracle.forces(so, crd, tab)
        assert ierr == 0
        g.bds_run(1, rng.standard_normal(n_noise))
        crd = g.get_coordinates()
        st.accumulate({"energy_bond": en.energy_bond, "energy_vdw": en.energy_vdw,
                       "energy_tot": en.energy_tot},
                      np.array(en.stress).reshape(3, 3, order="F"),
                      (en.bndlen, en.bndlen_min, en.bndlen_max), structure(crd))
        print(f"  step {step + 1}: rgsq {st.rgsq:.6e} energy_tot {st.acc['energy_tot']:.6e}", flush=True)
        _check(g.stats(), st)


def test_stats_isolated_chain(oracle):
    m = fene_wca_chain(100, 5)
    g = gpu_system(m, mob_fctr="CHOL", nts_mobsam=1, tim_stp=1e-3)
    g.stats_init((3, 2))
    st = S.Stats(3, 2)
    _run_and_compare(oracle, m, g, st, 7, 300, S.compute_ic0, 11)
    # molecule_com is the accumulated centring shift
    assert np.array_equal(g.stats()["molecule_com"], g.runinfo()["molecule_com"])


def test_stats_periodic_melt(oracle):
    """10 harmonic chains of 50 beads in a periodic box (compute_ic1), plus
    a single-bead molecule row, which compute_ic1 does not count."""
    nch, nper = 10, 50
    crd, L = periodic_melt(nch, nper, 0.3, 43)
    bonds = np.concatenate([chain_bonds(nper, offset=c * nper) for c in range(nch)])
    m = dict(crd=crd, bonds=bonds, bond_styles=[1], bond_params=[[25.0, 2.0]], vdw_styles=[1],
             vdw_params=[[1.0, 2.0, 5.0]], natom_types=1, excluded_atoms=1, rcutoff=5.0, tskin=0.0,
             mth_ptgen="DIR", imcon=1, box=(L, L, L))
    mol = np.array([(nper, 1 + c * nper) for c in range(nch)] + [(1, 1)])
    g = gpu_system(m, lhdia=False, tim_stp=1e-3)
    g.stats_init((2, 4), molecules=mol)
    st = S.Stats(2, 4)
    _run_and_compare(oracle, m, g, st, 6, 3 * nch * nper, lambda c: S.compute_ic1(c, m["box"], mol), 12)


def test_stats_do_not_change_the_run_and_chunks_do_not_matter():
    """The energies the statistics need are computed in the same force
    launch; coordinates stay bit-identical with statistics on, and a run
    split into calls gives the statistics of one call."""
    m = fene_wca_chain(100, 9)
    xi = np.random.default_rng(4).standard_normal(300 * 4 * 3)
    out = []
    for mode in ("off", "one", "split"):
        g = gpu_system(m, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=40, lanc_tol=1e-6)
        if mode != "off":
            g.stats_init((4, 3))
        if mode == "split":
            used = 0
            for k in (5, 2, 5):
                used += g.bds_run(k, xi[used:])
        else:
            g.bds_run(12, xi)
        out.append((g.get_coordinates(), None if mode == "off" else g.stats()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[1][0], out[2][0])
    for k, v in out[1][1].items():
        assert np.array_equal(v, out[2][1][k]), k


def test_rng_state_round_trip():
    """bd_get_rng gives the generator's whole state: a new context set to
    it continues the stream."""
    m = fene_wca_chain(50, 3)
    a = gpu_system(m, mob_fctr="LANC", nts_mobsam=2, lanc_mxitr=30, lanc_tol=1e-6)
    a.set_rng(2024, 0)
    a.bds_run(4)   # two whole mobility blocks
    sd, off = a.get_rng()
    assert sd == 2024 and off == 2 * (150 * 2)
    b = gpu_system(dict(m, crd=a.get_coordinates()), mob_fctr="LANC", nts_mobsam=2, lanc_mxitr=30,
                   lanc_tol=1e-6)
    b.set_rng(sd, off)
    a.reset_run()  # a fresh bds_run in both: centring, then a new block
    a.bds_run(4)
    b.bds_run(4)
    assert np.array_equal(a.get_coordinates(), b.get_coordinates())
