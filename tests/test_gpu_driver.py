"""The Fortran driver (iso_c_binding host side) runs the reference's inputs
(control.txt keys, .cfg sections) through libbdstep, matches the Python
mirror of the same ABI step for step, writes the reference's revive,
trajectory and statistics files, and continues a revived run exactly."""
import os
import subprocess

import numpy as np
import pytest

from helpers import fene_wca_chain, fortran_seed, gpu_system, rel, write_cfg

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "brownpak_amd", "_build", "bdrun")
from oracle import bdstats as S  # noqa: E402

CONTROL = ("lhdia T\nmob_fctr LANC\nnts_mobsam 4\nlanc_mxitr 40\nlanc_tol 1.0D-6\nbdintg EM\n"
           "mth_ptgen DIR\nrcutoff 2.2449241\nexcluded_atoms 0\nlvdw T\ntim_stp 1.0D-3\n")


def _drive(cwd, tag):
    print(f"  bdrun job_tag={tag} in {cwd}", flush=True)
    r = subprocess.run([DRIVER, "fn_control=control.txt", f"job_tag={tag}"], cwd=cwd,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return r.stdout


def test_fortran_driver_matches_python_api(tmp_path):
    m = fene_wca_chain(100, 77)
    write_cfg(tmp_path / "chain.cfg_t1", m)
    # bds_run steps until nts > nts_sim: 16 steps
    (tmp_path / "control.txt").write_text(CONTROL + "nts_sim 15\nnts_log 8\nfn_cfg chain.cfg\n"
                                          "fn_out final.xyz\nunknown_key 5\n")
    out = _drive(tmp_path, "_t1")
    assert "execution time(s)" in out
    lines = (tmp_path / "final.xyz_t1").read_text().splitlines()
    got = np.array([[float(x) for x in ln.split()[1:4]] for ln in lines[2:]])
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=40, lanc_tol=1e-6)
    g.set_rng(fortran_seed("_t1"))
    g.bds_run(8)
    g.bds_run(8)
    ref = g.get_coordinates()
    assert got.shape == ref.shape
    assert rel(got, ref) < 1e-12


def test_driver_files_and_revive(tmp_path):
    """Run A: 8 steps.  Run B: 4 steps, then a revived run of 4 more.  The
    revive, trajectory and statistics files of A and B are byte-identical,
    and equal the oracle's writers applied to the Python mirror's run."""
    m = fene_wca_chain(100, 31)
    files = ("stats.txt_r", "traj.bin_r", "revive.bin_r")
    ctl = CONTROL + ("stats_binsize 2 2\nnts_samp 2\nnts_dump 4\nnts_log 4\nwrite_traj T\n"
                     "fn_cfg chain.cfg\nfn_stats stats.txt\nfn_traj traj.bin\nfn_revive revive.bin\n")
    da, db = tmp_path / "a", tmp_path / "b"
    for d in (da, db):
        d.mkdir()
        write_cfg(d / "chain.cfg_r", m)
    (da / "control.txt").write_text(ctl + "nts_sim 7\n")
    _drive(da, "_r")
    (db / "control.txt").write_text(ctl + "nts_sim 3\n")
    _drive(db, "_r")
    (db / "control.txt").write_text(ctl + "nts_sim 7\nlrevive T\n")
    out = _drive(db, "_r")
    assert "read revive file" in out
    for f in files:
        assert (da / f).read_bytes() == (db / f).read_bytes(), f

    # the same run through the Python mirror, written by the oracle's writers
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=40, lanc_tol=1e-6)
    g.set_rng(fortran_seed("_r"))
    g.stats_init((2, 2))
    cfg = S.read_config(da / "chain.cfg_r")
    groups = S.stats_columns(num_bonds=cfg["bonds"].shape[1], lvdw=True, imcon=0)
    text = S.stats_header(groups)
    traj = S.traj_header(len(m["crd"]))
    for nts in (2, 4, 6, 8):
        g.bds_run(2)
        v = g.stats()
        st = S.Stats(2, 2)
        st.bndlen = (v["bndlen"], v["bndlen_min"], v["bndlen_max"])
        st.acc = {k: v[k] for k in S.ENERGIES}
        st.stress, st.rgsq, st.reedsq, st.span, st.reev = v["stress"], v["rgsq"], v["reedsq"], v["span"], v["reev"]
        text += S.stats_line(nts, st, groups, molecule_com=v["molecule_com"])
        traj += S.traj_frame(nts, g.get_coordinates())
    assert (da / "stats.txt_r").read_text() == text
    assert (da / "traj.bin_r").read_bytes() == traj
    cfg["coordinates"] = g.get_coordinates().T.copy()
    dump = S.dump_bytes(8, 0, cfg["basis"], cfg, molecule_com=g.runinfo()["molecule_com"], rng=g.get_rng())
    assert (da / "revive.bin_r").read_bytes() == dump
    nts, _, _, atc, com, rng = S.read_dump(dump)
    assert nts == 8 and rng == g.get_rng()


@pytest.mark.parametrize("mode", ["lanc_s4", "chol", "free"])
def test_driver_percall_composition_matches_bd_run(tmp_path, mode):
    """integrate_em composed in Fortran from the per-call-site entries
    (bd_rpy_assemble, bd_brownian, bd_forces, bd_mob_symv; update, centring
    and flow on the host) takes the same steps as bd_run on the same
    injected Gaussian stream."""
    m = fene_wca_chain(100, 12)
    # Lanczos with reorthogonalisation: without it the converged B.xi moves by
    # ~1e-7 under a rounding-level change of the coordinates (the centre of
    # mass is summed in another order on the host)
    ctl = {"lanc_s4": CONTROL + "lanc_reorth T\n",
           "chol": CONTROL.replace("mob_fctr LANC", "mob_fctr CHOL").replace("nts_mobsam 4", "nts_mobsam 1"),
           "free": CONTROL.replace("lhdia T", "lhdia F")}[mode]
    ctl += "nts_sim 11\nnts_log 4\nfn_cfg chain.cfg\nfn_xi xi.bin\nfn_out final.xyz\n"
    xi = np.random.default_rng(21).standard_normal(4000)
    out = {}
    for path in ("run", "percall"):
        d = tmp_path / path
        d.mkdir()
        write_cfg(d / "chain.cfg_p", m)
        xi.tofile(d / "xi.bin")
        (d / "control.txt").write_text(ctl + f"device_path {path}\n")
        _drive(d, "_p")
        lines = (d / "final.xyz_p").read_text().splitlines()
        out[path] = np.array([[float(x) for x in ln.split()[1:4]] for ln in lines[2:]])
    print(f"  {mode}: rel {rel(out['percall'], out['run']):.2e}")
    assert rel(out["percall"], out["run"]) < 1e-12
