"""The Fortran driver (iso_c_binding host side) runs the reference's inputs
(control.txt keys, .cfg sections) through libbdstep and matches the Python
mirror of the same ABI step for step."""
import os
import subprocess

import numpy as np
import pytest

from helpers import fene_wca_chain, fortran_seed, gpu_system, rel, write_cfg

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "brownpak_amd", "_build", "bdrun")


def test_fortran_driver_matches_python_api(tmp_path):
    m = fene_wca_chain(100, 77)
    write_cfg(tmp_path / "chain.cfg", m)
    (tmp_path / "control.txt").write_text(
        "lhdia T\nmob_fctr LANC\nnts_mobsam 4\nlanc_mxitr 40\nlanc_tol 1.0D-6\nbdintg EM\n"
        "mth_ptgen DIR\nrcutoff 2.2449241\nexcluded_atoms 0\nlvdw T\ntim_stp 1.0D-3\n"
        "nts_sim 16\nnts_log 8\nfn_cfg chain.cfg\nfn_out final.xyz\nunknown_key 5\n")
    r = subprocess.run([DRIVER, "fn_control=control.txt", "job_tag=_t1"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "execution time(s)" in r.stdout
    lines = (tmp_path / "final.xyz_t1").read_text().splitlines()
    out = np.array([[float(x) for x in ln.split()[1:4]] for ln in lines[2:]])
    # same run through the Python mirror (the rcutoff above is what the
    # driver read; the vdw parameters come from the .cfg)
    g = gpu_system(m, mob_fctr="LANC", nts_mobsam=4, lanc_mxitr=40, lanc_tol=1e-6)
    g.set_rng(fortran_seed("_t1"))
    g.bds_run(8)
    g.bds_run(8)
    ref = g.get_coordinates()
    assert out.shape == ref.shape
    assert rel(out, ref) < 1e-12
