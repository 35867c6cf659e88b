"""Pins of the statistics and file-format restatement (oracle/bdstats.py)
and the driver's file module (brownpak_amd/fortran/bdrun_io.f90) on the CPU.

The statistics are pinned by closed forms (a straight rod), a hand-worked
case of compute_ic1's imaging loop, and brute-force bin means.  The file
formats are pinned by item sizes worked out from config_io_m.f90,
trajectory_m.f90 and stats_m.f90, by ES14.6 strings printed by the Fortran
runtime of this image, and by byte equality between the Fortran writer and
the Python restatement.  The reference holds no written dump, trajectory
or statistics file, so the order of the items rests on the reading of
config_io_m.f90:176-243 and is shared by both writers."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
from oracle import bdstats as S  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden", "chns-10.cfg")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


# ------------------------------------------------------------ statistics

@pytest.mark.parametrize("n,b", [(2, 1.0), (7, 0.97), (64, 1.3)])
def test_ic0_straight_rod(n, b):
    d = np.array([2.0, -1.0, 0.5]) / np.linalg.norm([2.0, -1.0, 0.5])
    crd = np.array([3.0, 1.0, -2.0]) + b * np.arange(n)[:, None] * d
    st = S.compute_ic0(crd)
    assert np.isclose(st["rgsq"], b * b * (n * n - 1) / 12.0, rtol=1e-13)
    assert np.isclose(st["reedsq"], ((n - 1) * b) ** 2, rtol=1e-13)
    assert np.allclose(st["span"], (n - 1) * b * np.abs(d), rtol=1e-13)
    assert np.allclose(st["reev"], (n - 1) * b * d, rtol=1e-13)


def test_ic1_imaging_loop_hand_worked():
    """compute_ic1 zeroes bead 1 before shifting the others
    (stats_m.f90:383-386): a rod from (10,10,10) along x in a box of 100
    gives molbuf = [0, (11,10,10), (12,10,10), (13,10,10), (14,10,10)]:
    span (14,10,10), reev (14,10,10), reedsq 396, com (10,8,8),
    rgsq (228 + 9 + 12 + 17 + 24) / 5 = 58."""
    crd = np.array([[10.0 + k, 10.0, 10.0] for k in range(5)])
    st = S.compute_ic1(crd, (100.0, 100.0, 100.0), [(5, 1)])
    assert np.allclose(st["span"], [14, 10, 10])
    assert np.allclose(st["reev"], [14, 10, 10])
    assert st["reedsq"] == 396.0
    assert np.isclose(st["rgsq"], 58.0, rtol=1e-15)
    # two molecules average; a single-bead molecule is not counted
    crd2 = np.vstack([crd, crd + [0.0, 0.0, 1.0], [[1.0, 2.0, 3.0]]])
    st2 = S.compute_ic1(crd2, (100.0, 100.0, 100.0), [(5, 1), (5, 6), (1, 11)])
    # second molecule: molbuf = [0, (11,10,11), ..., (14,10,11)]
    assert np.allclose(st2["span"], [14, 10, 10.5])
    assert np.isclose(st2["reedsq"], (396.0 + 196 + 100 + 121) / 2)
    # imaging about the origin: bead at x = 6 in a box of 10 moves to -4
    crd3 = np.array([[1.0, 0.0, 0.0], [6.0, 0.0, 0.0], [6.5, 0.0, 0.0]])
    st3 = S.compute_ic1(crd3, (10.0, 10.0, 10.0), [(3, 1)])
    assert np.allclose(st3["reev"], [-3.5, 0, 0])
    assert np.allclose(st3["span"], [4.0, 0, 0])


def test_nint_is_half_away_from_zero():
    x = np.array([0.5, -0.5, 1.5, 2.5, -2.5, 0.49999999999999994, 2.4999999999999996, 3.0])
    assert S.nint(x).tolist() == [1, -1, 2, 3, -3, 0, 2, 3]


@pytest.mark.parametrize("bnsiz,bnsiz_s", [(1, 1), (3, 2), (4, 7), (0, 0)])
def test_accumulators_are_bin_means(bnsiz, bnsiz_s):
    """stats_accumulate's running mean over a bin equals the arithmetic mean
    of the steps since the bin started; a bin restarts on the step after
    bnsiz steps (binsize 0: never)."""
    rng = np.random.default_rng(1)
    T = 11
    x = rng.standard_normal((T, 7))
    sig = rng.standard_normal((T, 3, 3))
    geo = rng.random((T, 5))
    st = S.Stats(bnsiz, bnsiz_s)
    for t in range(T):
        e = dict(zip(S.ENERGIES, x[t]))
        st.accumulate(e, sig[t], (1, 0.5, 2), {"rgsq": geo[t, 0], "reedsq": geo[t, 1],
                                              "span": geo[t, 2:5], "reev": geo[t, 2:5]})
        b0 = 0 if bnsiz == 0 else (t // bnsiz) * bnsiz
        s0 = 0 if bnsiz_s == 0 else (t // bnsiz_s) * bnsiz_s
        assert np.allclose([st.acc[k] for k in S.ENERGIES], x[b0:t + 1].mean(axis=0), rtol=1e-13, atol=1e-15)
        assert np.allclose(st.stress, sig[s0:t + 1].mean(axis=0), rtol=1e-13, atol=1e-15)
        assert np.isclose(st.rgsq, geo[b0:t + 1, 0].mean(), rtol=1e-13)
        assert np.allclose(st.span, geo[b0:t + 1, 2:5].mean(axis=0), rtol=1e-13)
        assert np.allclose(st.reev, geo[t, 2:5])


def test_es14_matches_fortran_runtime():
    """Strings printed with (1x,es14.6) by this image's Fortran runtime."""
    want = {1.0: "  1.000000E+00", -2.5e-3: " -2.500000E-03", 1.234567891e123: "  1.234568+123",
            0.0: "  0.000000E+00", -1e-200: " -1.000000-200", 9.9999995: "  9.999999E+00"}
    for x, s in want.items():
        assert S.es14(x) == s, x


def test_stats_header_layout():
    groups = S.stats_columns(num_bonds=9, lvdw=True, imcon=0)
    cols = [c for g, _ in groups for c in g]
    h = S.stats_header(groups)
    assert h.startswith("         nts     bndlen_avg")
    # 15 characters a column, one blank after each unlimited-format statement
    assert len(h) == 12 + 15 * len(cols) + sum(u for _, u in groups) + 1
    assert h.split() == ["nts"] + cols
    assert cols[-6:] == ["reevx", "reevy", "reevz", "comx", "comy", "comz"]
    assert "reevx" not in [c for g, _ in S.stats_columns(num_bonds=9, imcon=1) for c in g]


# ------------------------------------------------------------ files

def _cfg_imcon0(path):
    """A two-type isolated chain with every optional section the dump
    writes when present (charges, molecules, flow field)."""
    path.write_text(
        "title\nSIMBOX 0\n0 0 0\n0 0 0\n0 0 0\n\nATOM_TYPES 2\nA 1 1.5\nBB 2 2.0\n\n"
        "ATOMS 4\n1 0.5 0.0 0.0 0.0\n2 -0.5 1.0 0.1 0.0  # comment\n1 0.0 2.0 0.0 0.3\n2 1.0 3.0 0.2 0.1\n\n"
        "BOND_TYPES 1\n1 2 30.0 1.5\n\nBONDS 3\n1 1 2\n1 2 3\n1 3 4\n\n"
        "MOLECULE_TYPES 1\nCHN 1\n\nMOLECULES 1\n1 4 1 3 1 0 0 0 0\n\n"
        "VDW 3\n1 1 1 3 1.0 1.0 1.12246\n1 2 1 3 1.0 1.0 1.12246\n2 2 1 3 1.0 1.0 1.12246\n\n"
        "FLOW_FIELD\n1 1 0.25\n")


def test_dump_item_sizes():
    c = S.read_config(GOLDEN)
    b = S.dump_bytes(7, c["imcon"], c["basis"], c)
    na, nb_, nan, nm = 100, 90, 80, 10
    size = (8 + 4 + 72                      # nts, imcon, basis
            + 4 + 1 * (8 + 4 + 8)           # atom types: name, style, mass
            + 4 + na * (4 + 24 + 8)         # atoms, coordinates, charge
            + 4 + 1 * (4 + 96) + 4 + nb_ * 12          # bond types, bonds
            + 4 + 1 * (4 + 96) + 4 + nan * 16          # angle types, angles
            + 4 + 4                          # dihedral types, dihedrals: none
            + 4                              # branches: none
            + 4 + 1 * (8 + 4) + 4 + nm * 36            # molecule types, molecules
            + 4 + 4                          # tether types, tethers: none
            + 4 + 1 * (8 + 4 + 96)          # vdw types: pairs, styles, params
            + 4                              # externals: none
            + 4)                             # flow style 0; no molecule_com (imcon 1)
    assert len(b) == size
    assert len(S.dump_bytes(7, c["imcon"], c["basis"], c, rng=(1, 2))) == size + 24
    nts, imcon, basis, atc, com, rng = S.read_dump(S.dump_bytes(7, 1, c["basis"], c, rng=(5, 6)))
    assert (nts, imcon, rng) == (7, 1, (5, 6)) and com is None
    assert np.array_equal(atc["coordinates"], c["coordinates"])
    assert np.array_equal(atc["angles"], c["angles"])


def test_traj_layout():
    crd = np.arange(12.0).reshape(4, 3)
    h = S.traj_header(4)
    assert struct.unpack("<7i", h) == (28, 8 + 96, 4, 1, 0, 0, 0)
    data = h + S.traj_frame(3, crd) + S.traj_frame(9, 2 * crd)
    na, fr = S.read_traj(data)
    assert na == 4 and [f[0] for f in fr] == [3, 9]
    assert np.array_equal(fr[1][1], 2 * crd)


@pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not in this image")
@pytest.mark.parametrize("case", ["chns10", "imcon0"])
def test_fortran_file_module_matches_restatement(tmp_path, case):
    exe = tmp_path / "io_check"
    fdir = os.path.join(ROOT, "brownpak_amd", "fortran")
    r = subprocess.run([FLANG, "-O1", "-module-dir", str(tmp_path), os.path.join(fdir, "bdrun_io.f90"),
                        os.path.join(ROOT, "tests", "fortran", "io_check.f90"), "-o", str(exe)],
                       capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 0, r.stderr[-3000:]
    if case == "chns10":
        cfg = tmp_path / "in.cfg"
        shutil.copy(GOLDEN, cfg)
    else:
        cfg = tmp_path / "in.cfg"
        _cfg_imcon0(cfg)
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([str(exe), str(cfg), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    c = S.read_config(cfg)
    com = np.array([0.5, -0.25, 1e-3])
    want = S.dump_bytes(1234, c["imcon"], c["basis"], c, molecule_com=com, rng=(42, 1000))
    got = (out / "a.dump").read_bytes()
    assert got == want
    assert (out / "b.dump").read_bytes() == want
    crd = c["coordinates"].T
    assert (out / "t.bin").read_bytes() == (S.traj_header(len(crd)) + S.traj_frame(2, crd)
                                           + S.traj_frame(4, 2 * crd) + S.traj_frame(6, 3 * crd))
    s = np.arange(1, 31) * 1.5 - 7.0
    s[4], s[5] = 1.234567891e123, -2.5e-201
    cols = S.stats_columns(num_bonds=c["bonds"].shape[1], num_angles=c.get("angles", np.zeros((4, 0))).shape[1],
                           lvdw=True, imcon=c["imcon"], flow_style=c["flow_style"])

    def line(nts, v):
        st = S.Stats(1, 1)
        st.bndlen = tuple(v[0:3])
        st.acc = dict(zip(S.ENERGIES, v[3:10]))
        st.stress = v[10:19].reshape(3, 3, order="F")
        st.rgsq, st.reedsq, st.span, st.reev = v[19], v[20], v[21:24], v[24:27]
        return S.stats_line(nts, st, cols, molecule_com=v[27:30])

    text = (out / "s.txt").read_text()
    assert text == S.stats_header(cols) + line(10, s) + line(20, s * 1e-3)
