"""Pins of the oracle's remaining force terms (oracle/bdoracle.c): angles
(ia_angle_m.f90), tethers (ia_tether_m.f90), external fields
(ia_external_m.f90) and vdw styles 2-6 with charges (ia_vdw_m.f90).

Each term is pinned by closed forms worked by hand, by F = -grad E
(central differences), by the cut-off conditions its setter derives, and
by the virial form of the stress, sum_i r_i (x) F_i."""
import numpy as np
import pytest


def _sys(O, crd, **kw):
    nb = len(crd)
    base = dict(bonds=[], bond_styles=[], bond_params=None, lvdw=False)
    base.update(kw)
    return O.System(nb, base.pop("bonds"), base.pop("bond_styles"), base.pop("bond_params"), **base)


def _fd_check(O, so, crd, table=None, h=1e-6, tol=1e-7, offset=None):
    """F = -grad E + offset (offset: a force the reference applies without
    its energy gradient)."""
    f, en, ierr = O.forces(so, crd, table)
    assert ierr == 0
    g = np.zeros_like(crd)
    for i in range(crd.shape[0]):
        for d in range(3):
            cp, cm = crd.copy(), crd.copy()
            cp[i, d] += h
            cm[i, d] -= h
            ep = O.forces(so, cp, table)[1].energy_tot
            em = O.forces(so, cm, table)[1].energy_tot
            g[i, d] = -(ep - em) / (2 * h)
    if offset is not None:
        g = g + offset
    scale = max(np.max(np.abs(f)), 1.0)
    assert np.max(np.abs(f - g)) < tol * scale, np.max(np.abs(f - g))
    return f, en


def test_angle_closed_forms(oracle):
    k = 2.5
    # right angle: cos theta = 0, E = k; f(i-1) = -k q2hat / |q1|, f(i+1) = k q1hat / |q2|
    crd = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [2.0, 0.5, 0.0]])
    so = _sys(oracle, crd, angles=[[1, 1, 2, 3]], angle_styles=[1], angle_params=[[k]])
    f, en, ierr = oracle.forces(so, crd)
    assert ierr == 0 and np.isclose(en.energy_angle, k) and np.isclose(en.energy_tot, k)
    assert np.allclose(f[0], [0, -k / 2.0, 0]) and np.allclose(f[2], [k / 0.5, 0, 0])
    assert np.allclose(f.sum(axis=0), 0)
    # straight: E = 0, no force
    crd2 = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 0.0], [3.0, 3.0, 0.0]])
    f2, en2, _ = oracle.forces(so, crd2)
    assert abs(en2.energy_angle) < 1e-15 and np.max(np.abs(f2)) < 1e-14
    # style 0 carries topology only
    so0 = _sys(oracle, crd, angles=[[1, 1, 2, 3]], angle_styles=[0], angle_params=[[k]])
    f0, en0, _ = oracle.forces(so0, crd)
    assert en0.energy_angle == 0.0 and not f0.any()


def test_tether_closed_forms(oracle):
    k, r0 = 3.0, 0.5
    crd = np.array([[1.0, 2.0, 3.0], [4.0, 2.0, 3.0]])
    so = _sys(oracle, crd, tethers=[[1, 2]], tether_points=[[1.0, 2.0, 3.0]], tether_styles=[2],
              tether_params=[[k, r0]])
    f, en, ierr = oracle.forces(so, crd)
    assert ierr == 0
    assert np.isclose(en.energy_tether, 0.5 * k * (3.0 - r0) ** 2)
    assert np.allclose(f[1], [-k * (3.0 - r0), 0, 0]) and not f[0].any()
    # the tether stress is (r - tp) (x) F
    assert np.isclose(en.stress[0], 3.0 * f[1, 0])
    # rigid tether: inside |r - r0| < eps no force, outside an error
    rig = _sys(oracle, crd, tethers=[[1, 2]], tether_points=[[1.0, 2.0, 3.0]], tether_styles=[1],
               tether_params=[[3.0, 0.1]])
    f, en, ierr = oracle.forces(rig, crd)
    assert ierr == 0 and not f.any()
    crd_far = crd.copy()
    crd_far[1, 0] += 0.2
    assert oracle.forces(rig, crd_far)[2] == 1


def test_external_closed_forms(oracle):
    crd = np.array([[1.0, 0.0, 0.0], [4.0, 1.0, 0.0], [2.0, -1.0, 5.0]])
    so = _sys(oracle, crd, external_styles=[1], external_params=[[3, 0.75]])
    f, en, ierr = oracle.forces(so, crd)
    assert ierr == 0
    assert np.isclose(en.energy_external, -0.75 * (2.0 - 1.0)) and np.isclose(en.energy_tot, -0.75)
    assert np.array_equal(f, [[0, 0, 0], [0, 0, 0], [0.75, 0, 0]])
    # wall on z: atoms must stay below 4.0 (sign -1) -- atom 3 is above
    wall = _sys(oracle, crd, external_styles=[2], external_params=[[3, 4.0, -1.0]])
    assert oracle.forces(wall, crd)[2] == 1
    wall_ok = _sys(oracle, crd, external_styles=[2], external_params=[[3, -2.0, 1.0]])
    assert oracle.forces(wall_ok, crd)[2] == 0


VDW_CASES = {
    # style: (params, charges)
    2: ([2.0, 0.7, 2.5], None),
    3: ([1.5, 2.2], None),
    4: ([1.0, 1.0, 1.12246, 3.0, 0.8, 1.3], [1.0, -2.0]),
    5: ([1.0, 1.0, 1.12246, 3.0, 0.8], [1.5, 0.5]),
    6: ([25.0, 1.5], None),
}


def _pair(O, style, r, params, charge, rc_list=4.0):
    crd = np.array([[0.0, 0.0, 0.0], [r, 0.0, 0.0]])
    so = _sys(O, crd, vdw_styles=[style], vdw_params=[params], lvdw=True, rcutoff=rc_list,
              charge=charge)
    tab = O.pairs(so, crd, "DIR")
    f, en, ierr = O.forces(so, crd, tab)
    assert ierr == 0
    return f, en


@pytest.mark.parametrize("style", sorted(VDW_CASES))
def test_vdw_styles_gradient_and_cutoffs(oracle, style):
    params, q = VDW_CASES[style]
    rc = params[2] if style in (2, 4, 5) else params[1]
    for r in (0.9, 1.05, 1.4, 2.0):
        crd = np.array([[0.0, 0.1, -0.2], [r, 0.3, 0.1]])
        so = _sys(oracle, crd, vdw_styles=[style], vdw_params=[params], lvdw=True, rcutoff=4.0, charge=q)
        _fd_check(oracle, so, crd, oracle.pairs(so, crd, "DIR"))
    # at the cut-off: the force-shifted styles (2) and DPD (6) vanish in
    # value and force; the cosine (3) has E(rc) = 0 and E(0+) -> 2A
    f, en = _pair(oracle, style, rc * (1 - 1e-12), params, q)
    if style in (2, 6):
        assert abs(en.energy_vdw) < 1e-9 and abs(f[0, 0]) < 1e-9
    if style == 3:
        assert abs(en.energy_vdw) < 1e-9
        assert np.isclose(_pair(oracle, 3, 1e-9, params, q)[1].energy_vdw, 2 * params[0])
    if style == 6:
        assert np.isclose(_pair(oracle, 6, 1e-12, params, q)[1].energy_vdw, 0.5 * params[0] * params[1])


def test_coulomb_closed_forms(oracle):
    """Beyond the LJ cut-off, with no Coulomb cut-off (rcut_coul = 0), the
    pair energy is C q_i q_j / r (style 5) and C q_i q_j exp(-kappa r) / r
    (style 4, Debye-Hueckel); with a Coulomb cut-off it is shifted to 0
    there."""
    C, kap, q = 0.8, 1.3, [1.0, -2.0]
    r = 2.5
    _, en5 = _pair(oracle, 5, r, [1.0, 1.0, 1.12246, 0.0, C], q)
    assert np.isclose(en5.energy_vdw, C * -2.0 / r, rtol=1e-15)
    _, en4 = _pair(oracle, 4, r, [1.0, 1.0, 1.12246, 0.0, C, kap], q)
    assert np.isclose(en4.energy_vdw, C * -2.0 * np.exp(-kap * r) / r, rtol=1e-15)
    rcc = 3.0
    _, e_at = _pair(oracle, 4, rcc * (1 - 1e-13), [1.0, 1.0, 1.12246, rcc, C, kap], q)
    assert abs(e_at.energy_vdw) < 1e-10
    # inside the LJ range the LJ part is added, shifted to 0 at its cut-off
    _, e_in = _pair(oracle, 5, 1.1224, [1.0, 1.0, 1.12246, 0.0, C], [0.0, 1.0])
    assert abs(e_in.energy_vdw) < 1e-3


def test_all_terms_gradient_and_virial(oracle):
    """A bent chain with bonds, angles, LJ + Debye pairs, a harmonic tether
    and an x force: F = -grad E (bar the x force's reference atom),
    energy_tot is the sum of the parts, and
    the stress is the virial sum_i r_i (x) F_i less tp (x) F_tether."""
    rng = np.random.default_rng(3)
    nb = 8
    crd = np.cumsum(rng.normal(size=(nb, 3)) * 0.6 + [1.0, 0, 0], axis=0)
    bonds = [[1, i + 1, i + 2] for i in range(nb - 1)]
    angles = [[1, i + 1, i + 2, i + 3] for i in range(nb - 2)]
    q = rng.uniform(-1, 1, nb)
    tp = crd[4] + [0.3, -0.2, 0.1]
    so = _sys(oracle, crd, bonds=bonds, bond_styles=[1], bond_params=[[30.0, 1.0]],
              angles=angles, angle_styles=[1], angle_params=[[4.0]],
              tethers=[[1, 5]], tether_points=[tp], tether_styles=[2], tether_params=[[7.0, 0.1]],
              external_styles=[1], external_params=[[8, 0.4]],
              vdw_styles=[4], vdw_params=[[1.0, 1.0, 1.12246, 0.0, 0.5, 0.9]], lvdw=True,
              excluded_atoms=1, rcutoff=2.5, charge=q)
    tab = oracle.pairs(so, crd, "DIR")
    # the x force's energy -f (x_8 - x_1) is measured from atom 1, whose
    # gradient term the reference does not apply (ia_external_m.f90:81-86)
    off = np.zeros_like(crd)
    off[0, 0] = 0.4
    f, en = _fd_check(oracle, so, crd, tab, offset=off)
    parts = (en.energy_bond + en.energy_angle + en.energy_dihedral + en.energy_vdw + en.energy_tether
             + en.energy_external)
    assert en.energy_tot == parts
    assert en.energy_angle > 0 and en.energy_tether > 0 and en.energy_vdw != 0
    f_ext = np.zeros_like(f)
    f_ext[7, 0] = 0.4
    f_teth = np.zeros(3)
    so_t = _sys(oracle, crd, tethers=[[1, 5]], tether_points=[tp], tether_styles=[2], tether_params=[[7.0, 0.1]])
    f_teth = oracle.forces(so_t, crd)[0][4]
    fi = f - f_ext   # the external force adds no stress
    vir = np.einsum("ia,ib->ab", crd, fi) - np.outer(tp, f_teth)
    st = np.array(en.stress).reshape(3, 3, order="F")
    assert np.allclose(st, vir, rtol=1e-12, atol=1e-12 * np.max(np.abs(vir)))


def test_tethered_run_is_not_recentred(oracle):
    """limud (bd_solver_m.f90:153-158) is false with tethers: no centring."""
    nb = 6
    crd = np.array([[5.0 + i, 1.0, 0.0] for i in range(nb)])
    so = _sys(oracle, crd, bonds=[[1, i + 1, i + 2] for i in range(nb - 1)], bond_styles=[1],
              bond_params=[[10.0, 1.0]], tethers=[[1, 1]], tether_points=[[4.9, 1.0, 0.0]],
              tether_styles=[2], tether_params=[[5.0, 0.0]])
    out, _, _ = oracle.bd_run(so, crd, 3, np.zeros(3 * nb * 3), lhdia=False, tim_stp=1e-3)
    assert out.mean(axis=0)[0] > 5.0   # the molecule stays where it is
    free = _sys(oracle, crd, bonds=[[1, i + 1, i + 2] for i in range(nb - 1)], bond_styles=[1],
                bond_params=[[10.0, 1.0]])
    out2, _, _ = oracle.bd_run(free, crd, 3, np.zeros(3 * nb * 3), lhdia=False, tim_stp=1e-3)
    assert np.allclose(out2.mean(axis=0), 0, atol=1e-12)
