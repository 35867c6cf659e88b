"""ctypes wrapper around the CPU oracle (oracle/bdoracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import
this module; the product (``brownpak_amd``) never does.

Arrays follow the reference's layout: coordinates are (N, 3) C-order here,
which is byte-identical to Fortran's column-major (3, N); the mobility matrix
is returned as an (n, n) array whose memory is column-major (n = 3N).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbdoracle.so")
_lib = None

MXPARAM = 12
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class OrcSystem(C.Structure):
    _fields_ = [
        ("nbeads", C.c_int), ("imcon", C.c_int), ("box", C.c_double * 3),
        ("nbonds", C.c_int), ("bonds", IP),
        ("nbond_types", C.c_int), ("bond_styles", IP), ("bond_params", DP),
        ("natom_types", C.c_int), ("atom_types", IP),
        ("nvdw_types", C.c_int), ("vdw_styles", IP), ("vdw_params", DP),
        ("lvdw", C.c_int), ("excluded_atoms", C.c_int), ("mth_ptgen", C.c_int),
        ("rcutoff", C.c_double), ("tskin", C.c_double),
        ("charge", DP),
        ("nangles", C.c_int), ("angles", IP), ("nangle_types", C.c_int), ("angle_styles", IP),
        ("angle_params", DP),
        ("ntethers", C.c_int), ("tethers", IP), ("tether_points", DP), ("ntether_types", C.c_int),
        ("tether_styles", IP), ("tether_params", DP),
        ("nexternals", C.c_int), ("external_styles", IP), ("external_params", DP),
    ]


class OrcEnergies(C.Structure):
    _fields_ = [
        ("energy_bond", C.c_double), ("energy_vdw", C.c_double), ("energy_tot", C.c_double),
        ("stress", C.c_double * 9),
        ("bndlen", C.c_double), ("bndlen_min", C.c_double), ("bndlen_max", C.c_double),
        ("energy_angle", C.c_double), ("energy_dihedral", C.c_double), ("energy_tether", C.c_double),
        ("energy_external", C.c_double),
    ]


class OrcBdopts(C.Structure):
    _fields_ = [
        ("lhdia", C.c_int), ("mob_fctr", C.c_int), ("nts_mobsam", C.c_int),
        ("lanc_mxitr", C.c_int), ("lanc_tol", C.c_double), ("tim_stp", C.c_double),
        ("keep_reuse_noise", C.c_int), ("flow_style", C.c_int), ("flow_rate", C.c_double),
        ("lanc_reorth", C.c_int), ("integrator", C.c_int),
        ("se_nlmxitr", C.c_int), ("se_kdmax", C.c_int), ("se_ftol", C.c_double), ("se_stptol", C.c_double),
    ]


class OrcRuninfo(C.Structure):
    _fields_ = [
        ("nts", C.c_longlong), ("molecule_com", C.c_double * 3),
        ("last_lanczos_iters", C.c_int), ("n_not_converged", C.c_int), ("ierr", C.c_int),
        ("total_lanczos_iters", C.c_int), ("n_mob_updates", C.c_int),
        ("se_nfe", C.c_int), ("se_nli", C.c_int), ("se_nni", C.c_int), ("se_nbt", C.c_int),
        ("se_iterm", C.c_int),
    ]


NIT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double))


class OrcVerlet(C.Structure):
    _fields_ = [("first_call", C.c_int), ("crd_save", DP),
                ("rskin_sq", C.c_double), ("tskin_sq", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(_LIB_PATH)
        L = _lib
        L.orc_rpy.argtypes = [C.c_int, DP, DP]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_symv_lower.argtypes = [C.c_int, C.c_double, DP, DP, C.c_double, DP]
        L.orc_chol_bdw.argtypes = [C.c_int, C.c_int, DP, DP]
        L.orc_chol_bdw.restype = C.c_int
        L.orc_lanczos.argtypes = [C.c_int, DP, DP, C.c_int, C.c_double, IP, IP]
        L.orc_block_lanczos.argtypes = [C.c_int, C.c_int, DP, DP, C.c_int, C.c_double, IP, IP]
        L.orc_brn_calc_bdw.argtypes = [C.c_int, C.c_int, C.c_int, DP, DP, C.c_int,
                                       C.c_double, IP, IP]
        L.orc_brn_calc_bdw_ex.argtypes = [C.c_int, C.c_int, C.c_int, DP, DP, C.c_int,
                                          C.c_double, C.c_int, IP, IP]
        L.orc_geqp3.argtypes = [C.c_int, C.c_int, DP, IP, DP]
        L.orc_orgqr.argtypes = [C.c_int, C.c_int, DP, DP]
        L.orc_syev.argtypes = [C.c_int, DP, DP]
        L.orc_setup_params.argtypes = [C.POINTER(OrcSystem)]
        L.orc_exclusions.argtypes = [C.c_int, C.c_int, IP, C.c_int, IP, IP, C.c_int]
        L.orc_pairs_n2.argtypes = [C.POINTER(OrcSystem), IP, IP, DP, C.c_double, IP, IP, C.c_int]
        L.orc_pairs_verlet.argtypes = [C.POINTER(OrcSystem), IP, IP, DP, C.POINTER(OrcVerlet),
                                       IP, IP, C.c_int, IP]
        L.orc_pairs_cell_list.argtypes = [C.POINTER(OrcSystem), IP, IP, DP, IP, IP, C.c_int]
        L.orc_forces.argtypes = [C.POINTER(OrcSystem), DP, IP, IP, DP, C.POINTER(OrcEnergies)]
        L.orc_get_image.argtypes = [DP, DP]
        L.orc_wrap_all.argtypes = [DP, C.c_int, DP]
        L.orc_to_origin.argtypes = [C.c_int, DP, DP]
        L.orc_rgsq.argtypes = [C.c_int, DP]
        L.orc_rgsq.restype = C.c_double
        L.orc_bd_run.argtypes = [C.POINTER(OrcSystem), C.POINTER(OrcBdopts), DP, C.c_int, DP,
                                 C.c_longlong, C.POINTER(OrcRuninfo)]
        L.orc_bd_run.restype = C.c_longlong
        L.orc_nitsol.argtypes = [C.c_int, DP, NIT_FN, C.c_void_p, C.c_double, C.c_double, IP, IP]
        L.orc_nitsol.restype = C.c_int
        L.orc_dlaic1.argtypes = [C.c_int, C.c_int, DP, C.c_double, DP, C.c_double, DP, DP, DP]
        L.orc_bd_run_ex.argtypes = [C.POINTER(OrcSystem), C.POINTER(OrcBdopts), DP, C.c_int, DP,
                                    C.c_longlong, DP, C.POINTER(OrcRuninfo)]
        L.orc_bd_run_ex.restype = C.c_longlong
    return _lib


def _dp(a):
    return a.ctypes.data_as(DP)


def _ip(a):
    return a.ctypes.data_as(IP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- mobility
def set_threads(threads):
    """Host threads of the RPY assembly and the lower-triangle symv: 1 (the
    default, used by the tests) is the serial code; more runs them with
    OpenMP (the symv's sums then come in another order), for timing the
    restatement on all host cores."""
    lib().orc_set_threads(int(threads))


def rpy(crd):
    """calc_rpy_tensor: returns the full n x n mobility (symmetric)."""
    crd = _f64(crd)
    nb = crd.shape[0]
    mob = np.empty((3 * nb, 3 * nb), dtype=np.float64)
    lib().orc_rpy(nb, _dp(crd), _dp(mob))
    return mob.T.copy()  # column-major memory -> logical (row, col)


def symv_lower(mob_colmajor_logical, x, alpha=1.0, beta=0.0, y=None):
    a = np.asfortranarray(mob_colmajor_logical, dtype=np.float64)
    n = a.shape[0]
    x = _f64(x)
    y = np.zeros(n) if y is None else _f64(y).copy()
    lib().orc_symv_lower(n, alpha, a.ctypes.data_as(DP), _dp(x), beta, _dp(y))
    return y


# ------------------------------------------------------------ Brownian B.xi
CHOL, LANC, CHOL_LT = 0, 1, 2
EM, PC, SE = 0, 1, 2   # integrator: integrate_em, predictor-corrector, integrate_se


def brn_calc_bdw(method, mob, xi, lanc_mxitr=50, lanc_tol=1e-6, reorth=False):
    """brn_calc_bdw.  mob (n,n) logical, xi (n,s).  Returns (bdw, lconv, iters, ierr)."""
    a = np.asfortranarray(mob, dtype=np.float64).copy(order="F")
    xi = np.asfortranarray(xi, dtype=np.float64)
    if xi.ndim == 1:
        xi = xi[:, None]
    n, s = xi.shape
    b = xi.copy(order="F")
    lconv = C.c_int(0)
    iters = C.c_int(0)
    ierr = lib().orc_brn_calc_bdw_ex(method, n, s, a.ctypes.data_as(DP), b.ctypes.data_as(DP),
                                     lanc_mxitr, lanc_tol, int(bool(reorth)), C.byref(lconv),
                                     C.byref(iters))
    return b, bool(lconv.value), iters.value, ierr


def geqp3(a):
    a = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    m, n = a.shape
    jpvt = np.zeros(n, dtype=np.int32)
    tau = np.zeros(min(m, n))
    lib().orc_geqp3(m, n, a.ctypes.data_as(DP), _ip(jpvt), _dp(tau))
    return a, jpvt, tau


def orgqr(a, tau):
    a = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    m, n = a.shape
    lib().orc_orgqr(m, n, a.ctypes.data_as(DP), _dp(_f64(tau)))
    return a


def syev(a):
    a = np.asfortranarray(a, dtype=np.float64).copy(order="F")
    n = a.shape[0]
    w = np.zeros(n)
    info = lib().orc_syev(n, a.ctypes.data_as(DP), _dp(w))
    return w, a, info


# ----------------------------------------------------------------- system
class System:
    """Holds the topology/parameter arrays alive for an OrcSystem struct.

    bonds: (nb, 3) int [type, i, j], 1-based (memory == Fortran (3, nb)).
    bond_params / vdw_params: (ntypes, 12) (memory == Fortran (12, ntypes)).
    """

    def __init__(self, nbeads, bonds, bond_styles, bond_params, atom_types=None,
                 vdw_styles=(), vdw_params=None, natom_types=1, lvdw=True,
                 excluded_atoms=0, mth_ptgen=0, rcutoff=0.0, tskin=0.0,
                 imcon=0, box=(0.0, 0.0, 0.0), charge=None, angles=None, angle_styles=(),
                 angle_params=None, tethers=None, tether_points=None, tether_styles=(),
                 tether_params=None, external_styles=(), external_params=None):
        self.bonds = _i32(np.asarray(bonds).reshape(-1, 3)) if len(bonds) else np.zeros((0, 3), np.int32)
        self.bond_styles = _i32(bond_styles) if len(bond_styles) else np.zeros(1, np.int32)
        bp = np.zeros((max(len(bond_styles), 1), MXPARAM))
        if bond_params is not None and len(bond_styles):
            bpi = np.atleast_2d(np.asarray(bond_params, dtype=np.float64))
            bp[:, :bpi.shape[1]] = bpi
        self.bond_params = bp
        self.atom_types = _i32(atom_types if atom_types is not None else np.ones(nbeads, np.int32))
        self.vdw_styles = _i32(vdw_styles) if len(vdw_styles) else np.zeros(1, np.int32)
        vp = np.zeros((max(len(vdw_styles), 1), MXPARAM))
        if vdw_params is not None and len(vdw_styles):
            vpi = np.atleast_2d(np.asarray(vdw_params, dtype=np.float64))
            vp[:, :vpi.shape[1]] = vpi
        self.vdw_params = vp
        s = OrcSystem()
        s.nbeads = nbeads
        s.imcon = imcon
        s.box = (C.c_double * 3)(*box)
        s.nbonds = self.bonds.shape[0]
        s.bonds = _ip(self.bonds)
        s.nbond_types = len(bond_styles)
        s.bond_styles = _ip(self.bond_styles)
        s.bond_params = _dp(self.bond_params)
        s.natom_types = natom_types
        s.atom_types = _ip(self.atom_types)
        s.nvdw_types = len(vdw_styles)
        s.vdw_styles = _ip(self.vdw_styles)
        s.vdw_params = _dp(self.vdw_params)
        s.lvdw = int(bool(lvdw))
        s.excluded_atoms = excluded_atoms
        s.mth_ptgen = mth_ptgen
        s.rcutoff = rcutoff
        s.tskin = tskin
        # angles (n, 4), tethers (n, 2) + points (n, 3), externals: rows of
        # the reference's column-major arrays; parameters (ntypes, <= 12)
        def table(styles, params):
            st = _i32(styles) if len(styles) else np.zeros(1, np.int32)
            p = np.zeros((max(len(styles), 1), MXPARAM))
            if params is not None and len(styles):
                pi = np.atleast_2d(np.asarray(params, dtype=np.float64))
                p[:, :pi.shape[1]] = pi
            return st, p
        self.charge = None if charge is None else _f64(charge)
        s.charge = _dp(self.charge) if self.charge is not None else None
        self.angles = _i32(np.asarray(angles).reshape(-1, 4)) if angles is not None and len(angles) \
            else np.zeros((1, 4), np.int32)
        self.angle_styles, self.angle_params = table(angle_styles, angle_params)
        s.nangles = 0 if angles is None else len(angles)
        s.angles = _ip(self.angles)
        s.nangle_types = len(angle_styles)
        s.angle_styles = _ip(self.angle_styles)
        s.angle_params = _dp(self.angle_params)
        self.tethers = _i32(np.asarray(tethers).reshape(-1, 2)) if tethers is not None and len(tethers) \
            else np.zeros((1, 2), np.int32)
        self.tether_points = _f64(np.asarray(tether_points).reshape(-1, 3)) if tethers is not None and len(tethers) \
            else np.zeros((1, 3))
        self.tether_styles, self.tether_params = table(tether_styles, tether_params)
        s.ntethers = 0 if tethers is None else len(tethers)
        s.tethers = _ip(self.tethers)
        s.tether_points = _dp(self.tether_points)
        s.ntether_types = len(tether_styles)
        s.tether_styles = _ip(self.tether_styles)
        s.tether_params = _dp(self.tether_params)
        self.external_styles, self.external_params = table(external_styles, external_params)
        s.nexternals = len(external_styles)
        s.external_styles = _ip(self.external_styles)
        s.external_params = _dp(self.external_params)
        self.s = s
        lib().orc_setup_params(C.byref(s))

    @property
    def nbeads(self):
        return self.s.nbeads


def exclusions(sysobj):
    nb = sysobj.nbeads
    cap = nb * 64 + 16
    ptr = np.zeros(nb + 1, np.int32)
    col = np.zeros(cap, np.int32)
    m = lib().orc_exclusions(nb, sysobj.bonds.shape[0], _ip(sysobj.bonds),
                             sysobj.s.excluded_atoms, _ip(ptr), _ip(col), cap)
    assert m >= 0
    return ptr, col[:m].copy()


def pairs(sysobj, crd, method="DIR", cap=None, verlet_state=None):
    """Pair table as (row_ptr, col) with 1-based columns (pairtab_m.fpp)."""
    crd = _f64(crd)
    nb = sysobj.nbeads
    cap = cap or max(nb * nb // 2 + 16, 1024)
    ex_ptr, ex_col = exclusions(sysobj)
    ex_col = ex_col if ex_col.size else np.zeros(1, np.int32)
    ptr = np.zeros(nb + 1, np.int32)
    col = np.zeros(cap, np.int32)
    L = lib()
    if method == "DIR":
        np_ = L.orc_pairs_n2(C.byref(sysobj.s), _ip(ex_ptr), _ip(ex_col), _dp(crd),
                             sysobj.s.rcutoff ** 2, _ip(ptr), _ip(col), cap)
    elif method == "VER":
        st = verlet_state
        rebuilt = C.c_int(0)
        np_ = L.orc_pairs_verlet(C.byref(sysobj.s), _ip(ex_ptr), _ip(ex_col), _dp(crd),
                                 C.byref(st.v), _ip(ptr), _ip(col), cap, C.byref(rebuilt))
        if np_ == -2:
            return None
    elif method == "CL":
        np_ = L.orc_pairs_cell_list(C.byref(sysobj.s), _ip(ex_ptr), _ip(ex_col), _dp(crd),
                                    _ip(ptr), _ip(col), cap)
    else:
        raise ValueError(method)
    if np_ < 0:
        raise RuntimeError(f"pair build failed ({np_})")
    return ptr, col[:np_].copy()


class VerletState:
    def __init__(self, sysobj):
        nb = sysobj.nbeads
        self.save = np.zeros(3 * nb)
        self.v = OrcVerlet()
        self.v.first_call = 1
        self.v.crd_save = _dp(self.save)
        self.v.rskin_sq = (sysobj.s.rcutoff + sysobj.s.tskin) ** 2
        self.v.tskin_sq = sysobj.s.tskin ** 2


def pair_set(ptr, col):
    """Set of unordered 1-based pairs (min, max) from a CSR pair table."""
    out = set()
    for i in range(len(ptr) - 1):
        for k in range(ptr[i], ptr[i + 1]):
            j = int(col[k])
            a, b = (i + 1, j) if i + 1 < j else (j, i + 1)
            out.add((a, b))
    return out


def forces(sysobj, crd, table=None):
    """ia_calc_forces over the given pair table.  Returns (frc (N,3), energies)."""
    crd = _f64(crd)
    nb = sysobj.nbeads
    frc = np.zeros((nb, 3))
    en = OrcEnergies()
    if table is None:
        ptr_p = None
        col_p = None
        ierr = lib().orc_forces(C.byref(sysobj.s), _dp(crd), None, None, _dp(frc), C.byref(en))
    else:
        ptr, col = table
        col = col if col.size else np.zeros(1, np.int32)
        ptr_p, col_p = _ip(ptr), _ip(col)
        ierr = lib().orc_forces(C.byref(sysobj.s), _dp(crd), ptr_p, col_p, _dp(frc), C.byref(en))
    return frc, en, ierr


def rgsq(crd):
    crd = _f64(crd)
    return lib().orc_rgsq(crd.shape[0], _dp(crd))


def bd_run(sysobj, crd, nsteps, xi, lhdia=True, mob_fctr=CHOL, nts_mobsam=1,
           lanc_mxitr=50, lanc_tol=1e-6, tim_stp=1e-3, keep_reuse_noise=False,
           flow_style=0, flow_rate=0.0, lanc_reorth=False, integrator=EM,
           se_nlmxitr=0, se_kdmax=0, se_tol=(1e-10, 1e-12)):
    """bds_run for nsteps with the injected Gaussian stream xi; integrator
    EM (integrate_em), PC (predictor-corrector, not in the reference) or SE
    (integrate_se with NITSOL; se_* are the control keys)."""
    crd = _f64(crd).copy()
    xi = _f64(np.ravel(xi))
    opt = OrcBdopts(int(lhdia), mob_fctr, nts_mobsam, lanc_mxitr, lanc_tol, tim_stp,
                    int(keep_reuse_noise), flow_style, flow_rate, int(bool(lanc_reorth)), integrator,
                    se_nlmxitr, se_kdmax, se_tol[0], se_tol[1])
    info = OrcRuninfo()
    used = lib().orc_bd_run(C.byref(sysobj.s), C.byref(opt), _dp(crd), nsteps, _dp(xi),
                            xi.size, C.byref(info))
    return crd, used, info


def nitsol(fn, x0, ftol, stptol, input10=None):
    """NITSOL (restated, oracle/bdnitsol.c) on a Python function f(x) -> F.
    Returns (x, iterm, info[6] = nfe, njve, nrpre, nli, nni, nbt)."""
    x = _f64(x0).copy()
    n = x.size

    def cb(ctx, nn, xp, fp):
        xv = np.ctypeslib.as_array(xp, shape=(nn,)).copy()
        try:
            fv = np.asarray(fn(xv), dtype=np.float64)
        except Exception:
            return 1
        np.ctypeslib.as_array(fp, shape=(nn,))[:] = fv
        return 0

    cfn = NIT_FN(cb)
    inp = np.zeros(10, np.int32) if input10 is None else _i32(input10)
    info = np.zeros(6, np.int32)
    iterm = lib().orc_nitsol(n, _dp(x), cfn, None, ftol, stptol, _ip(inp), _ip(info))
    return x, iterm, info


def dlaic1(job, x, sest, w, gamma):
    """The dlaic1 restatement: (sestpr, s, c)."""
    x, w = _f64(x), _f64(w)
    out = np.zeros(3)
    o = [C.c_double() for _ in range(3)]
    lib().orc_dlaic1(job, x.size, _dp(x), sest, _dp(w), gamma, C.byref(o[0]), C.byref(o[1]), C.byref(o[2]))
    return o[0].value, o[1].value, o[2].value


def bd_run_replay(sysobj, crd, disp, lhdia=True, nts_mobsam=1, tim_stp=1e-3):
    """bds_run with the Brownian term of every step replaced by disp (K, N, 3)."""
    crd = _f64(crd).copy()
    d = _f64(disp)
    nsteps = d.shape[0]
    opt = OrcBdopts(int(lhdia), CHOL, nts_mobsam, 0, 0.0, tim_stp, 0, 0, 0.0, 0)
    info = OrcRuninfo()
    r = lib().orc_bd_run_ex(C.byref(sysobj.s), C.byref(opt), _dp(crd), nsteps, None, 0, _dp(d),
                            C.byref(info))
    if r < 0:
        raise RuntimeError(f"orc_bd_run_ex failed ({r})")
    return crd
