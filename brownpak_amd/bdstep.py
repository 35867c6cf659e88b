"""Host-side mirror of BROWNPAK's BD-step interface over libbdstep.so.

Method names follow the reference routines they replace:

=====================  ==================================================
``BdStep`` method      reference routine (file:line)
=====================  ==================================================
``__init__``           bds_init + brn_init (bd_solver_m.f90:50-104,
                       brown_m.f90:73-118)
``ia_setup``           ia_setup + pt_init (interaction_m.f90:30-71,
                       pairtab_m.fpp:36-98)
``calc_rpy_tensor``    calc_rpy_tensor (bd_solver_m.f90:458-522)
``brn_calc_bdw``       brn_calc_bdw (brown_m.f90:147-181)
``ia_calc_forces``     ia_calc_forces (interaction_m.f90:84-147)
``mob_symv``           dsymv/dsymm('L') (bd_solver_m.f90:446)
``bds_run``            bds_run loop body (bd_solver_m.f90:167-189)
``close``              bds_finish / brn_finish / ia_finish
=====================  ==================================================

Errors are raised as :class:`BdError` carrying the reference-style ``ierr``
code (1 bond overstretch, 2 factorisation failure, 3 Lanczos not
converged, negative values for argument/HIP/state errors).  The library must
be present: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libbdstep.so")

BD_OK, BD_ERR_BOND, BD_ERR_FACTOR, BD_WARN_NOTCONV = 0, 1, 2, 3
BD_ERR_ARG, BD_ERR_HIP, BD_ERR_STATE, BD_ERR_CELL, BD_ERR_CAPACITY = -1, -2, -3, -4, -5
CHOL, LANC, CHOL_LT = 0, 1, 2
FP64, FP32 = 0, 1
PT = {"DIR": 0, "VER": 1, "CL": 2}
MOB_FCTR = {"CHOL": CHOL, "LANC": LANC, "CHOL_LT": CHOL_LT}
INTEGRATOR = {"EM": 0, "PC": 1, "SE": 2}
MOB_STORAGE = {"auto": 0, "tiles": 1, "matfree": 2}
MXPARAM = 12
# work-unit shapes (tile rows, tile columns) of the mobility product, in the
# order of the batch autotuner's timings (BD_TILE_SHAPES in bdstep.h)
TILE_SHAPES = ((2, 4), (1, 4), (4, 4), (2, 2), (2, 5), (2, 6), (2, 8))

DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)

# Every symbol include/bdstep.h declares (checked by the CPU test suite).
EXPORTS = (
    "bd_ctx_create", "bd_ctx_destroy", "bd_last_error", "bd_set_box", "bd_set_topology",
    "bd_set_coordinates", "bd_get_coordinates", "bd_rpy_assemble", "bd_get_mobility",
    "bd_mob_symv", "bd_brownian", "bd_forces", "bd_pair_table", "bd_run", "bd_set_rng",
    "bd_get_runinfo", "bd_reset_run", "bd_rgsq", "bd_sync", "bd_prof_enable", "bd_prof_get",
    "bd_set_theta", "bd_run_record", "bd_run_replay", "bd_adjoint_rgsq",
    "bd_batch_create", "bd_batch_destroy", "bd_batch_set_model", "bd_batch_set_coordinates",
    "bd_batch_get_coordinates", "bd_batch_set_rng", "bd_batch_run", "bd_batch_info", "bd_batch_prof",
    "bd_batch_tile_shape", "bd_batch_set_tile_shape", "bd_plan_layout", "bd_rpy_apply",
    "bd_rpy_apply_noalloc", "bd_get_rng", "bd_resume_run", "bd_stats_enable", "bd_stats_get",
    "bd_set_angles", "bd_set_tethers", "bd_set_externals", "bd_set_charges", "bd_energies",
    "bd_se_info",
)

# bd_stats_get's output, in stats_write's column order (stats_m.f90:235-290)
STATS_FIELDS = (
    ("bndlen", 1), ("bndlen_min", 1), ("bndlen_max", 1), ("energy_bond", 1), ("energy_angle", 1),
    ("energy_dihedral", 1), ("energy_vdw", 1), ("energy_tether", 1), ("energy_external", 1),
    ("energy_tot", 1), ("stress", 9), ("rgsq", 1), ("reedsq", 1), ("span", 3), ("reev", 3),
    ("molecule_com", 3),
)


class BdOpts(C.Structure):
    _fields_ = [
        ("lhdia", C.c_int), ("mob_fctr", C.c_int), ("nts_mobsam", C.c_int),
        ("lanc_mxitr", C.c_int), ("lanc_tol", C.c_double), ("tim_stp", C.c_double),
        ("precision", C.c_int), ("keep_reuse_noise", C.c_int), ("flow_style", C.c_int),
        ("flow_rate", C.c_double), ("lanc_reorth", C.c_int), ("integrator", C.c_int),
        ("mob_storage", C.c_int), ("se_nlmxitr", C.c_int), ("se_kdmax", C.c_int),
        ("se_ftol", C.c_double), ("se_stptol", C.c_double),
    ]


class BdRuninfo(C.Structure):
    _fields_ = [
        ("nts", C.c_longlong), ("molecule_com", C.c_double * 3),
        ("last_lanczos_iters", C.c_int), ("total_lanczos_iters", C.c_int),
        ("n_mob_updates", C.c_int), ("n_not_converged", C.c_int), ("n_pair_rebuilds", C.c_int),
    ]


class BdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def plan_layout(nbeads: int, cus: int = 256):
    """The single-chain work decomposition of the mobility product (host
    only): ((chr, wsc), units (n, 16) in execution order, rowblk (nt, 8))."""
    L = load_library()
    a, b, nu, nr = C.c_int(0), C.c_int(0), C.c_int(0), C.c_int(0)
    rc = L.bd_plan_layout(int(nbeads), int(cus), C.byref(a), C.byref(b), C.byref(nu), None, 0, C.byref(nr), None, 0)
    if rc != 0:
        raise BdError(rc, L.bd_last_error().decode())
    units = np.zeros((nu.value, 16), dtype=np.int32)
    rowblk = np.zeros((nr.value, 8), dtype=np.int32)
    rc = L.bd_plan_layout(int(nbeads), int(cus), C.byref(a), C.byref(b), C.byref(nu),
                          units.ctypes.data_as(IP), units.size, C.byref(nr), rowblk.ctypes.data_as(IP), rowblk.size)
    if rc != 0:
        raise BdError(rc, L.bd_last_error().decode())
    return (a.value, b.value), units, rowblk


def load_library(path: str = LIB_PATH):
    """dlopen libbdstep.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with brownpak_amd.build.build_library()")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.bd_ctx_create.argtypes = [C.c_int, C.POINTER(BdOpts)]
    L.bd_ctx_create.restype = vp
    L.bd_ctx_destroy.argtypes = [vp]
    L.bd_ctx_destroy.restype = None
    L.bd_last_error.restype = C.c_char_p
    L.bd_set_box.argtypes = [vp, C.c_int, DP]
    L.bd_set_topology.argtypes = [vp, C.c_int, IP, C.c_int, IP, DP, C.c_int, IP, C.c_int, IP, DP,
                                  C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.bd_set_coordinates.argtypes = [vp, DP]
    L.bd_get_coordinates.argtypes = [vp, DP]
    L.bd_rpy_assemble.argtypes = [vp]
    L.bd_get_mobility.argtypes = [vp, DP]
    L.bd_mob_symv.argtypes = [vp, DP, C.c_int, DP]
    L.bd_rpy_apply.argtypes = [C.c_int, vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp]
    L.bd_rpy_apply_noalloc.argtypes = [C.c_int, vp, vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, vp]
    L.bd_brownian.argtypes = [vp, DP, C.c_int, C.c_int, DP, IP, IP]
    L.bd_forces.argtypes = [vp, DP, DP, DP, DP]
    L.bd_pair_table.argtypes = [vp, IP, C.c_int, IP, IP]
    L.bd_run.argtypes = [vp, C.c_int, DP, C.c_longlong, C.POINTER(C.c_longlong)]
    L.bd_set_rng.argtypes = [vp, C.c_uint64, C.c_uint64]
    L.bd_get_rng.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.bd_resume_run.argtypes = [vp, DP]
    L.bd_stats_enable.argtypes = [vp, C.c_int, C.c_int, C.c_int, IP]
    L.bd_stats_get.argtypes = [vp, DP]
    L.bd_set_angles.argtypes = [vp, C.c_int, IP, C.c_int, IP, DP]
    L.bd_set_tethers.argtypes = [vp, C.c_int, IP, DP, C.c_int, IP, DP]
    L.bd_set_externals.argtypes = [vp, C.c_int, IP, DP]
    L.bd_set_charges.argtypes = [vp, DP]
    L.bd_energies.argtypes = [vp, DP]
    L.bd_se_info.argtypes = [vp, IP, C.POINTER(C.c_longlong)]
    L.bd_get_runinfo.argtypes = [vp, C.POINTER(BdRuninfo)]
    L.bd_reset_run.argtypes = [vp]
    L.bd_rgsq.argtypes = [vp, DP]
    L.bd_sync.argtypes = [vp]
    L.bd_prof_enable.argtypes = [vp, C.c_int]
    L.bd_prof_get.argtypes = [vp, DP, C.POINTER(C.c_longlong), DP]
    L.bd_set_theta.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int]
    L.bd_run_record.argtypes = [vp, C.c_int, DP, C.c_longlong, C.POINTER(C.c_longlong), DP]
    L.bd_run_replay.argtypes = [vp, C.c_int, DP]
    L.bd_adjoint_rgsq.argtypes = [vp, C.c_int, DP, C.c_int, C.c_int, DP, DP]
    L.bd_batch_create.argtypes = [C.c_int, IP, C.POINTER(BdOpts)]
    L.bd_batch_create.restype = vp
    L.bd_batch_destroy.argtypes = [vp]
    L.bd_batch_destroy.restype = None
    L.bd_batch_set_model.argtypes = [vp, C.c_int, DP, C.c_int, DP, C.c_double, C.c_int]
    L.bd_batch_set_coordinates.argtypes = [vp, DP]
    L.bd_batch_get_coordinates.argtypes = [vp, DP]
    L.bd_batch_set_rng.argtypes = [vp, C.c_uint64, C.c_uint64]
    L.bd_batch_run.argtypes = [vp, C.c_int, DP, C.c_longlong, C.POINTER(C.c_longlong)]
    L.bd_batch_info.argtypes = [vp, C.POINTER(C.c_longlong), IP, IP, IP]
    L.bd_batch_prof.argtypes = [vp, C.c_int, DP, C.POINTER(C.c_longlong), DP]
    L.bd_batch_tile_shape.argtypes = [vp, IP, IP, DP]
    L.bd_batch_set_tile_shape.argtypes = [vp, C.c_int, C.c_int]
    L.bd_plan_layout.argtypes = [C.c_int, C.c_int, IP, IP, IP, IP, C.c_int, IP, IP, C.c_int]
    for name in EXPORTS:
        if name not in ("bd_ctx_create", "bd_ctx_destroy", "bd_last_error", "bd_batch_create",
                        "bd_batch_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(DP)


def _ip(a):
    return a.ctypes.data_as(IP)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _params(rows, ntypes):
    out = np.zeros((max(ntypes, 1), MXPARAM))
    if ntypes:
        r = np.atleast_2d(np.asarray(rows, dtype=np.float64))
        out[:, : r.shape[1]] = r
    return out


class BdStep:
    """One BD system on the current HIP device (the reference's module state)."""

    def __init__(self, nbeads, *, lhdia=True, mob_fctr="LANC", nts_mobsam=1, lanc_mxitr=50,
                 lanc_tol=1e-6, tim_stp=1e-3, precision="fp64", keep_reuse_noise=False,
                 flow_style=0, flow_rate=0.0, lanc_reorth=False, integrator="EM", mob_storage="auto",
                 se_nlmxitr=0, se_kdmax=0, se_tol=(1e-10, 1e-12)):
        self.L = load_library()
        self.nbeads = int(nbeads)
        self.n = 3 * self.nbeads
        self.s = int(nts_mobsam) if lhdia else 1
        mf = MOB_FCTR[mob_fctr] if isinstance(mob_fctr, str) else int(mob_fctr)
        prec = {"fp64": FP64, "fp32": FP32}[precision] if isinstance(precision, str) else int(precision)
        self.opts = BdOpts(int(bool(lhdia)), mf, int(nts_mobsam), int(lanc_mxitr), float(lanc_tol),
                           float(tim_stp), prec, int(bool(keep_reuse_noise)), int(flow_style),
                           float(flow_rate), int(bool(lanc_reorth)),
                           INTEGRATOR[integrator] if isinstance(integrator, str) else int(integrator),
                           MOB_STORAGE[mob_storage] if isinstance(mob_storage, str) else int(mob_storage),
                           int(se_nlmxitr), int(se_kdmax), float(se_tol[0]), float(se_tol[1]))
        self.h = self.L.bd_ctx_create(self.nbeads, C.byref(self.opts))
        if not self.h:
            raise BdError(BD_ERR_ARG, self.L.bd_last_error().decode())

    # -- plumbing
    def _chk(self, rc, ok=(BD_OK,)):
        if rc not in ok:
            raise BdError(rc, self.L.bd_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.bd_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- setup
    def set_box(self, imcon=0, box=(0.0, 0.0, 0.0)):
        b = _f64(box)
        self._chk(self.L.bd_set_box(self.h, int(imcon), _dp(b)))

    def ia_setup(self, bonds, bond_styles, bond_params, *, atom_types=None, natom_types=1,
                 vdw_styles=(), vdw_params=None, lvdw=True, excluded_atoms=0, mth_ptgen="DIR",
                 rcutoff=0.0, tskin=0.0):
        """Topology and potentials as read from a .cfg file (1-based indices)."""
        self._bonds = _i32(np.asarray(bonds).reshape(-1, 3)) if len(bonds) else np.zeros((1, 3), np.int32)
        nbond = len(bonds)
        self._bstyles = _i32(bond_styles) if len(bond_styles) else np.zeros(1, np.int32)
        self._bparams = _params(bond_params, len(bond_styles))
        at = np.ones(self.nbeads, np.int32) if atom_types is None else _i32(atom_types)
        self._atypes = at
        self._vstyles = _i32(vdw_styles) if len(vdw_styles) else np.zeros(1, np.int32)
        self._vparams = _params(vdw_params, len(vdw_styles))
        mth = PT[mth_ptgen] if isinstance(mth_ptgen, str) else int(mth_ptgen)
        self._chk(self.L.bd_set_topology(
            self.h, nbond, _ip(self._bonds), len(bond_styles), _ip(self._bstyles), _dp(self._bparams),
            int(natom_types), _ip(self._atypes), len(vdw_styles), _ip(self._vstyles),
            _dp(self._vparams), int(bool(lvdw)), int(excluded_atoms), mth, float(rcutoff),
            float(tskin)))

    def set_angles(self, angles, styles, params):
        """angles: rows [type, i-1, i, i+1] 1-based; params (ntypes, <= 12)."""
        a = _i32(np.reshape(angles, (-1, 4))) if len(angles) else np.zeros((1, 4), np.int32)
        st = _i32(styles) if len(styles) else np.zeros(1, np.int32)
        p = _params(params, len(styles))
        self._chk(self.L.bd_set_angles(self.h, len(angles), _ip(a), len(styles), _ip(st), _dp(p)))

    def set_tethers(self, tethers, points, styles, params):
        """tethers: rows [type, atom] 1-based; points (n, 3)."""
        t = _i32(np.reshape(tethers, (-1, 2))) if len(tethers) else np.zeros((1, 2), np.int32)
        pts = _f64(np.reshape(points, (-1, 3))) if len(tethers) else np.zeros((1, 3))
        st = _i32(styles) if len(styles) else np.zeros(1, np.int32)
        p = _params(params, len(styles))
        self._chk(self.L.bd_set_tethers(self.h, len(tethers), _ip(t), _dp(pts), len(styles), _ip(st), _dp(p)))

    def set_externals(self, styles, params):
        st = _i32(styles) if len(styles) else np.zeros(1, np.int32)
        p = _params(params, len(styles))
        self._chk(self.L.bd_set_externals(self.h, len(styles), _ip(st), _dp(p)))

    def set_charges(self, charge):
        if charge is None:
            self._chk(self.L.bd_set_charges(self.h, None))
        else:
            q = _f64(charge)
            assert q.size == self.nbeads
            self._chk(self.L.bd_set_charges(self.h, _dp(q)))

    def se_info(self):
        """NITSOL counts (nfe, njve, nrpre, nli, nni, nbt) of the last
        semi-implicit step and of the run."""
        last = np.zeros(6, np.int32)
        tot = np.zeros(6, np.int64)
        self._chk(self.L.bd_se_info(self.h, _ip(last), tot.ctypes.data_as(C.POINTER(C.c_longlong))))
        return last, tot

    def energies(self):
        """Energies of the last evaluation: dict in stats order."""
        e = np.zeros(7)
        self._chk(self.L.bd_energies(self.h, _dp(e)))
        return dict(zip(("energy_bond", "energy_angle", "energy_dihedral", "energy_vdw", "energy_tether",
                         "energy_external", "energy_tot"), e))

    def set_coordinates(self, crd):
        crd = _f64(crd).reshape(self.nbeads, 3)
        self._chk(self.L.bd_set_coordinates(self.h, _dp(crd)))

    def get_coordinates(self):
        out = np.empty((self.nbeads, 3))
        self._chk(self.L.bd_get_coordinates(self.h, _dp(out)))
        return out

    # -- the hot path pieces
    def calc_rpy_tensor(self):
        self._chk(self.L.bd_rpy_assemble(self.h))

    def get_mobility(self):
        m = np.empty((self.n, self.n))
        self._chk(self.L.bd_get_mobility(self.h, _dp(m)))
        return m.T.copy()  # column-major memory -> logical (row, col)

    def mob_symv(self, x):
        x = np.asarray(x, dtype=np.float64)
        vec = x.ndim == 1
        xm = np.asfortranarray(x.reshape(self.n, -1))
        s = xm.shape[1]
        y = np.zeros((self.n, s), order="F")
        self._chk(self.L.bd_mob_symv(self.h, xm.ctypes.data_as(DP), s, y.ctypes.data_as(DP)))
        return y[:, 0].copy() if vec else y

    def brn_calc_bdw(self, xi, method="LANC"):
        """B.xi for xi (n, s); returns (bdw, lconv, iters)."""
        xm = np.asfortranarray(np.asarray(xi, dtype=np.float64).reshape(self.n, -1))
        s = xm.shape[1]
        out = np.zeros((self.n, s), order="F")
        lconv = C.c_int(0)
        iters = C.c_int(0)
        m = MOB_FCTR[method] if isinstance(method, str) else int(method)
        self._chk(self.L.bd_brownian(self.h, xm.ctypes.data_as(DP), s, m, out.ctypes.data_as(DP),
                                     C.byref(lconv), C.byref(iters)), ok=(BD_OK, BD_WARN_NOTCONV))
        return out, bool(lconv.value), iters.value

    def ia_calc_forces(self):
        """Returns (forces (N,3), energies [bond, vdw, tot], stress (3,3), bndlen [mean,min,max])."""
        f = np.empty((self.nbeads, 3))
        e = np.empty(3)
        st = np.empty(9)
        bl = np.empty(3)
        self._chk(self.L.bd_forces(self.h, _dp(f), _dp(e), _dp(st), _dp(bl)))
        return f, e, st.reshape(3, 3, order="F"), bl

    def pair_table(self):
        """Unordered pairs (i < j, 1-based) of the last force evaluation; (P, 2) int."""
        npairs = C.c_int(0)
        rebuilt = C.c_int(0)
        rc = self.L.bd_pair_table(self.h, None, 0, C.byref(npairs), C.byref(rebuilt))
        if rc not in (BD_OK, BD_ERR_CAPACITY):
            self._chk(rc)
        buf = np.zeros((max(npairs.value, 1), 2), np.int32)
        self._chk(self.L.bd_pair_table(self.h, _ip(buf), npairs.value, C.byref(npairs), C.byref(rebuilt)))
        return buf[: npairs.value].copy(), bool(rebuilt.value)

    def bds_run(self, nsteps, xi=None):
        """nsteps EM steps; xi = injected Gaussian stream or None (device Philox)."""
        used = C.c_longlong(0)
        if xi is None:
            self._chk(self.L.bd_run(self.h, int(nsteps), None, 0, C.byref(used)))
        else:
            xa = _f64(np.ravel(xi))
            self._chk(self.L.bd_run(self.h, int(nsteps), _dp(xa), xa.size, C.byref(used)))
        return used.value

    def set_rng(self, seed, offset=0):
        self._chk(self.L.bd_set_rng(self.h, int(seed), int(offset)))

    def get_rng(self):
        """(seed, offset): the generator's whole state."""
        sd, off = C.c_uint64(0), C.c_uint64(0)
        self._chk(self.L.bd_get_rng(self.h, C.byref(sd), C.byref(off)))
        return sd.value, off.value

    def resume_run(self, molecule_com=(0.0, 0.0, 0.0)):
        """Continue a run from a revive: molecule_com restored, no second
        centring of the coordinates just set."""
        a = _f64(molecule_com)
        assert a.size == 3
        self._chk(self.L.bd_resume_run(self.h, _dp(a)))

    def stats_init(self, binsize=(1, 1), molecules=None):
        """stats_init (stats_m.f90:53-105): accumulate on the device from now
        on.  molecules: (nmol, 2) rows (atom count, first atom 1-based),
        needed for a periodic box (compute_ic1)."""
        mol = np.zeros((0, 2), np.int32) if molecules is None else _i32(np.reshape(molecules, (-1, 2)))
        self._chk(self.L.bd_stats_enable(self.h, int(binsize[0]), int(binsize[1]), mol.shape[0],
                                         _ip(mol) if mol.size else None))

    def stats(self):
        """The accumulators stats_write prints, by name (STATS_FIELDS)."""
        out = np.zeros(30)
        self._chk(self.L.bd_stats_get(self.h, _dp(out)))
        d, k = {}, 0
        for name, w in STATS_FIELDS:
            d[name] = out[k] if w == 1 else out[k:k + w].copy()
            k += w
        d["stress"] = d["stress"].reshape(3, 3, order="F")
        return d

    def runinfo(self):
        ri = BdRuninfo()
        self._chk(self.L.bd_get_runinfo(self.h, C.byref(ri)))
        return {f: (list(getattr(ri, f)) if f == "molecule_com" else getattr(ri, f))
                for f, _ in BdRuninfo._fields_}

    def reset_run(self):
        self._chk(self.L.bd_reset_run(self.h))

    def rgsq(self):
        v = C.c_double(0.0)
        self._chk(self.L.bd_rgsq(self.h, C.byref(v)))
        return v.value

    def sync(self):
        self._chk(self.L.bd_sync(self.h))

    # -- differentiable trajectory (fixed noise)
    def set_theta(self, eps, sigma, kbond, tie_rc=True):
        self._chk(self.L.bd_set_theta(self.h, float(eps), float(sigma), float(kbond), int(bool(tie_rc))))

    def run_record(self, nsteps, xi=None):
        """bds_run that also returns the per-step displacements (nsteps, N, 3)."""
        used = C.c_longlong(0)
        out = np.zeros((int(nsteps), self.nbeads, 3))
        if xi is None:
            self._chk(self.L.bd_run_record(self.h, int(nsteps), None, 0, C.byref(used), _dp(out)))
        else:
            xa = _f64(np.ravel(xi))
            self._chk(self.L.bd_run_record(self.h, int(nsteps), _dp(xa), xa.size, C.byref(used), _dp(out)))
        return out

    def run_replay(self, disp):
        d = _f64(disp).reshape(-1, self.nbeads, 3)
        self._chk(self.L.bd_run_replay(self.h, d.shape[0], _dp(d)))

    def adjoint_rgsq(self, disp, ckpt=20, tie_rc=True):
        """(final Rg^2, d/d(eps, sigma, k)) of a replayed trajectory."""
        d = _f64(disp).reshape(-1, self.nbeads, 3)
        loss = C.c_double(0.0)
        g = np.zeros(3)
        self._chk(self.L.bd_adjoint_rgsq(self.h, d.shape[0], _dp(d), int(ckpt), int(bool(tie_rc)),
                                         C.byref(loss), _dp(g)))
        return loss.value, g

    def prof_enable(self, on=True, width=0):
        """Bracket tile-product launches with HIP events (width > 0: only the
        launches of blocks with that many columns)."""
        self._chk(self.L.bd_prof_enable(self.h, (1 + int(width) if width else 1) if on else 0))

    def prof_get(self):
        """(summed tile-product kernel ms, launches, stored tile bytes per launch)."""
        ms = C.c_double(0.0)
        cnt = C.c_longlong(0)
        tb = C.c_double(0.0)
        self._chk(self.L.bd_prof_get(self.h, C.byref(ms), C.byref(cnt), C.byref(tb)))
        return ms.value, cnt.value, tb.value


class BdBatch:
    """Many independent linear chains (RPY + Lanczos noise) advanced together.

    Each chain follows bds_run (bd_solver_m.f90:128-238) for an isolated chain
    exactly as one BdStep would; the batch shares every launch.
    """

    def __init__(self, nbeads, *, nts_mobsam=1, lanc_mxitr=50, lanc_tol=1e-6, tim_stp=1e-3,
                 precision="fp64", keep_reuse_noise=False, mob_storage="auto"):
        self.L = load_library()
        self.nbeads = np.asarray(nbeads, dtype=np.int32).ravel().copy()
        self.nchains = int(self.nbeads.size)
        self.s = int(nts_mobsam)
        prec = {"fp64": FP64, "fp32": FP32}[precision] if isinstance(precision, str) else int(precision)
        self.opts = BdOpts(1, LANC, int(nts_mobsam), int(lanc_mxitr), float(lanc_tol), float(tim_stp), prec,
                           int(bool(keep_reuse_noise)), 0, 0.0, 0)
        # "auto" and "matfree": the RPY blocks formed inside every product;
        # "tiles": the stored lower triangles with the autotuned unit shape
        self.opts.mob_storage = MOB_STORAGE[mob_storage] if isinstance(mob_storage, str) else int(mob_storage)
        self.h = self.L.bd_batch_create(self.nchains, _ip(self.nbeads), C.byref(self.opts))
        if not self.h:
            raise BdError(BD_ERR_ARG, self.L.bd_last_error().decode())

    def _chk(self, rc, ok=(BD_OK,)):
        if rc not in ok:
            raise BdError(rc, self.L.bd_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.bd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_model(self, bond_style, bond_params, *, lvdw=True, vdw_params=None, rcutoff=0.0,
                  excluded_atoms=0):
        """One bond type for every bond i-(i+1), LJ style 1 for every pair (.cfg slots)."""
        bp = _params(bond_params, 1)
        vp = _params(vdw_params if vdw_params is not None else [], 1)
        self._chk(self.L.bd_batch_set_model(self.h, int(bond_style), _dp(bp), int(bool(lvdw)), _dp(vp),
                                            float(rcutoff), int(excluded_atoms)))

    def set_coordinates(self, chains):
        """chains: list of (nb_c, 3) arrays (or one flat array, chains back to back)."""
        flat = _f64(np.concatenate([np.ravel(c) for c in chains]) if isinstance(chains, (list, tuple))
                    else np.ravel(chains))
        if flat.size != 3 * int(self.nbeads.sum()):
            raise ValueError("coordinate count does not match the chain lengths")
        self._chk(self.L.bd_batch_set_coordinates(self.h, _dp(flat)))

    def get_coordinates(self):
        flat = np.zeros(3 * int(self.nbeads.sum()))
        self._chk(self.L.bd_batch_get_coordinates(self.h, _dp(flat)))
        out, p = [], 0
        for nb in self.nbeads:
            out.append(flat[p:p + 3 * nb].reshape(nb, 3).copy())
            p += 3 * nb
        return out

    def set_rng(self, seed, offset=0):
        self._chk(self.L.bd_batch_set_rng(self.h, int(seed), int(offset)))

    def run(self, nsteps, xi=None):
        used = C.c_longlong(0)
        if xi is None:
            self._chk(self.L.bd_batch_run(self.h, int(nsteps), None, 0, C.byref(used)))
        else:
            xa = _f64(np.ravel(xi))
            self._chk(self.L.bd_batch_run(self.h, int(nsteps), _dp(xa), xa.size, C.byref(used)))
        return used.value

    def info(self):
        nts = C.c_longlong(0)
        upd, it, nc = C.c_int(0), C.c_int(0), C.c_int(0)
        self._chk(self.L.bd_batch_info(self.h, C.byref(nts), C.byref(upd), C.byref(it), C.byref(nc)))
        return {"nts": nts.value, "mob_updates": upd.value, "total_lanczos_iters": it.value,
                "n_not_converged": nc.value}

    def tile_shape(self):
        """((chr, wsc) in use, autotuner timings in us of TILE_SHAPES)."""
        a, b = C.c_int(0), C.c_int(0)
        t = np.zeros(len(TILE_SHAPES))
        self._chk(self.L.bd_batch_tile_shape(self.h, C.byref(a), C.byref(b), _dp(t)))
        return (a.value, b.value), t

    def set_tile_shape(self, chr_, wsc):
        self._chk(self.L.bd_batch_set_tile_shape(self.h, int(chr_), int(wsc)))

    def prof(self, on=-1):
        ms = C.c_double(0.0)
        cnt = C.c_longlong(0)
        tb = C.c_double(0.0)
        self._chk(self.L.bd_batch_prof(self.h, int(on), C.byref(ms), C.byref(cnt), C.byref(tb)))
        return ms.value, cnt.value, tb.value
