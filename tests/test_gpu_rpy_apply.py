"""bd_rpy_apply: the context-free matrix-free RPY product on device buffers,
y = D(crd) x, against the oracle's RPY matrix (calc_rpy_tensor restated,
bd_solver_m.f90:458-522) times x.  Bars: FP64 1e-13 relative (norm-wise,
per column block); FP32 arithmetic 3e-7 (about 8e-8 measured)."""
import numpy as np
import pytest

from helpers import rel
from brownpak_amd.configs import random_walk_chain

pytestmark = pytest.mark.gpu


def _apply(fn, crd, x, ldx, ldy, prec):
    """x (n, s) host -> y (n, s) through fn on device buffers with leading
    dimensions ldx, ldy (column-major)."""
    from devbuf import DevArray
    from brownpak_amd import load_library
    from brownpak_amd.bdstep import BD_OK, BdError
    n, s = x.shape
    xs = np.zeros((s, ldx))
    xs[:, :n] = x.T
    dc, dx, dy = DevArray(crd.ravel()), DevArray(xs), DevArray(np.full((s, ldy), 7.0))
    rc = fn(n // 3, dc.ptr, dx.ptr, ldx, s, dy.ptr, ldy, prec, None)
    if rc != BD_OK:
        raise BdError(rc, load_library().bd_last_error().decode())
    out = dy.get()
    for b in (dc, dx, dy):
        b.free()
    assert np.all(out[:, n:] == 7.0)       # nothing written past n rows
    return out[:, :n].T


@pytest.mark.parametrize("nb,s", [(100, 1), (300, 4), (2000, 4), (700, 8)])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_rpy_apply_matches_oracle(oracle, nb, s, prec):
    from brownpak_amd import load_library
    from brownpak_amd.bdstep import FP32, FP64
    L = load_library()
    crd = random_walk_chain(nb, nb + s)
    x = np.random.default_rng(nb * s).standard_normal((3 * nb, s))
    y = _apply(L.bd_rpy_apply, crd, x, 3 * nb + 5, 3 * nb + 3, FP64 if prec == "fp64" else FP32)
    want = oracle.rpy(crd) @ x
    err = rel(y, want)
    print(f"nb {nb} s {s} {prec}: rel {err:.2e}", flush=True)
    assert err < (1e-13 if prec == "fp64" else 3e-7)


def test_rpy_apply_rejects_bad_arguments():
    from brownpak_amd import load_library
    from brownpak_amd.bdstep import BD_ERR_ARG, FP64
    L = load_library()
    assert L.bd_rpy_apply(10, None, None, 30, 4, None, 30, FP64, None) == BD_ERR_ARG
    assert L.bd_rpy_apply(10, 1, 1, 29, 4, 1, 30, FP64, None) == BD_ERR_ARG      # ldx < n
    assert L.bd_rpy_apply(10, 1, 1, 30, 9, 1, 30, FP64, None) == BD_ERR_ARG      # s > 8
