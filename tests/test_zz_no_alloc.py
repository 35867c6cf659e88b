"""bd_rpy_apply_noalloc (static device scratch, no allocation) against
bd_rpy_apply (the same product with scratch it allocates): bit-identical
outputs, on two calls in a row, the second smaller than the first (the
scratch keeps the first call's contents), and its size limit."""
import numpy as np
import pytest

from brownpak_amd.configs import random_walk_chain

pytestmark = pytest.mark.gpu


def _run(fn, crd, x, ldx, ldy, prec):
    from devbuf import DevArray
    from brownpak_amd import load_library
    from brownpak_amd.bdstep import BD_OK, BdError
    n, s = x.shape
    xs = np.zeros((s, ldx))
    xs[:, :n] = x.T
    dc, dx, dy = DevArray(crd.ravel()), DevArray(xs), DevArray(np.full((s, ldy), -3.0))
    rc = fn(n // 3, dc.ptr, dx.ptr, ldx, s, dy.ptr, ldy, prec, None)
    if rc != BD_OK:
        raise BdError(rc, load_library().bd_last_error().decode())
    out = dy.get()
    for b in (dc, dx, dy):
        b.free()
    return out


def test_noalloc_matches_allocating_entry_on_consecutive_calls():
    from brownpak_amd import load_library
    from brownpak_amd.bdstep import FP32, FP64
    L = load_library()
    cases = [(2000, 4, FP64), (300, 2, FP64), (1500, 8, FP32), (64, 1, FP32)]
    for nb, s, prec in cases:
        print(f"case: noalloc nb={nb} s={s} prec={prec}", flush=True)
        crd = random_walk_chain(nb, nb + 1)
        x = np.random.default_rng(nb).standard_normal((3 * nb, s))
        ref = _run(L.bd_rpy_apply, crd, x, 3 * nb + 2, 3 * nb + 1, prec)
        got = _run(L.bd_rpy_apply_noalloc, crd, x, 3 * nb + 2, 3 * nb + 1, prec)
        assert np.array_equal(got, ref), (nb, s, prec)
        assert np.all(got[:, 3 * nb:] == -3.0)


def test_noalloc_rejects_sizes_beyond_its_scratch():
    from brownpak_amd import load_library
    from brownpak_amd.bdstep import BD_ERR_ARG, FP64
    L = load_library()
    assert L.bd_rpy_apply_noalloc(8193, 1, 1, 3 * 8193, 4, 1, 3 * 8193, FP64, None) == BD_ERR_ARG
    assert b"8192" in L.bd_last_error()
