"""Work decomposition of the symmetric tile product (CPU, no GPU).

The unit and output-block records a single-chain context would hand to
symm_tiles_kernel / the reduction kernels (bd_plan_layout) are replayed on
the host with the kernels' own semantics: a unit adds D_IJ x_J to the row
partial (I, super-column) and D_IJ^T x_I to the column partial (canonical
unit, J - J0) for each stored tile of its mask, and writes every row
partial of its rows and every column partial up to its super-column's width
(zero where it holds no tile); output block K sums the row partials of the
super-columns up to its own and the column partials of its super-column's
units at its offset.  The result must equal D x, every lower tile must
belong to exactly one unit, and the execution order must put at most two
units on a CU (blocks b and b + CUs share one)."""
import numpy as np
import pytest

from brownpak_amd.bdstep import plan_layout

TILE = 64


def _replay(nbeads, cus, S=2, seed=0, dense=True):
    (chr_, wsc), units, rowblk = plan_layout(nbeads, cus)
    n = 3 * nbeads
    nt = (n + TILE - 1) // TILE
    npad = nt * TILE
    assert rowblk.shape[0] == nt
    owner = {}
    rowpart, colpart = {}, {}
    rng = np.random.default_rng(seed)
    if dense:
        A = rng.standard_normal((npad, npad))
        D = A + A.T
        x = rng.standard_normal((npad, S))
    for pos, rec in enumerate(units):
        ch, sc, Ib, Ie, ntile, nsc, mask, J0 = (int(v) for v in rec[:8])
        cid, w = int(rec[14]), int(rec[15])
        assert ch == 0 and ntile == nt and 1 <= w <= wsc and Ib >= J0
        for ii in range(chr_):
            I = Ib + ii
            if I < Ie:
                rowpart.setdefault((I, sc), 0.0)
            for jj in range(wsc):
                if not (mask >> (ii * wsc + jj)) & 1:
                    continue
                J = J0 + jj
                assert J <= I < Ie and J < J0 + w, (pos, I, J)
                assert (I, J) not in owner, (I, J)
                owner[(I, J)] = pos
                if dense:
                    T = D[I * TILE:(I + 1) * TILE, J * TILE:(J + 1) * TILE]
                    rowpart[(I, sc)] = rowpart[(I, sc)] + T @ x[J * TILE:(J + 1) * TILE]
                    if I != J:
                        colpart[(cid, jj)] = colpart.get((cid, jj), 0.0) + T.T @ x[I * TILE:(I + 1) * TILE]
        for jj in range(w):
            colpart.setdefault((cid, jj), 0.0)
    # every stored tile exactly once
    assert len(owner) == nt * (nt + 1) // 2
    # canonical indices are a permutation
    assert sorted(int(c) for c in units[:, 14]) == list(range(len(units)))
    y = np.zeros((npad, S)) if dense else None
    for K, rb in enumerate(rowblk):
        ch, K2, nsc, u0, u1, rpo, yrow, scoff = (int(v) for v in rb)
        scK, off = scoff & 0xFFFF, scoff >> 16
        assert K2 == K and yrow == K * TILE and u0 < u1
        parts = [rowpart[(K, s)] for s in range(scK + 1)] + [colpart[(u, off)] for u in range(u0, u1)]
        if dense:
            y[K * TILE:(K + 1) * TILE] = sum(parts)
    return (chr_, wsc), units, (y, D @ x) if dense else None


@pytest.mark.parametrize("nbeads,cus", [(20, 256), (150, 256), (700, 256), (1300, 256), (700, 16), (1300, 80)])
def test_plan_replay_equals_dense_product(nbeads, cus):
    _, _, (y, ref) = _replay(nbeads, cus)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("nbeads", [2000, 3000])
def test_plan_covers_every_tile_once(nbeads):
    _replay(nbeads, 256, dense=False)


def test_plan_pairs_units_on_cus():
    """N = 2000 on 256 CUs: one round (at most two units per CU), and the
    tiles on any CU stay below the 20 that 470 uniform 2 x 5 units give."""
    shape, units, _ = _replay(2000, 256, dense=False)
    U = len(units)
    assert U <= 512
    tiles = np.array([bin(int(m)).count("1") for m in units[:, 6]])
    per_cu = tiles[:256].copy()
    per_cu[: max(0, U - 256)] += tiles[256:]
    assert per_cu.max() < 20, (shape, U, per_cu.max())


def test_plan_is_a_function_of_length_and_cus():
    a = plan_layout(2000, 256)
    b = plan_layout(2000, 256)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
