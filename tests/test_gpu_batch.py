"""Batched ensemble (bd_batch_*): every chain of a batch follows its own
bds_run exactly.

Bars:
  * per chain against the oracle, same injected xi: 1e-10 relative at a fixed
    Lanczos iteration count (lanc_tol = 0), as the single-chain tests;
  * converged runs: each chain's Brownian block within 10 * lanc_tol of the
    exact square root (eigen-decomposition of the oracle's RPY matrix);
  * chain independence at the batched-ensemble sizes (lengths U[50, 800]):
    a chain's trajectory is bit-identical whether it runs alone or inside
    the batch (the product partials and reductions are per chain and fixed
    order, so the other chains cannot leak into its arithmetic).
"""
import numpy as np
import pytest

from helpers import fene_wca_chain, oracle_system, rel
from brownpak_amd.configs import fene_wca_params, random_walk_chain

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _batch(lengths, seed, overlap_free=False, **opts):
    from brownpak_amd import BdBatch
    from brownpak_amd.configs import overlap_free_chain
    b = BdBatch(lengths, **opts)
    bond, vdw = fene_wca_params()
    b.set_model(2, bond, lvdw=True, vdw_params=vdw, rcutoff=vdw[2], excluded_atoms=0)
    gen = overlap_free_chain if overlap_free else random_walk_chain
    chains = [gen(int(nb), seed + c) for c, nb in enumerate(lengths)]
    b.set_coordinates(chains)
    return b, chains


def _split_xi(lengths, s, nupd, xi):
    """Per-chain streams of the batch's xi layout (chain by chain per update)."""
    per = [[] for _ in lengths]
    p = 0
    for _ in range(nupd):
        for c, nb in enumerate(lengths):
            k = 3 * nb * s
            per[c].append(xi[p:p + k])
            p += k
    return [np.concatenate(x) for x in per]


@pytest.mark.parametrize("storage", ["matfree", "tiles"])
@pytest.mark.parametrize("s,prec", [(1, "fp64"), (4, "fp64"), (3, "fp64"), (4, "fp32")])
def test_batch_matches_oracle_fixed_iterations(oracle, s, prec, storage):
    lengths = [40, 17, 64, 23, 50, 130]
    nsteps, mx = 2 * s, 5
    nupd = (nsteps + s - 1) // s
    _say(f"case: batch fixed-iteration s={s} {prec} {storage} lengths={lengths}")
    tot = sum(3 * nb * s for nb in lengths) * nupd
    xi = np.random.default_rng(7 + s).standard_normal(tot)
    b, chains = _batch(lengths, 300, nts_mobsam=s, lanc_mxitr=mx, lanc_tol=0.0, precision=prec,
                       mob_storage=storage)
    used = b.run(nsteps, xi)
    assert used == tot
    out = b.get_coordinates()
    per = _split_xi(lengths, s, nupd, xi)
    tol = 1e-10 if prec == "fp64" else 1e-6
    for c, nb in enumerate(lengths):
        m = fene_wca_chain(nb, 0)
        m["crd"] = chains[c]
        so = oracle_system(oracle, m)
        ref, used_c, _ = oracle.bd_run(so, chains[c], nsteps, per[c], mob_fctr=oracle.LANC, nts_mobsam=s,
                                       lanc_mxitr=mx, lanc_tol=0.0)
        assert used_c == per[c].size
        assert rel(out[c], ref) < tol, (c, nb, rel(out[c], ref))
    info = b.info()
    assert info["nts"] == nsteps and info["mob_updates"] == nupd
    assert info["total_lanczos_iters"] == nupd * len(lengths) * (mx + 1)


@pytest.mark.parametrize("s", [1, 2, 4])
def test_batch_converged_noise_accuracy(oracle, s):
    """One step; the displacement minus the exact drift is the Brownian term
    (s = 1: the two-norm stopping test; s > 1: the column-sum one)."""
    lengths = [30, 55, 21, 80]
    tol = 1e-6
    _say(f"case: batch converged noise s={s}")
    xi = np.random.default_rng(11).standard_normal(sum(3 * nb * s for nb in lengths))
    b, chains = _batch(lengths, 500, nts_mobsam=s, lanc_mxitr=60, lanc_tol=tol, tim_stp=0.5)
    # keep_reuse_noise off: steps 2..s of a block carry no noise; one step is enough
    before = [c - c.mean(axis=0) for c in chains]
    b.run(1, xi)
    after = b.get_coordinates()
    per = _split_xi(lengths, s, 1, xi)
    for c, nb in enumerate(lengths):
        n = 3 * nb
        m = fene_wca_chain(nb, 0)
        m["crd"] = chains[c]
        so = oracle_system(oracle, m)
        D = oracle.rpy(chains[c])
        w, U = np.linalg.eigh(D)
        X = per[c].reshape(s, n).T
        exact = (U * np.sqrt(w)) @ (U.T @ X)
        # drift of this step from the oracle's bond + pair forces
        f = oracle.forces(so, chains[c], oracle.pairs(so, chains[c], "DIR"))[0].ravel()
        disp = after[c].ravel() - before[c].ravel()
        brown = np.sqrt(2 * 0.5) * exact[:, 0]
        expect = 0.5 * (D @ f) + brown
        expect = expect.reshape(nb, 3)
        expect -= expect.mean(axis=0)
        err = np.linalg.norm(disp - expect.ravel()) / np.linalg.norm(brown)
        assert err < 10 * tol, (c, err)
    assert b.info()["n_not_converged"] == 0


@pytest.mark.parametrize("storage", ["matfree", "tiles"])
def test_batch_chain_independence_ensemble(storage):
    """Lengths U[50, 800] (seeded), FP32, s = 4, tol 1e-6, device RNG
    replaced by an injected stream so single runs see the same numbers."""
    rng = np.random.default_rng(2024)
    lengths = [int(x) for x in rng.integers(50, 801, size=24)]
    s, nsteps = 4, 8
    _say(f"case: batch independence {storage} {len(lengths)} chains, {sum(lengths)} beads")
    tot = sum(3 * nb * s for nb in lengths) * (nsteps // s)
    xi = rng.standard_normal(tot)
    b, chains = _batch(lengths, 900, nts_mobsam=s, lanc_mxitr=50, lanc_tol=1e-6, precision="fp32",
                       mob_storage=storage)
    if storage == "tiles":
        b.set_tile_shape(2, 4)   # the single-chain batches below use the same unit shape
    b.run(nsteps, xi)
    out = b.get_coordinates()
    assert all(np.isfinite(o).all() for o in out)
    for o in out:
        assert np.abs(o.mean(axis=0)).max() < 1e-9
    per = _split_xi(lengths, s, nsteps // s, xi)
    for c in (0, 7, len(lengths) - 1):
        one, _ = _batch([lengths[c]], 900 + c, nts_mobsam=s, lanc_mxitr=50, lanc_tol=1e-6, precision="fp32",
                        mob_storage=storage)
        if storage == "tiles":
            one.set_tile_shape(2, 4)
        one.run(nsteps, per[c])
        assert np.array_equal(one.get_coordinates()[0], out[c]), c


@pytest.mark.parametrize("shape", [(2, 4), (1, 4), (4, 4), (2, 2), (2, 5), (2, 6), (2, 8)])
def test_batch_tile_shapes_match_oracle(oracle, shape):
    """Every work-unit shape of the tile product gives the oracle's chains
    (fixed Lanczos iteration count, FP64, 1e-10)."""
    lengths = [40, 17, 64, 23, 50, 90]
    s, nsteps, mx = 4, 4, 5
    _say(f"case: batch tile shape {shape}")
    tot = sum(3 * nb * s for nb in lengths)
    xi = np.random.default_rng(17).standard_normal(tot)
    b, chains = _batch(lengths, 310, nts_mobsam=s, lanc_mxitr=mx, lanc_tol=0.0, mob_storage="tiles")
    b.set_tile_shape(*shape)
    assert b.tile_shape()[0] == shape
    b.run(nsteps, xi)
    out = b.get_coordinates()
    per = _split_xi(lengths, s, 1, xi)
    for c, nb in enumerate(lengths):
        m = fene_wca_chain(nb, 0)
        m["crd"] = chains[c]
        so = oracle_system(oracle, m)
        ref, _, _ = oracle.bd_run(so, chains[c], nsteps, per[c], mob_fctr=oracle.LANC, nts_mobsam=s,
                                  lanc_mxitr=mx, lanc_tol=0.0)
        assert rel(out[c], ref) < 1e-10, (shape, c, rel(out[c], ref))


def test_batch_autotuner_keeps_the_fastest_shape():
    from brownpak_amd import BdError
    rng = np.random.default_rng(5)
    lengths = [int(x) for x in rng.integers(50, 801, size=32)]
    b, _ = _batch(lengths, 60, nts_mobsam=4, lanc_mxitr=20, lanc_tol=1e-6, precision="fp32",
                  mob_storage="tiles")
    shape, t = b.tile_shape()
    from brownpak_amd.bdstep import TILE_SHAPES
    shapes = list(TILE_SHAPES)
    assert len(t) == len(shapes) == 7
    assert shape in shapes and np.all(t > 0)
    assert shape == shapes[int(np.argmin(t))]
    # a second batch of the same lengths reuses the choice
    b2, _ = _batch(lengths, 61, nts_mobsam=4, lanc_mxitr=20, lanc_tol=1e-6, precision="fp32",
                   mob_storage="tiles")
    assert b2.tile_shape()[0] == shape
    with pytest.raises(BdError):
        b.set_tile_shape(3, 4)
    # a matrix-free batch has no tile shape to report or set
    m, _ = _batch(lengths[:4], 62, nts_mobsam=4, lanc_mxitr=20, lanc_tol=1e-6, precision="fp32",
                  mob_storage="matfree")
    with pytest.raises(BdError):
        m.tile_shape()
    with pytest.raises(BdError):
        m.set_tile_shape(2, 4)


@pytest.mark.parametrize("storage", ["matfree", "tiles"])
def test_batch_graph_blocks_match_step_by_step(storage):
    """Device noise: whole mobility blocks run as one HIP graph launch each;
    the same seed stepped launch by launch (profiling forces that path)
    gives bit-identical chains and the same Lanczos tallies, also across a
    partial block."""
    rng = np.random.default_rng(77)
    lengths = [int(x) for x in rng.integers(50, 401, size=16)]
    _say(f"case: batch graph vs eager {storage} {len(lengths)} chains")
    outs, infos = [], []
    for eager in (False, True):
        b, _ = _batch(lengths, 700, nts_mobsam=4, lanc_mxitr=50, lanc_tol=1e-6, precision="fp32",
                      mob_storage=storage)
        if storage == "tiles":
            b.set_tile_shape(2, 4)
        b.set_rng(99, 0)
        if eager:
            b.prof(1)
        b.run(8)
        b.run(6)     # one graph block, then two single steps
        outs.append(b.get_coordinates())
        infos.append(b.info())
    for a, e in zip(*outs):
        assert np.array_equal(a, e)
    assert infos[0] == infos[1]
    assert infos[0]["nts"] == 14 and infos[0]["mob_updates"] == 4


@pytest.mark.parametrize("path", ["graph", "eager"])
def test_batch_overstretch_reports_bond_error(path):
    """A bond beyond rmax at the first step: BD_ERR_BOND, no step counted,
    and every chain left where the failing step found it (bds_run returns
    before the update, bd_solver_m.f90:174; the centre of mass was moved to
    the origin once before the loop, bd_solver_m.f90:161)."""
    from brownpak_amd import BdError
    lengths = [20, 30, 25]
    b, chains = _batch(lengths, 77, nts_mobsam=4, lanc_mxitr=20, lanc_tol=1e-6)
    chains[1][15:] += np.array([6.0, 0.0, 0.0])
    b.set_coordinates(chains)
    with pytest.raises(BdError) as ei:
        if path == "graph":
            b.set_rng(5, 0)
            b.run(8)
        else:
            b.run(8, np.random.default_rng(5).standard_normal(3 * sum(lengths) * 4 * 2))
    assert ei.value.code == 1
    assert b.info()["nts"] == 0
    for o, c in zip(b.get_coordinates(), chains):
        assert np.allclose(o, c - c.mean(axis=0), rtol=0, atol=1e-12)


C3_CHAINS = 512


@pytest.mark.parametrize("storage", ["matfree", "tiles"])
def test_batch_c3_full_size(oracle, storage):
    """The dilute-ensemble configuration at its full size: 512 chains with
    lengths from U[50, 800] (seed 3), FENE/WCA/RPY, block Lanczos s = 4,
    tol 1e-6, FP32 products (matrix-free or stored tiles).  (1) Device noise, whole mobility blocks as HIP
    graph launches, 8 steps: every chain finite, centre of mass at the
    origin, every solve converged.  (2) Injected xi, one mobility block:
    six sampled chains bit-identical to the same chain run alone, and three
    chains against the oracle at a fixed Lanczos iteration count (FP32
    products: 1e-6)."""
    rng = np.random.default_rng(3)
    lengths = [int(x) for x in rng.integers(50, 801, size=C3_CHAINS)]
    s = 4
    _say(f"case: batch C3 full size {storage}, {sum(lengths)} beads")
    b, chains = _batch(lengths, 1000, overlap_free=True, nts_mobsam=s, lanc_mxitr=50, lanc_tol=1e-6,
                       precision="fp32", mob_storage=storage)
    b.set_rng(11, 0)
    b.run(8)
    out = b.get_coordinates()
    info = b.info()
    assert info["nts"] == 8 and info["mob_updates"] == 2 and info["n_not_converged"] == 0
    for o in out:
        assert np.isfinite(o).all()
        assert np.abs(o.mean(axis=0)).max() < 1e-9
    # (2) injected xi, chain independence and oracle parity at fixed iterations
    shape = b.tile_shape()[0] if storage == "tiles" else None
    tot = sum(3 * nb * s for nb in lengths)
    xi = rng.standard_normal(tot)
    per = _split_xi(lengths, s, 1, xi)
    for tol, mx in ((1e-6, 50), (0.0, 5)):
        bb, _ = _batch(lengths, 1000, overlap_free=True, nts_mobsam=s, lanc_mxitr=mx, lanc_tol=tol,
                       precision="fp32", mob_storage=storage)
        if shape:
            bb.set_tile_shape(*shape)
        bb.run(s, xi)
        res = bb.get_coordinates()
        if tol > 0:
            for c in (0, 101, 202, 303, 404, C3_CHAINS - 1):
                one, _ = _batch([lengths[c]], 1000 + c, overlap_free=True, nts_mobsam=s, lanc_mxitr=mx,
                                lanc_tol=tol, precision="fp32", mob_storage=storage)
                if shape:
                    one.set_tile_shape(*shape)
                one.run(s, per[c])
                assert np.array_equal(one.get_coordinates()[0], res[c]), c
        else:
            for c in (5, 250, 500):
                m = fene_wca_chain(lengths[c], 0)
                m["crd"] = chains[c]
                so = oracle_system(oracle, m)
                ref, _, _ = oracle.bd_run(so, chains[c], s, per[c], mob_fctr=oracle.LANC, nts_mobsam=s,
                                          lanc_mxitr=mx, lanc_tol=0.0)
                _say(f"  chain {c} (N={lengths[c]}): rel {rel(res[c], ref):.2e}")
                assert rel(res[c], ref) < 1e-6, c


def test_batch_rejects_unsupported_options():
    from brownpak_amd import BdBatch, BdError
    with pytest.raises(BdError):
        BdBatch([10, 20], nts_mobsam=6)
    with pytest.raises(BdError):
        BdBatch([1, 20])
