"""Synthetic starting configurations for the BD step (no file I/O).

Lengths are in the reference's native unit, the bead hydrodynamic radius
a = 1, so the bead diameter sigma is 2 (bd_solver_m.f90:484-491 tests r >= 2
for the RPY overlap branch).  The random walk follows the 'efjc' recipe of the
reference's generators (utils/models/configuration.py:121-172, overlap test
has_overlap 458-472): each new bead sits at a uniformly random direction on a
sphere of radius ``bond`` around the previous bead, redrawn up to 100 times
while it comes closer than ``sep`` to any earlier bead.
"""
from __future__ import annotations

import numpy as np

SIGMA = 2.0


def _unit_vectors(rng, k):
    v = rng.standard_normal((k, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def random_walk_chain(nbeads: int, seed: int, bond: float = 0.97 * SIGMA,
                      sep: float = 1.9, maxitr: int = 100) -> np.ndarray:
    """Seeded self-avoiding-ish random walk; returns (nbeads, 3) float64."""
    rng = np.random.default_rng(seed)
    crd = np.zeros((nbeads, 3))
    sepsq = sep * sep
    for i in range(1, nbeads):
        prev = crd[i - 1]
        for _ in range(maxitr):
            ri = prev + bond * _unit_vectors(rng, 1)[0]
            d = crd[:i] - ri
            if np.min(np.einsum("ij,ij->i", d, d)) >= sepsq:
                break
        crd[i] = ri
    return crd


def overlap_free_chain(nbeads: int, seed: int, bond: float = 0.97 * SIGMA,
                       sep: float = 1.9, maxitr: int = 5000) -> np.ndarray:
    """random_walk_chain, regrown from derived seeds (seed, attempt) until no
    two non-bonded beads are closer than sep.  The plain walk accepts its last
    trial position when a bead is trapped, which occasionally leaves an
    overlap whose WCA force overstretches a bond on the first step; long
    ensembles need every chain clean."""
    for attempt in range(1000):
        s = seed if attempt == 0 else int(np.random.SeedSequence([seed, attempt]).generate_state(1)[0])
        crd = random_walk_chain(nbeads, s, bond, sep, maxitr)
        d = crd[:, None, :] - crd[None, :, :]
        r2 = np.einsum("ijk,ijk->ij", d, d)
        iu = np.triu_indices(nbeads, 2)
        if nbeads < 3 or r2[iu].min() >= sep * sep * (1.0 - 1e-12):
            return crd
    raise RuntimeError(f"no overlap-free walk of {nbeads} beads from seed {seed}")


def chain_bonds(nbeads: int, offset: int = 0, btype: int = 1) -> np.ndarray:
    """(nbeads-1, 3) int32 rows [type, i, j], 1-based, for a linear chain."""
    i = np.arange(1, nbeads, dtype=np.int32) + offset
    return np.stack([np.full_like(i, btype), i, i + 1], axis=1).astype(np.int32)


def fene_wca_params(sigma: float = SIGMA, eps: float = 1.0):
    """FENE (k = 30 eps/sigma^2, Q0 = 1.5 sigma, r0 = 0) and WCA
    (rc = 2^(1/6) sigma) in reference units: bond style 2 params (k, rmax, r0)
    and vdw style 1 params (eps, sigma, rcut)."""
    bond = np.array([30.0 * eps / sigma ** 2, 1.5 * sigma, 0.0])
    vdw = np.array([eps, sigma, 2.0 ** (1.0 / 6.0) * sigma])
    return bond, vdw


def periodic_melt(nchains: int, nper: int, density: float, seed: int,
                  bond: float = SIGMA, sep: float = 0.9 * SIGMA):
    """Chains in a cubic periodic box at the given bead number density
    (beads per sigma^3); positions are wrapped into [0, L).  Returns
    (crd (N,3), box L)."""
    nb = nchains * nper
    L = (nb / density) ** (1.0 / 3.0) * SIGMA
    rng = np.random.default_rng(seed)
    crd = np.zeros((nb, 3))
    # coarse cell grid for the overlap test
    ncell = max(int(L // sep), 1)
    cs = L / ncell
    grid: dict = {}

    def key(p):
        c = np.floor(p / cs).astype(int) % ncell
        return int(c[0]), int(c[1]), int(c[2])

    def overlaps(p):
        c = key(p)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    kk = ((c[0] + dx) % ncell, (c[1] + dy) % ncell, (c[2] + dz) % ncell)
                    for j in grid.get(kk, ()):
                        d = crd[j] - p
                        d -= L * np.round(d / L)
                        if d @ d < sep * sep:
                            return True
        return False

    k = 0
    for _ in range(nchains):
        while True:
            p = rng.random(3) * L
            if not overlaps(p):
                break
        crd[k] = p
        grid.setdefault(key(p), []).append(k)
        k += 1
        for _ in range(1, nper):
            for _it in range(200):
                p = (crd[k - 1] + bond * _unit_vectors(rng, 1)[0]) % L
                if not overlaps(p):
                    break
            crd[k] = p
            grid.setdefault(key(p), []).append(k)
            k += 1
    return crd, L
