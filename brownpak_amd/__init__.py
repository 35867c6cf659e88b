"""brownpak_amd: an MI355X-native Brownian-dynamics step.

The hot path (RPY mobility, Cholesky / block-Lanczos B.xi, bond and pair
forces, the Euler-Maruyama update) runs as hand-written HIP kernels for
gfx950 in ``_build/libbdstep.so`` behind the C ABI of ``include/bdstep.h``.
``bdstep.BdStep`` is the Python mirror of the reference's call boundary;
``bdstep.BdBatch`` advances many independent chains together.
"""
from .bdstep import BdBatch, BdError, BdStep, load_library  # noqa: F401
from .configs import chain_bonds, fene_wca_params, random_walk_chain  # noqa: F401

__all__ = ["BdStep", "BdBatch", "BdError", "load_library", "random_walk_chain", "chain_bonds",
           "fene_wca_params"]
