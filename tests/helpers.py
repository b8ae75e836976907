"""Shared inputs for the tests: seeded systems, hypers and sparse sets built
with the FP64 oracle (the reference of every parity test)."""
import functools

import numpy as np

from turbogap_amd import structures as S


@functools.lru_cache(maxsize=None)
def carbon216(n_sparse=2000):
    from oracle import oracle as O
    sysm = S.carbon216_system()
    h = S.soap_carbon()
    Qs, al = S.sparse_set(h, n_sparse, O.block_descriptors, S.carbon216_system, seed=5001)
    return sysm, h, Qs, al


@functools.lru_cache(maxsize=None)
def subset(n_sparse=500, n=4096):
    """Two-species box at the density of the 65,536-atom box with both central-species blocks."""
    from oracle import oracle as O
    sysm = S.binary_system(n, seed=2002)
    models = []
    for k, h in enumerate(S.soap_binary()):
        Qs, al = S.sparse_set(h, n_sparse, O.block_descriptors, lambda s: S.binary_system(1024, seed=s),
                              seed=7001 + 100 * k)
        models.append((h, Qs, al))
    return sysm, models


def small_box(n=24, L=5.5, seed=11):
    """A box shorter than 2 rcut: needs the supercell and the brute-force branch."""
    pos = S.rsa(n, L, 1.2, seed)
    return S.System(pos, np.full(n, S.CARBON, np.int32), S.cubic_box(L))


def load_models(ctx, models):
    return [ctx.soap_model(h, Qs, al) for h, Qs, al in models]
