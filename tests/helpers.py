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


GOLD = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)), "golden")


def silicon_crystal(perturb=0.05, seed=4):
    """216-atom diamond silicon (a = 5.43 Å, 3x3x3 cells), seeded rattle."""
    a = 5.43
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    cells = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)])
    pos = ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3)) * a
    if perturb:
        pos = pos + np.random.default_rng(seed).normal(scale=perturb, size=pos.shape)
    return S.System(pos, np.full(pos.shape[0], 14, np.int32), S.cubic_box(3 * a))


@functools.lru_cache(maxsize=None)
def silicon_compressed():
    """The compressed soap_turbo block of the reference suite's silicon.gap
    (tests/stopping/1/gap_files: hypers, compress.dat, the 9,000 alphas of
    alphas_mb.dat) on a rattled 216-atom crystal. Its desc_sparse file is not in
    the reference tree: the sparse set is the oracle's compressed descriptors of
    seeded rattled crystals, one row per alpha."""
    import os
    from oracle import oracle as O
    from turbogap_amd import gap_file as G
    blocks = G.read_gap_file(os.path.join(GOLD, "silicon.gap"))
    (b,) = [x for x in blocks if x.kind == "soap_turbo"]
    h = G.compressed_hypers(b, GOLD)
    al = np.loadtxt(os.path.join(GOLD, b.file_alphas))
    rows, k = [], 0
    while sum(r.shape[0] for r in rows) < al.shape[0]:
        rows.append(O.block_descriptors(h, silicon_crystal(0.15, 9100 + k)))
        k += 1
    Qs = np.concatenate(rows)[:al.shape[0]].copy()
    return silicon_crystal(0.1, 77), h, Qs, al
