"""Compressed soap_turbo blocks on the device against the oracle (C10).

  * The reference suite's silicon potential block (compress.dat keeps 72 of
    324 components, zeta = 6, radial_enhancement = 1, its 9,000 alphas) on a
    rattled 216-atom crystal: descriptors and their derivatives to 1e-12, and E,
    F, virial of the FP64 plain and adjoint paths to the parity bar (E 1e-8
    relative, F and virial 1e-6 in the potential's own force scale).
  * A two-species block with a general matrix (three entries per compressed
    component, random weights) on the 4,096-atom box: both FP64 paths.
  * The mixed path refuses a compressed block.
"""
import numpy as np
import pytest

from tests import helpers as H
from turbogap_amd import structures as S
from turbogap_amd.lib import TgxError

pytestmark = pytest.mark.gpu


def test_silicon_compressed_descriptors(ctx, orc):
    sysm, h, Qs, al = H.silicon_compressed()
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    nl = ctx.build_neighbors_list(h.rcut_max)
    mid = ctx.soap_model(h, Qs, al)
    ctx.soap_backend_begin(mid)
    q, der = ctx.get_soap(derivatives=True)
    ctx.soap_backend_end()
    assert q.shape == (sysm.n, 72)
    spi = orc.species_index(nl["neighbor_species"], h.species_types)
    qo, dero = orc.get_soap(h, nl["n_neigh"], spi, nl["rjs"], nl["xyz"], derivatives=True)
    np.testing.assert_allclose(q, qo, rtol=0, atol=1e-12)
    np.testing.assert_allclose(der, dero, rtol=0, atol=1e-11)


@pytest.mark.parametrize("path", ["fp64_plain", "fp64_adjoint"])
def test_silicon_compressed_gap(ctx, orc, path):
    sysm, h, Qs, al = H.silicon_compressed()
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(h.rcut_max, fetch=False)
    ctx.soap_model(h, Qs, al)
    E, W = ctx.compute(path)
    en, F = ctx.results()
    Eo, eno, Fo, Wo, _ = orc.compute(sysm, [(h, Qs, al)])
    fs = max(1.0, np.abs(Fo).max())
    assert abs(E - Eo) <= 1e-8 * abs(Eo), (E, Eo)
    np.testing.assert_allclose(en, eno, rtol=1e-8, atol=1e-10)
    assert np.abs(F - Fo).max() <= 1e-6 * fs, np.abs(F - Fo).max()
    assert np.abs(W - Wo).max() <= 1e-6 * max(1.0, np.abs(Wo).max())


def _binary_compressed(dim=300, seed=31):
    sysm, models = H.subset(300, 4096)
    out = []
    for k, (h, _, _) in enumerate(models):
        rng = np.random.default_rng(seed + k)
        P_i = np.repeat(np.arange(1, dim + 1), 3)
        P_j = rng.integers(1, h.full_dim + 1, size=P_i.shape[0])
        P_el = rng.normal(size=P_i.shape[0])
        hc = h.with_compression(dim, P_i, P_j, P_el)
        from oracle import oracle as O
        Qs, al = S.sparse_set(hc, 300, O.block_descriptors, lambda s: S.binary_system(1024, seed=s),
                              seed=7001 + 100 * k)
        out.append((hc, Qs, al))
    return sysm, out


@pytest.mark.parametrize("path", ["fp64_plain", "fp64_adjoint"])
def test_two_species_general_matrix(ctx, orc, path):
    sysm, models = _binary_compressed()
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(3.7, fetch=False)
    H.load_models(ctx, models)
    E, W = ctx.compute(path)
    en, F = ctx.results()
    Eo, eno, Fo, Wo, _ = orc.compute(sysm, models)
    assert abs(E - Eo) <= 1e-8 * abs(Eo)
    np.testing.assert_allclose(en, eno, rtol=1e-8, atol=1e-12)
    assert np.abs(F - Fo).max() <= 1e-6 and np.abs(W - Wo).max() <= 1e-6


def test_mixed_path_refuses_compressed_block(ctx):
    sysm, h, Qs, al = H.silicon_compressed()
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(h.rcut_max, fetch=False)
    ctx.soap_model(h, Qs, al)
    with pytest.raises(TgxError, match="FP64 paths only"):
        ctx.compute("mixed")
