"""The GPU path reproduces the committed 216-atom carbon fixture (tests/golden):
neighbour list bit for bit, energies to 1e-8 relative, forces to 1e-6 eV/Å."""
import os

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "carbon216_golden.npz")


@pytest.mark.parametrize("path", ["fp64_plain", "fp64_adjoint"])
def test_gpu_reproduces_golden(ctx, path):
    ref = np.load(GOLDEN)
    sysm, h, Qs, al = H.carbon216(200)
    np.testing.assert_allclose([Qs.sum(), al.sum()], ref["Qs_checksum"], rtol=1e-13)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    nl = ctx.build_neighbors_list(3.7)
    for k in ("n_neigh", "neighbors_list", "rjs", "xyz"):
        assert np.array_equal(nl[k], ref[k]), k
    ctx.soap_model(h, Qs, al)
    E, W = ctx.compute(path)
    en, F = ctx.results()
    np.testing.assert_allclose(en, ref["energies"], rtol=1e-8, atol=1e-12)
    assert np.abs(F - ref["forces"]).max() <= 1e-6
    assert np.abs(W - ref["virial"]).max() <= 1e-6
