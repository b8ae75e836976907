"""The oracle reproduces the committed 216-atom carbon fixture (tests/golden):
neighbour list bit for bit, energies/forces/virial to 1e-12."""
import os

import numpy as np

from tests.golden import make_golden

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "carbon216_golden.npz")


def test_oracle_reproduces_golden(orc):
    ref = np.load(GOLDEN)
    got = make_golden.build()
    for k in ("n_neigh", "neighbors_list", "rjs", "xyz"):
        assert np.array_equal(got[k], ref[k]), k
    np.testing.assert_allclose(got["Qs_checksum"], ref["Qs_checksum"], rtol=1e-13)
    for k in ("energies", "forces", "virial", "energy", "soap_sites"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-12, atol=1e-14)
