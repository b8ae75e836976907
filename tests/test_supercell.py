"""The library's supercell (host code, no GPU): number_of_unit_cells
(src/neighbors.f90:175-236) and the replication of src/read_files.f90:320-367,
against the oracle and closed forms."""
import numpy as np

from turbogap_amd import structures as S
from turbogap_amd.context import build_supercell, number_of_unit_cells

BOXES = [
    (S.cubic_box(12.91), 3.7),
    (S.cubic_box(86.8), 3.7),
    (np.diag([10.0, 5.0, 3.0]), 3.7),
    (S.cubic_box(7.4), 3.7),  # 2 rc equals the edge
    (np.diag([7.4, 3.7, 2.4666666666666666]), 3.7),  # 2 rc on a multiple of each height
    (np.array([[9.0, 0.0, 0.0], [2.1, 8.5, 0.0], [1.3, -1.7, 9.4]]), 3.0),
    (np.array([[5.2, 0.0, 0.0], [1.4, 4.9, 0.0], [-0.9, 1.1, 5.6]]), 3.2),
    (np.array([[3.567, 0.0, 0.0], [1.7835, 3.0891, 0.0], [1.7835, 1.0297, 2.9124]]), 2.5),
]


def test_unit_cells_match_oracle_and_heights(orc):
    for box, rc in BOXES:
        ind = number_of_unit_cells(box, rc)
        assert ind.tolist() == orc.number_of_unit_cells(box, rc).tolist(), (box, rc)
        # closed form: ceil(2 rc / h) with h the cell height over each face pair
        vol = abs(np.linalg.det(box))
        a, b, c = box
        h = np.array([vol / np.linalg.norm(np.cross(b, c)), vol / np.linalg.norm(np.cross(c, a)),
                      vol / np.linalg.norm(np.cross(a, b))])
        x = 2 * rc / h
        ok = np.ceil(x - 1e-9) <= ind
        assert ok.all() and (ind <= np.ceil(x + 1e-9)).all(), (ind, x)
    assert number_of_unit_cells(np.diag([10.0, 5.0, 3.0]), 3.7).tolist() == [1, 2, 3]


def test_supercell_matches_oracle(orc):
    for box, rc in BOXES[2:]:
        pos = np.random.default_rng(1).random((7, 3)) @ box
        sp = np.arange(7, dtype=np.int32) % 3 + 1
        ind = number_of_unit_cells(box, rc)
        psc, ssc, bsc = build_supercell(pos, sp, box, ind)
        po, so, bo = orc.build_supercell(pos, sp, box, ind)
        assert np.array_equal(psc, po) and np.array_equal(ssc, so) and np.array_equal(bsc, bo)
        m = int(np.prod(ind))
        assert psc.shape == (7 * m, 3) and np.array_equal(psc[:7], pos)
        # copy (i, j, k), i slowest, is the cell shifted by i a + j b + k c
        i, j, k = ind - 1
        last = pos + i * box[0] + j * box[1] + k * box[2]
        np.testing.assert_allclose(psc[-7:], last, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(bsc, box * ind[:, None])
