"""CPU pins of the oracle's neighbour list (src/neighbors.f90:47-778).

Checked against things that do not restate the oracle:
  * an O(N^2) numpy minimum-image search (round-based MIC, all images for small
    and triclinic cells): same neighbour sets and distances;
  * the list properties the reference's build guarantees: self first with
    zero displacement, symmetric relation, each cell's atoms visited as one run
    in descending index order (the head/next linked list), angles consistent
    with the displacement;
  * closed forms of number_of_unit_cells and of the per-pair byte model
    (96,512 B per pair at n_max 16, l_max 8, D 1224, two species).
"""
import itertools

import numpy as np
import pytest

from turbogap_amd import structures as S


def brute_pairs(pos, box, rc, n_sites=None):
    """All (i, j, d) with 0 < |r_j - r_i + image| < rc over lattice images."""
    n = pos.shape[0] if n_sites is None else n_sites
    inv = np.linalg.inv(box.T)
    out = {}
    reach = [int(np.ceil(rc / (abs(np.linalg.det(box)) / np.linalg.norm(np.cross(box[(k + 1) % 3], box[(k + 2) % 3])))))
             + 1 for k in range(3)]
    images = np.array(list(itertools.product(*[range(-r, r + 1) for r in reach])), float) @ box
    for i in range(n):
        d = pos[None, :, :] - pos[i][None, None, :] + images[:, None, :]
        dist = np.linalg.norm(d, axis=2)
        m = (dist < rc) & (dist > 1e-9)
        for a, j in zip(*np.nonzero(m)):
            out.setdefault(i, []).append((j, dist[a, j]))
    return out


def _check_against_brute(nl, pos, box, rc, n_fold):
    br = brute_pairs(pos, box, rc)
    beg = np.concatenate([[0], np.cumsum(nl["n_neigh"])])
    for i in range(len(nl["n_neigh"])):
        lst = nl["neighbors_list"][beg[i]:beg[i + 1]]
        assert lst[0] == i + 1 and nl["rjs"][beg[i]] == 0.0
        got = sorted(zip(((lst[1:] - 1) % n_fold).tolist(), nl["rjs"][beg[i] + 1:beg[i + 1]].round(9).tolist()))
        exp = sorted((j, round(d, 9)) for j, d in br.get(i, []))
        assert got == exp, f"atom {i}"


def test_carbon216_matches_brute_force(orc):
    sysm = S.carbon216_system()
    nl = orc.build_neighbors_list(sysm.positions, sysm.species, sysm.box, 3.7)
    _check_against_brute(nl, sysm.positions, sysm.box, 3.7, sysm.n)
    # xyz is the minimum-image displacement and |xyz| = rjs
    assert np.allclose(np.linalg.norm(nl["xyz"], axis=1), nl["rjs"], rtol=0, atol=1e-13)


def test_list_properties_and_cell_order(orc):
    sysm = S.binary_system(2048, seed=77)
    rc = 3.7
    nl = orc.build_neighbors_list(sysm.positions, sysm.species, sysm.box, rc)
    beg = np.concatenate([[0], np.cumsum(nl["n_neigh"])])
    L = sysm.box[0, 0]
    m = int(L / rc)
    cell = np.floor(np.mod(sysm.positions, L) / L * m).astype(int)
    cid = cell[:, 0] + m * cell[:, 1] + m * m * cell[:, 2]
    pairs = set()
    for i in range(sysm.n):
        lst = nl["neighbors_list"][beg[i] + 1:beg[i + 1]] - 1
        for j in lst:
            pairs.add((i, int(j)))
        # runs of one cell, each cell once, indices descending inside a run
        runs = []
        for j in lst:
            if runs and runs[-1][0] == cid[j]:
                assert j < runs[-1][1][-1], "indices must descend within a cell"
                runs[-1][1].append(j)
            else:
                runs.append((cid[j], [j]))
        cells = [c for c, _ in runs]
        assert len(cells) == len(set(cells)), "a cell visited twice"
        # angles consistent with the displacement
        x = nl["xyz"][beg[i] + 1:beg[i + 1]]
        r = nl["rjs"][beg[i] + 1:beg[i + 1]]
        np.testing.assert_allclose(np.cos(nl["thetas"][beg[i] + 1:beg[i + 1]]), x[:, 2] / r, atol=1e-14)
        np.testing.assert_allclose(nl["phis"][beg[i] + 1:beg[i + 1]], np.arctan2(x[:, 1], x[:, 0]), atol=1e-15)
    assert all((j, i) in pairs for i, j in pairs), "neighbour relation must be symmetric"


def test_small_box_supercell_matches_images(orc):
    L = 5.5
    pos = S.rsa(24, L, 1.2, 11)
    sp = np.full(24, 6, np.int32)
    box = S.cubic_box(L)
    ind = orc.number_of_unit_cells(box, 3.7)
    assert ind.tolist() == [2, 2, 2]  # ceil(2 * 3.7 / 5.5)
    psc, ssc, bsc = orc.build_supercell(pos, sp, box, ind)
    nl = orc.build_neighbors_list(psc, ssc, bsc, 3.7, n_sites=24, indices=ind)
    _check_against_brute(nl, pos, box, 3.7, 24)


def test_triclinic_cell_matches_images(orc):
    box = np.array([[9.0, 0.0, 0.0], [2.1, 8.5, 0.0], [1.3, -1.7, 9.4]])
    rng = np.random.default_rng(5)
    frac = rng.random((60, 3))
    pos = frac @ box
    sp = np.full(60, 6, np.int32)
    nl = orc.build_neighbors_list(pos, sp, box, 3.0)
    _check_against_brute(nl, pos, box, 3.0, 60)


def test_unit_cells_closed_form(orc):
    assert orc.number_of_unit_cells(S.cubic_box(12.91), 3.7).tolist() == [1, 1, 1]
    assert orc.number_of_unit_cells(S.cubic_box(86.8), 3.7).tolist() == [1, 1, 1]
    assert orc.number_of_unit_cells(np.diag([10.0, 5.0, 3.0]), 3.7).tolist() == [1, 2, 3]


def test_batch_memory_model_closed_form(orc):
    import ctypes as C
    bp, bs = C.c_double(), C.c_double()
    orc.load().orc_soap_batch_memory_model(8, 16, 1224, 2, C.byref(bp), C.byref(bs))
    k_max = 45
    assert bp.value == 48 * k_max * 16 + 64 * k_max + 48 * 1224 + 16 * 16 + 4 * 2 + 56 == 96512
    assert bs.value == 16 * k_max * 16 + 8 * 1224 + 16


def test_batches_partition_sites(orc):
    import ctypes as C
    sysm = S.carbon216_system()
    nl = orc.build_neighbors_list(sysm.positions, sysm.species, sysm.box, 3.7)
    lib = orc.load()
    n = sysm.n
    for gb in (1e3, 0.05, 0.01):
        nc = lib.orc_get_number_of_atom_pairs(n, nl["n_neigh"].ctypes.data_as(C.POINTER(C.c_int)), nl["n_atom_pairs"],
                                              nl["rjs"].ctypes.data_as(C.POINTER(C.c_double)), 3.7, 8, 8, 324, 1,
                                              gb, 0, None, None, None, None)
        ib = np.zeros(nc, np.int32); ie = np.zeros(nc, np.int32); jb = np.zeros(nc, np.int64); je = np.zeros(nc, np.int64)
        lib.orc_get_number_of_atom_pairs(n, nl["n_neigh"].ctypes.data_as(C.POINTER(C.c_int)), nl["n_atom_pairs"],
                                         nl["rjs"].ctypes.data_as(C.POINTER(C.c_double)), 3.7, 8, 8, 324, 1, gb, nc,
                                         ib.ctypes.data_as(C.POINTER(C.c_int)), ie.ctypes.data_as(C.POINTER(C.c_int)),
                                         jb.ctypes.data_as(C.POINTER(C.c_long)), je.ctypes.data_as(C.POINTER(C.c_long)))
        assert ib[0] == 1 and ie[-1] == n and jb[0] == 1 and je[-1] == nl["n_atom_pairs"]
        assert all(ib[k + 1] == ie[k] + 1 and jb[k + 1] == je[k] + 1 for k in range(nc - 1))
        beg = np.concatenate([[0], np.cumsum(nl["n_neigh"])])
        assert all(je[k] == beg[ie[k]] for k in range(nc))
    assert nc > 1
