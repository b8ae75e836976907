"""SOAP compression (compress_soap) on the host and in the oracle, no GPU.

  * The compression file of the reference suite's silicon potential
    (tests/stopping/1/gap_files/compress.dat, the older index-list form) reads
    as src/read_files.f90:5166-5211 reads it; the "trivial" recipe of the
    reference's tools/compress_indices/compress_indices.py, which wrote that
    file, gives the same matrix, and in the two-species case keeps exactly the
    components whose n or n' opens a species block.
  * A P_transformation file reads into its entries.
  * silicon.gap loads all four blocks: the soap_turbo block compressed to 72
    components with its 9,000 alphas (its desc_sparse file is absent from the
    reference tree and is synthesised here), with a SparseSetWarning.
  * Oracle (C10, compression before normalisation): the identity matrix gives
    the uncompressed descriptor bit for bit; a selection gives the renormalised
    selected components of the uncompressed descriptor; a general matrix gives
    q_c = P q / |P q| and dq_c/dr = (I - q_c q_c^T) P dq/dr / |P q| from the
    uncompressed q (the other route to the same numbers, the one the device
    takes); and the compressed derivative matches central finite differences.
"""
import os
import shutil
import warnings

import numpy as np
import pytest

from tests import helpers as H
from turbogap_amd import gap_file as G
from turbogap_amd import structures as S
from turbogap_amd.hypers import compress_trivial, read_compress_file

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_compress_file_and_trivial_recipe():
    dim, P_i, P_j, P_el, am, lm = read_compress_file(os.path.join(GOLD, "gap_files", "compress.dat"), 1)
    assert (dim, am, lm) == (72, [8], 8)
    assert P_i == list(range(1, 73)) and P_el == [1.0] * 72
    assert P_j == list(range(1, 73))
    assert compress_trivial([8], 8) == (dim, P_i, P_j, P_el)


def test_trivial_recipe_two_species():
    am, lm = [6, 4], 5
    dim, P_i, P_j, P_el = compress_trivial(am, lm)
    N = sum(am)
    # decode every component index of the full spectrum (n <= n', l fastest)
    comp = [(n, m, l) for n in range(1, N + 1) for m in range(n, N + 1) for l in range(lm + 1)]
    opens = {1, 1 + am[0]}
    want = [k + 1 for k, (n, m, l) in enumerate(comp) if n in opens or m in opens]
    assert P_j == want and P_i == list(range(1, dim + 1)) and dim == len(want)


def test_p_transformation_file(tmp_path):
    p = tmp_path / "c.dat"
    p.write_text("4 3 5\nP_transformation\n1 2 0.5\n1 7 -1.5d0\n3 40 2.0\n5 1 1\nend_transformation\n")
    dim, P_i, P_j, P_el, am, lm = read_compress_file(str(p), 1)
    assert (dim, am, lm) == (5, [4], 3)
    assert P_i == [1, 1, 3, 5] and P_j == [2, 7, 40, 1] and P_el == [0.5, -1.5, 2.0, 1.0]


def test_silicon_gap_loads_all_blocks(tmp_path):
    shutil.copy(os.path.join(GOLD, "silicon.gap"), tmp_path)
    shutil.copytree(os.path.join(GOLD, "gap_files"), tmp_path / "gap_files")
    rng = np.random.default_rng(3)
    np.savetxt(tmp_path / "gap_files" / "silicon.xml.sparseX.GAP_2020_10_15_180_13_48_39_9743",
               rng.random(9000 * 72), fmt="%.6f")
    with pytest.warns(G.SparseSetWarning):
        soap = G.load_soap_blocks(str(tmp_path / "silicon.gap"), base_dir=str(tmp_path))
    pairs = G.load_pair_blocks(str(tmp_path / "silicon.gap"), base_dir=str(tmp_path))
    assert sorted(k for k, _ in pairs) == ["angle_3b", "core_pot", "distance_2b"]
    ((h, Qs, al),) = soap
    assert h.compressed and h.dim == 72 and h.full_dim == 324
    assert Qs.shape == (9000, 72) and al.shape == (9000,)
    assert h.zeta == 6.0 and h.radial_enhancement == 1 and h.central_species == 1


def _env(seed, n=14):
    """One centre with n neighbours in two species inside 3.7 Å."""
    rng = np.random.default_rng(seed)
    xyz = np.zeros((n + 1, 3))
    k = 1
    while k <= n:
        v = rng.uniform(-3.6, 3.6, 3)
        if 1.0 < np.linalg.norm(v) < 3.6:
            xyz[k] = v
            k += 1
    spec = np.array([0] + [k % 2 for k in range(n)], np.int32)
    return xyz, spec


def _soap(orc, h, xyz, spec, der=True):
    r = np.linalg.norm(xyz, axis=1)
    return orc.get_soap(h, np.array([xyz.shape[0]], np.int32), spec, r, xyz, derivatives=der)


def _random_matrix(h, dim, seed):
    rng = np.random.default_rng(seed)
    P_i = np.repeat(np.arange(1, dim + 1), 3)
    P_j = rng.integers(1, h.full_dim + 1, size=P_i.shape[0])
    P_el = rng.normal(size=P_i.shape[0])
    return P_i, P_j, P_el


def test_oracle_compression_identity_and_selection(orc):
    h = S.soap_hypers([1, 2], n_max=4, l_max=4)
    xyz, spec = _env(1)
    q, d = _soap(orc, h, xyz, spec)
    D = h.full_dim
    hi = h.with_compression(D, np.arange(1, D + 1), np.arange(1, D + 1), np.ones(D))
    qi, di = _soap(orc, hi, xyz, spec)
    assert np.array_equal(qi, q) and np.array_equal(di, d)
    sel = np.random.default_rng(2).choice(D, 60, replace=False) + 1
    hs = h.with_compression(60, np.arange(1, 61), sel, np.ones(60))
    qs, _ = _soap(orc, hs, xyz, spec)
    want = q[0, sel - 1] / np.linalg.norm(q[0, sel - 1])
    np.testing.assert_allclose(qs[0], want, rtol=0, atol=1e-15)


def test_oracle_compression_from_normalised_descriptor(orc):
    h = S.soap_hypers([1, 2], n_max=4, l_max=4)
    xyz, spec = _env(4)
    q, d = _soap(orc, h, xyz, spec)
    P_i, P_j, P_el = _random_matrix(h, 50, 5)
    hc = h.with_compression(50, P_i, P_j, P_el)
    qc, dc = _soap(orc, hc, xyz, spec)
    Pm = np.zeros((50, h.full_dim))
    np.add.at(Pm, (P_i - 1, P_j - 1), P_el)
    y = Pm @ q[0]
    want_q = y / np.linalg.norm(y)
    np.testing.assert_allclose(qc[0], want_q, rtol=0, atol=1e-14)
    proj = (np.eye(50) - np.outer(want_q, want_q)) / np.linalg.norm(y)
    for p in range(d.shape[0]):
        want_d = proj @ Pm @ d[p]
        np.testing.assert_allclose(dc[p], want_d, rtol=0, atol=1e-12 * max(1.0, np.abs(want_d).max()))


def test_oracle_compressed_derivative_fd(orc):
    h = S.soap_hypers([1, 2], n_max=4, l_max=4)
    xyz, spec = _env(7)
    P_i, P_j, P_el = _random_matrix(h, 40, 8)
    hc = h.with_compression(40, P_i, P_j, P_el)
    _, dc = _soap(orc, hc, xyz, spec)
    step = 1e-5
    for p in (1, 5, 9):
        for a in range(3):
            xp, xm = xyz.copy(), xyz.copy()
            xp[p, a] += step
            xm[p, a] -= step
            fd = (_soap(orc, hc, xp, spec, False)[0][0] - _soap(orc, hc, xm, spec, False)[0][0]) / (2 * step)
            np.testing.assert_allclose(dc[p, :, a], fd, rtol=0, atol=2e-8)
