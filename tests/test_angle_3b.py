"""angle_3b, the three-body GAP family (src/gap.f90:724-1112).

CPU (oracle pins):
  * the silicon potential's angle_3b block loads with the reference's
    defaults (buffer 0.5, e0 0) and its 200-point sparse set;
  * cov_pp is (1 - r)^4 (4 r + 1) clamped to [0, 1] and cov_pp_der is its
    derivative divided by r (central differences, and the closed form
    -20 (1 - r)^3);
  * the list-driven oracle's energies equal an independent evaluation from
    positions over every periodic image (1e-12 of the summed term sizes), for
    "exp" and "pol" kernels, on silicon and on a two-species system whose
    centre species differs from the pair (species selection both ways round);
  * forces are minus the gradient of that independent energy and the virial is
    -dE/deps (central differences, h = 1e-5 A);
  * a perfect silicon crystal gives zero forces and an isotropic virial.
GPU (parity): the library's kernel against the oracle on the same lists: the
silicon set (whose alphas reach 4e4 and cancel; energies to 1e-12 of the
summed term sizes), well-conditioned "exp" and "pol" sets, a two-species
system, and a dense case whose centres have more than 128 pairs inside the
cutoff (the large-centre pass): energies to 1e-12 relative, forces and virial
to 1e-11 of their largest entry.
"""
import os

import numpy as np
import pytest

from oracle import angle_3b as A3
from oracle import oracle as O
from turbogap_amd import gap_file as G
from turbogap_amd import structures as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SI = 14


def silicon_crystal(cells=2, perturb=0.05, seed=4):
    a = 5.43
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    grid = np.array([[i, j, k] for i in range(cells) for j in range(cells) for k in range(cells)])
    pos = ((grid[:, None, :] + base[None, :, :]).reshape(-1, 3)) * a
    if perturb:
        pos = pos + np.random.default_rng(seed).normal(scale=perturb, size=pos.shape)
    return S.System(pos, np.full(pos.shape[0], SI, np.int32), S.cubic_box(cells * a))


def silicon_3b():
    return dict(G.load_pair_blocks(os.path.join(GOLD, "silicon.gap"), base_dir=GOLD))["angle_3b"]


def synthetic_3b(kernel_type, sc=SI, s1=SI, s2=SI, rcut=4.0, n=40, seed=21):
    """Sparse points spread over the descriptor range of real triplets, alphas of order one."""
    rng = np.random.default_rng(seed)
    Qs = np.stack([rng.uniform(3.5, 8.0, n), rng.uniform(0.0, 2.0, n), rng.uniform(1.5, 7.5, n)], axis=1)
    sigma = [1.5, 1.0, 1.5] if kernel_type == "exp" else [4.0, 3.0, 4.0]
    return dict(species_center=sc, species1=s1, species2=s2, rcut=rcut, buffer=0.5, delta=0.5, sigma=sigma,
                e0=-0.25, kernel_type=kernel_type, Qs=Qs, alphas=rng.normal(size=n),
                cutoff=rng.uniform(0.5, 1.0, n))


def cancel_scale(kw):
    """Size of the terms a triplet sum can cancel: round-off in any order is relative to it."""
    return max(1.0, 2 * kw["delta"] ** 2 * np.abs(kw["alphas"] * kw["cutoff"]).sum())


def _args(kw):
    return (kw["species_center"], kw["species1"], kw["species2"], kw["rcut"], kw["buffer"], kw["delta"],
            kw["sigma"], kw["e0"], kw["Qs"], kw["alphas"], kw["cutoff"], kw["kernel_type"])


def oracle_eval(kw, sysm):
    nl = O.build_neighbors_list(sysm.positions, sysm.species, sysm.box, kw["rcut"])
    return A3.angle_3b_energy_and_forces(nl, sysm.species, sysm.n, *_args(kw))


def brute_energy(kw, sysm):
    return A3.angle_3b_energy_brute(sysm.positions, sysm.species, sysm.box, *_args(kw))


def two_species(n=256, seed=33):
    sysm = S.binary_system(n, seed=seed)
    return sysm, synthetic_3b("exp", sc=S.SPECIES_B, s1=S.SPECIES_A, s2=S.SPECIES_B, rcut=4.0)


def test_loaded_silicon_angle_3b_block():
    kw = silicon_3b()
    assert kw["species_center"] == kw["species1"] == kw["species2"] == SI
    assert kw["rcut"] == 4.0 and kw["buffer"] == 0.5 and kw["e0"] == 0.0 and kw["delta"] == 0.01
    assert kw["sigma"] == [4.0, 4.0, 4.0] and kw["kernel_type"] == "pol"
    assert kw["Qs"].shape == (200, 3) and kw["alphas"].shape == kw["cutoff"].shape == (200,)
    first = [float(t) for t in open(os.path.join(GOLD, "gap_files",
                                                 "silicon.xml.sparseX.GAP_2020_10_15_180_13_48_39_9742")).read().split()[:3]]
    assert np.array_equal(kw["Qs"][0], first)


def test_cov_pp_is_the_compact_polynomial():
    r = np.concatenate([[-0.5, 0.0, 1.0, 1.5], np.random.default_rng(3).uniform(1e-3, 1 - 1e-3, 200)])
    c, d = A3.cov_pp(r), A3.cov_pp_der(r)
    assert c[0] == c[1] == 1.0 and c[2] == c[3] == 0.0 and not d[:4].any()
    x = r[4:]
    np.testing.assert_allclose(c[4:], (1 - x) ** 4 * (4 * x + 1), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(d[4:], -20 * (1 - x) ** 3, rtol=1e-9, atol=1e-12)
    h = 1e-6
    fd = (A3.cov_pp(x + h) - A3.cov_pp(x - h)) / (2 * h) / x
    np.testing.assert_allclose(d[4:], fd, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("case", ["silicon", "exp", "pol", "two_species"])
def test_oracle_energy_matches_brute_force(case):
    if case == "two_species":
        sysm, kw = two_species()
    else:
        sysm = silicon_crystal()
        kw = silicon_3b() if case == "silicon" else synthetic_3b(case)
    E, _, _ = oracle_eval(kw, sysm)
    Eb = brute_energy(kw, sysm)
    assert np.abs(E - kw["e0"]).max() > 1e-6 * cancel_scale(kw)  # triplets contribute
    if case == "two_species":  # centres of the other species keep e0
        assert np.all(E[sysm.species != kw["species_center"]] == kw["e0"])
    assert np.abs(E - Eb).max() <= 1e-12 * cancel_scale(kw)


@pytest.mark.parametrize("case", ["exp", "pol", "two_species"])
def test_forces_and_virial_are_energy_derivatives(case):
    if case == "two_species":
        sysm, kw = two_species(n=128, seed=5)
    else:
        sysm, kw = silicon_crystal(), synthetic_3b(case)
    _, F, W = oracle_eval(kw, sysm)
    h = 1e-5
    rng = np.random.default_rng(2)
    for i in rng.choice(sysm.n, 4, replace=False):
        for c in range(3):
            ep = []
            for s in (1, -1):
                p = sysm.positions.copy()
                p[i, c] += s * h
                ep.append(brute_energy(kw, S.System(p, sysm.species, sysm.box)).sum())
            assert abs(-(ep[0] - ep[1]) / (2 * h) - F[i, c]) < 1e-6 * max(1.0, abs(F[i, c])), (i, c)
    for c in range(3):
        ep = []
        for s in (1, -1):
            d = np.eye(3)
            d[c, c] += s * h
            ep.append(brute_energy(kw, S.System(sysm.positions @ d, sysm.species, sysm.box @ d)).sum())
        assert abs(-(ep[0] - ep[1]) / (2 * h) - W[c, c]) < 1e-5 * max(1.0, abs(W[c, c])), c


@pytest.mark.parametrize("which", ["silicon", "pol"])
def test_perfect_crystal(which):
    kw = silicon_3b() if which == "silicon" else synthetic_3b("pol")
    _, F, W = oracle_eval(kw, silicon_crystal(perturb=0.0))
    scale = max(1.0, np.abs(W).max(), cancel_scale(kw))
    assert np.abs(F).max() < 1e-12 * scale
    assert np.abs(W - np.diag(np.diag(W))).max() < 1e-12 * scale and np.ptp(np.diag(W)) < 1e-12 * scale


def _gpu_vs_oracle(ctx, kw, sysm, scale=1.0):
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(kw["rcut"], fetch=False)
    e, f, w = ctx.pair_energy_and_forces(ctx.angle_3b_model(**kw))
    Eo, Fo, Wo = oracle_eval(kw, sysm)
    np.testing.assert_allclose(e, Eo, rtol=1e-12, atol=1e-12 * scale)
    # forces and virial relative to their own largest entry (the virial of a
    # 216-atom cell reaches 3e4 eV)
    assert np.abs(f - Fo).max() <= 1e-11 * max(scale, np.abs(Fo).max())
    assert np.abs(w - Wo).max() <= 1e-11 * max(scale, np.abs(Wo).max())


@pytest.mark.gpu
def test_gpu_angle_3b_silicon_set(ctx):
    kw = silicon_3b()
    _gpu_vs_oracle(ctx, kw, silicon_crystal(cells=3), cancel_scale(kw))


@pytest.mark.gpu
@pytest.mark.parametrize("kernel_type", ["exp", "pol"])
def test_gpu_angle_3b_matches_oracle(ctx, kernel_type):
    _gpu_vs_oracle(ctx, synthetic_3b(kernel_type), silicon_crystal(cells=3))


@pytest.mark.gpu
def test_gpu_angle_3b_two_species(ctx):
    sysm, kw = two_species(n=256, seed=9)
    _gpu_vs_oracle(ctx, kw, sysm)


@pytest.mark.gpu
def test_gpu_angle_3b_large_centres(ctx):
    """Four centres of a third species in the two-species system with a 7 A
    cutoff: about 140 selected pairs each, more than the first pass keeps in
    LDS, so they take the large-centre pass; every other atom only receives
    forces."""
    sysm = S.binary_system(512, seed=12)
    sp = sysm.species.copy()
    sp[[3, 100, 250, 400]] = 3
    sysm = S.System(sysm.positions, sp, sysm.box)
    kw = synthetic_3b("pol", sc=3, s1=S.SPECIES_A, s2=S.SPECIES_B, rcut=7.0, n=8)
    nl = O.build_neighbors_list(sysm.positions, sysm.species, sysm.box, kw["rcut"])
    assert nl["n_neigh"][sp == 3].min() - 4 > 128
    _gpu_vs_oracle(ctx, kw, sysm)
