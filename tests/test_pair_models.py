"""distance_2b and core_pot (src/gap.f90:400-721).

CPU (oracle pins):
  * the core_pot spline (src/read_files.f90:5454-5510) is the interpolating
    cubic spline with the given end slopes (scipy's CubicSpline with clamped
    or natural ends, 1e-11), and spline / spline_der extrapolate linearly;
  * forces are minus the gradient of the total energy (central differences,
    h = 1e-5 A) and the virial is the strain derivative W = -dE/deps, on a
    perturbed silicon crystal, for the reference suite's silicon core_pot table
    and a well-conditioned distance_2b set;
  * a perfect crystal gives zero forces and an isotropic virial (silicon sets).
GPU (parity): the library's kernels against the oracle on the same lists:
the silicon core_pot and a two-species 2b model (pair selection by species)
to 1e-12 (energies, relative) and 1e-10 (forces, virial); the silicon 2b set,
whose energies are differences of terms up to 5e4 eV, to 1e-12 of the summed
term magnitudes.
"""
import os

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import oracle as O
from oracle import pair_oracle as PO
from turbogap_amd import gap_file as G
from turbogap_amd import structures as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SI = 14


def silicon_crystal(perturb=0.05, seed=4):
    a = 5.43
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                     [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
    cells = np.array([[i, j, k] for i in range(3) for j in range(3) for k in range(3)])
    pos = ((cells[:, None, :] + base[None, :, :]).reshape(-1, 3)) * a
    if perturb:
        pos = pos + np.random.default_rng(seed).normal(scale=perturb, size=pos.shape)
    return S.System(pos, np.full(pos.shape[0], SI, np.int32), S.cubic_box(3 * a))


def silicon_models():
    blocks = G.load_pair_blocks(os.path.join(GOLD, "silicon.gap"), base_dir=GOLD)
    return dict(blocks)


def cancel_scale(kind, kw):
    """Size of the terms a 2b sum cancels (the silicon set's alphas reach 5e4): round-off in
    any summation order is relative to it."""
    return 1.0 if kind == "core_pot" else 64 * kw["delta"] ** 2 * np.abs(kw["alphas"] * kw["cutoff"]).sum()


def oracle_eval(kind, kw, sysm, rc):
    nl = O.build_neighbors_list(sysm.positions, sysm.species, sysm.box, rc)
    if kind == "distance_2b":
        return PO.distance_2b_energy_and_forces(nl, sysm.species, sysm.n, kw["species1"], kw["species2"], kw["rcut"],
                                                kw["buffer"], kw["delta"], kw["sigma"], kw["e0"], kw["Qs"],
                                                kw["alphas"], kw["cutoff"])
    model = PO.core_pot_model(kw["x"], kw["V"], kw["yp1"], kw["ypn"], kw["cutoff"], kw["buffer"])
    return PO.core_pot_energy_and_forces(nl, sysm.species, sysm.n, kw["species1"], kw["species2"], model)


def test_loaded_silicon_blocks():
    m = silicon_models()
    b2, cp = m["distance_2b"], m["core_pot"]
    assert b2["species1"] == b2["species2"] == SI and b2["rcut"] == 5.0 and b2["buffer"] == 0.5
    assert b2["Qs"].shape == b2["alphas"].shape == b2["cutoff"].shape and b2["Qs"].shape[0] > 10
    assert cp["x"].shape == (26,) and cp["yp1"] == 1.0e50 and cp["ypn"] == 0.0


@pytest.mark.parametrize("ends", ["clamped", "natural"])
def test_core_pot_spline_is_the_interpolating_cubic(ends):
    cp = silicon_models()["core_pot"]
    x, V = cp["x"], cp["V"]
    yp1, ypn = (-3.0, 0.0) if ends == "clamped" else (1e50, 1e50)
    model = PO.core_pot_model(x, V, yp1, ypn)
    bc = ((1, yp1), (1, ypn)) if ends == "clamped" else "natural"
    ref = CubicSpline(x, V, bc_type=bc)
    r = np.random.default_rng(1).uniform(x[0], x[-1] - 1e-9, 200)
    np.testing.assert_allclose([PO.spline(model, t) for t in r], ref(r), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose([PO.spline_der(model, t) for t in r], ref(r, 1), rtol=1e-10, atol=1e-10)
    if ends == "natural":  # the slopes a natural end reports are the spline's own
        assert abs(model[3] - ref(x[0], 1)) < 1e-10 and abs(model[4] - ref(x[-1], 1)) < 1e-10
    # linear extrapolation below the table
    assert abs(PO.spline(model, x[0] - 0.1) - (V[0] - 0.1 * model[3])) < 1e-12


def well_conditioned_2b():
    """A distance_2b model whose energy is not a difference of large terms (the silicon
    set's alphas reach 5e4 and cancel, which limits a finite difference to ~1e-5)."""
    rng = np.random.default_rng(12)
    return dict(species1=SI, species2=SI, rcut=5.0, buffer=0.5, delta=0.8, sigma=0.5, e0=0.0,
                Qs=rng.uniform(2.0, 5.0, 25), alphas=rng.normal(size=25), cutoff=rng.uniform(0.5, 1.0, 25))


@pytest.mark.parametrize("kind", ["distance_2b", "core_pot"])
def test_forces_and_virial_are_energy_derivatives(kind):
    kw = well_conditioned_2b() if kind == "distance_2b" else silicon_models()[kind]
    sysm = silicon_crystal()
    rc = 5.0 if kind == "distance_2b" else float(kw["x"].max())
    E, F, W = oracle_eval(kind, kw, sysm, rc)
    h = 1e-5
    rng = np.random.default_rng(2)
    for i in rng.choice(sysm.n, 4, replace=False):
        for c in range(3):
            ep = []
            for s in (1, -1):
                p = sysm.positions.copy()
                p[i, c] += s * h
                ep.append(oracle_eval(kind, kw, S.System(p, sysm.species, sysm.box), rc)[0].sum())
            assert abs(-(ep[0] - ep[1]) / (2 * h) - F[i, c]) < 1e-6 * max(1.0, abs(F[i, c]))
    for c in range(3):  # W_cc = -dE/deps_cc
        ep = []
        for s in (1, -1):
            d = np.eye(3)
            d[c, c] += s * h
            ep.append(oracle_eval(kind, kw, S.System(sysm.positions @ d, sysm.species, sysm.box @ d), rc)[0].sum())
        assert abs(-(ep[0] - ep[1]) / (2 * h) - W[c, c]) < 1e-5 * max(1.0, abs(W[c, c]))


@pytest.mark.parametrize("kind", ["distance_2b", "core_pot"])
def test_perfect_crystal(kind):
    kw = silicon_models()[kind]
    sysm = silicon_crystal(perturb=0.0)
    rc = 5.0 if kind == "distance_2b" else float(kw["x"].max())
    E, F, W = oracle_eval(kind, kw, sysm, rc)
    scale = max(1.0, np.abs(W).max(), cancel_scale(kind, kw))
    assert np.abs(F).max() < 1e-12 * scale
    assert np.abs(W - np.diag(np.diag(W))).max() < 1e-12 * scale and np.ptp(np.diag(W)) < 1e-12 * scale


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["distance_2b", "core_pot"])
def test_gpu_pair_model_matches_oracle(ctx, kind):
    kw = silicon_models()[kind]
    sysm = silicon_crystal()
    rc = 5.0 if kind == "distance_2b" else float(kw["x"].max())
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(rc, fetch=False)
    mid = ctx.distance_2b_model(**kw) if kind == "distance_2b" else ctx.core_pot_model(**kw)
    e, f, w = ctx.pair_energy_and_forces(mid)
    Eo, Fo, Wo = oracle_eval(kind, kw, sysm, rc)
    scale = cancel_scale(kind, kw)
    np.testing.assert_allclose(e, Eo, rtol=1e-12, atol=1e-12 * scale)
    assert np.abs(f - Fo).max() <= 1e-10 * scale
    assert np.abs(w - Wo).max() <= 1e-10 * scale


@pytest.mark.gpu
def test_gpu_distance_2b_two_species(ctx):
    """Pairs of A with B only (both directions), A-A and B-B ignored; e0 on every centre."""
    sysm = S.binary_system(2048, seed=77)
    rng = np.random.default_rng(8)
    kw = dict(species1=S.SPECIES_A, species2=S.SPECIES_B, rcut=4.0, buffer=0.5, delta=0.7, sigma=0.4, e0=-0.1,
              Qs=rng.uniform(1.2, 4.0, 30), alphas=rng.normal(size=30), cutoff=rng.uniform(0.5, 1.0, 30))
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(4.0, fetch=False)
    e, f, w = ctx.pair_energy_and_forces(ctx.distance_2b_model(**kw))
    Eo, Fo, Wo = oracle_eval("distance_2b", kw, sysm, 4.0)
    np.testing.assert_allclose(e, Eo, rtol=1e-12, atol=1e-12)
    assert np.abs(f - Fo).max() <= 1e-10 and np.abs(w - Wo).max() <= 1e-10
