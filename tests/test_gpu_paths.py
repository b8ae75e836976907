"""GPU: the adjoint, mixed-precision and finite-difference checks.

The adjoint kernel's forces are checked against central finite differences of
the FP64 energy (h = 1e-4 Å, positions away from the cutoff) and the virial
against a strain finite difference, W = -dE/deps. The mixed path (FP32
descriptors, BF16 MFMA contraction) is held to error bounds against the FP64
oracle: per-atom energy MAE <= 0.1 meV/atom and force RMSE <= 1 meV/Å.
"""
import numpy as np
import pytest

from tests import helpers as H
from turbogap_amd import structures as S

pytestmark = pytest.mark.gpu


def _energy(ctx, pos, box, path):
    ctx.set_positions(pos)
    ctx.build_neighbors_list(3.7, fetch=False)
    E, _ = ctx.compute(path)
    return E


def test_adjoint_forces_finite_difference(ctx):
    sysm, h, Qs, al = H.carbon216(500)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.soap_model(h, Qs, al)
    ctx.build_neighbors_list(3.7, fetch=False)
    ctx.compute("fp64_adjoint")
    _, F = ctx.results()
    hstep = 1e-4
    rng = np.random.default_rng(3)
    for i in rng.choice(sysm.n, 4, replace=False):
        for k in range(3):
            p = sysm.positions.copy()
            p[i, k] += hstep
            ep = _energy(ctx, p, sysm.box, "fp64_adjoint")
            p[i, k] -= 2 * hstep
            em = _energy(ctx, p, sysm.box, "fp64_adjoint")
            fd = -(ep - em) / (2 * hstep)
            assert abs(fd - F[i, k]) <= 2e-6 + 1e-4 * abs(F[i, k]), (i, k, fd, F[i, k])


def test_virial_strain_finite_difference(ctx):
    sysm, h, Qs, al = H.carbon216(500)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.soap_model(h, Qs, al)
    ctx.build_neighbors_list(3.7, fetch=False)
    _, W = ctx.compute("fp64_adjoint")
    eps = 1e-5
    for k in range(3):
        vals = []
        for sgn in (1, -1):
            S = np.eye(3)
            S[k, k] += sgn * eps
            box = sysm.box @ S
            pos = sysm.positions @ S
            ctx.set_system(pos, sysm.species, box)
            ctx.build_neighbors_list(3.7, fetch=False)
            E, _ = ctx.compute("fp64_adjoint")
            vals.append(E)
        fd = -(vals[0] - vals[1]) / (2 * eps)
        assert abs(fd - W[k, k]) <= 1e-5 + 1e-4 * abs(W[k, k]), (k, fd, W[k, k])


def test_mixed_path_error_bounds(ctx, orc):
    sysm, models = H.subset(500, 4096)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(3.7, fetch=False)
    H.load_models(ctx, models)
    E, W = ctx.compute("mixed")
    en, F = ctx.results()
    Eo, eno, Fo, Wo, _ = orc.compute(sysm, models)
    assert np.mean(np.abs(en - eno)) <= 1e-4
    assert np.sqrt(np.mean((F - Fo) ** 2)) <= 1e-3
    assert np.abs(W - Wo).max() <= 0.05


@pytest.mark.parametrize("n_max,l_max", [(6, 6), (5, 4)])
def test_mixed_path_other_basis_sizes(ctx, orc, n_max, l_max):
    """The FP32 path at l_max other than 8 (the descriptor kernel's other template
    instances, the general adjoint) against the FP64 oracle, with the bounds above."""
    from oracle import oracle as O
    sysm = S.carbon216_system()
    h = S.soap_hypers([S.CARBON], n_max=n_max, l_max=l_max)
    Qs, al = S.sparse_set(h, 300, O.block_descriptors, S.carbon216_system, seed=5003)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(h.rcut_max, fetch=False)
    ctx.soap_model(h, Qs, al)
    E, W = ctx.compute("mixed")
    en, F = ctx.results()
    Eo, eno, Fo, Wo, _ = orc.compute(sysm, [(h, Qs, al)])
    assert np.isfinite(F).all()
    assert np.mean(np.abs(en - eno)) <= 1e-4
    assert np.sqrt(np.mean((F - Fo) ** 2)) <= 1e-3


def test_mixed_path_16k_atoms(ctx):
    """16,384 atoms, 4,000 sparse points per block (the contraction shapes of the
    benchmark's blocks at a quarter of the rows): the energy and force bounds
    above against the FP64 adjoint path on the same box; the virial within 1% of
    its largest entry (0.4% measured: a sum over 360k pair terms)."""
    _, models = H.subset(4000, 4096)
    sysm = S.binary_system(16384, seed=2003)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(3.7, fetch=False)
    H.load_models(ctx, models)
    E64, W64 = ctx.compute("fp64_adjoint")
    e64, F64 = ctx.results()
    E, W = ctx.compute("mixed")
    en, F = ctx.results()
    assert np.mean(np.abs(en - e64)) <= 1e-4
    assert np.sqrt(np.mean((F - F64) ** 2)) <= 1e-3
    assert np.abs(W - W64).max() <= 1e-2 * np.abs(W64).max()


def test_fp64_paths_agree_binary_density(ctx):
    """Plain (materialised dq/dr) and adjoint forces agree to round-off."""
    sysm, models = H.subset(300, 4096)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(3.7, fetch=False)
    H.load_models(ctx, models)
    Ep, Wp = ctx.compute("fp64_plain")
    ep, Fp = ctx.results()
    Ea, Wa = ctx.compute("fp64_adjoint")
    ea, Fa = ctx.results()
    assert abs(Ep - Ea) <= 1e-12 * max(1.0, abs(Ep))
    assert np.abs(Fp - Fa).max() <= 1e-11
    assert np.abs(Wp - Wa).max() <= 1e-10


def test_block_streams_above_two_launch_scan_limit(ctx):
    """81,920 atoms (two central species): the block selection scans more values
    than the two-launch scan takes, so each block falls back to the three-launch
    scan with its own block totals. Running the blocks on streams of their own
    (forced on the mixed path, which runs them in turn by default) gives
    bit-identical energies, forces and virial to the blocks in turn."""
    _, models = H.subset(500, 1024)
    sysm = S.binary_system(81920, seed=2003)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(3.7, fetch=False)
    H.load_models(ctx, models)
    for path in ("mixed", "fp64_adjoint"):
        out = []
        for on in (2, 0, 2):
            ctx.set_block_streams(on)
            E, W = ctx.compute(path)
            out.append((E, W) + ctx.results())
        ctx.set_block_streams(True)
        ref = out[1]
        for o in (out[0], out[2]):
            assert o[0] == ref[0] and np.array_equal(o[1], ref[1]), path
            assert np.array_equal(o[2], ref[2]) and np.array_equal(o[3], ref[3]), path
        assert np.isfinite(ref[3]).all() and ref[3].shape == (sysm.n, 3)
