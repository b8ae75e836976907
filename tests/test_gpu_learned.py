"""GPU tests of the learned path and the device MD.

  * MLP: energies, input gradients, losses and weight gradients of the library's
    kernels against oracle/mlp_oracle.py with the same BF16 operand rounding
    (relative tolerance 1e-2 on each tensor: BF16 roundings of intermediate
    activations can differ by one BF16 ulp where FP32 and FP64 accumulations
    straddle a rounding boundary).
  * Surrogate: trained on FP64-GAP MD environments of a 4,096-atom box, it
    beats the untrained network on a held-out snapshot (energy MAE below
    0.7x, descriptor-gradient RMSE below 0.5x); with the ensemble gate
    forced on every centre the surrogate path returns the FP64 GAP results.
  * MD: one velocity-Verlet step and a Berendsen rescale against hand
    arithmetic (src/md.f90:226-313).
"""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from tests import helpers as H
from turbogap_amd import structures as S
from turbogap_amd import surrogate as SG
from turbogap_amd.md import MD, maxwell_velocities, KB

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


def test_mlp_kernels_match_oracle(ctx):
    rng = np.random.default_rng(0)
    D, Hd, B = 300, 128, 128
    net = SG.MLP(ctx, D, Hd, 256, seed=3)
    w = {"W1": rng.normal(size=(Hd, D)).astype(np.float32), "b1": (0.1 * rng.normal(size=Hd)).astype(np.float32),
         "W2": (rng.normal(size=(Hd, Hd)) / np.sqrt(Hd)).astype(np.float32),
         "b2": (0.1 * rng.normal(size=Hd)).astype(np.float32),
         "w3": (rng.normal(size=Hd) / np.sqrt(Hd)).astype(np.float32), "b3": np.array([0.02], np.float32)}
    net.set_weights(w)
    q = rng.normal(size=(B, D))
    q /= np.linalg.norm(q, axis=1)[:, None]
    q = q.astype(np.float32)
    e_ref = rng.normal(size=B).astype(np.float32)
    g_ref = rng.normal(size=(B, D)).astype(np.float32) * 0.3
    e, g = net.forward(q)
    f = MO.forward(w, q, emulate_bf16=True)
    assert _rel(e, f["E"]) < 1e-2 and _rel(g, f["g"]) < 1e-2
    grads, le, lg = net.gradients(q, e_ref, g_ref, 0.7, 1.3)
    le_o, lg_o, go = MO.loss_and_grads(w, q.astype(np.float64), e_ref.astype(np.float64),
                                       g_ref.astype(np.float64), 0.7, 1.3, emulate_bf16=True)
    assert abs(le - le_o) <= 1e-2 * le_o and abs(lg - lg_o) <= 1e-2 * lg_o
    for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
        assert _rel(grads[k], go[k]) < 2e-2, (k, _rel(grads[k], go[k]))


def _small_models(ctx, n_sparse=300):
    sysm, models = H.subset(n_sparse, 4096)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ids = H.load_models(ctx, models)
    return sysm, models, ids


def test_surrogate_learns_and_gate_falls_back(ctx):
    sysm, models, ids = _small_models(ctx)
    D = models[0][0].dim
    nets = [[SG.MLP(ctx, D, 256, 1024, seed=11 + k)] for k in range(2)]
    untrained = [[SG.MLP(ctx, D, 256, 1024, seed=11 + k)] for k in range(2)]
    SG.md_dataset(ctx, ids, sysm, nets, temperatures=(500.0, 2500.0), steps=10)
    n_train = [nets[k][0].add_samples(ids[k], 1 << 30, 1 << 30) for k in range(2)]
    # held-out snapshot at 1500 K appended after the training samples
    mass = S.masses_for(sysm.species)
    md = MD(ctx, mass, maxwell_velocities(mass, 1500.0, seed=5), 3.7, "fp64_plain")
    for _ in range(10):
        md.step(1.0, T0=1500.0)
    md.forces("fp64_plain")
    for k in range(2):
        nets[k][0].add_samples(ids[k], 1, 0)
        untrained[k][0].add_samples(ids[k], 1, 0)
    n_all = [nets[k][0].add_samples(ids[k], 1 << 30, 1 << 30) for k in range(2)]
    SG.train_ensemble(nets, n_train, epochs=100, batch=1024)
    for k in range(2):
        mu, sd = nets[k][0].dataset_stats(n_train[k])
        untrained[k][0].set_normalization(mu, sd)
        ev = nets[k][0].evaluate(n_train[k], n_all[k])
        ev0 = untrained[k][0].evaluate(0, n_all[k] - n_train[k])
        print(f"block {k}: trained {ev}, untrained {ev0}", flush=True)
        assert ev["e_mae"] < 0.7 * ev0["e_mae"] and ev["g_rmse"] < 0.5 * ev0["g_rmse"], (k, ev, ev0)
        nets[k][0].install(ids[k], 0)
        untrained[k][0].install(ids[k], 1)
    held, _ = md.state()
    ctx.set_positions(held)
    ctx.build_neighbors_list(3.7, fetch=False)
    E64, W64 = ctx.compute("fp64_adjoint")
    e64, F64 = ctx.results()
    SG.set_gate(ctx, 1, 1e9)
    Es, Ws = ctx.compute("surrogate")
    es, Fs = ctx.results()
    assert np.isfinite(es).all() and np.sqrt(np.mean((Fs - F64) ** 2)) < 0.05
    # two very different members and a zero threshold: every centre falls back to the FP64 GAP
    SG.set_gate(ctx, 2, 1e-12)
    Eg, Wg = ctx.compute("surrogate")
    eg, Fg = ctx.results()
    frac, spread = SG.gate_stats(ctx)
    assert frac == 1.0 and spread > 0
    np.testing.assert_allclose(eg, e64, atol=1e-10)
    assert np.abs(Fg - F64).max() < 1e-4


def test_velocity_verlet_and_berendsen(ctx):
    sysm, h, Qs, al = H.carbon216(300)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.soap_model(h, Qs, al)
    mass = S.masses_for(sysm.species)
    v0 = maxwell_velocities(mass, 800.0, seed=3)
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint")
    dt = 1.0
    md.forces()
    _, F0 = ctx.results()
    from turbogap_amd import lib as L
    L.check(L.load().tgx_velocity_verlet(ctx.h, dt, 1))
    x1, v1 = md.state()
    np.testing.assert_allclose(x1, sysm.positions + v0 * dt + 0.5 * F0 / mass[:, None] * dt * dt, atol=1e-12)
    np.testing.assert_allclose(v1, v0, atol=0)
    md.forces()
    _, F1 = ctx.results()
    L.check(L.load().tgx_velocity_verlet(ctx.h, dt, 0))
    x2, v2 = md.state()
    v1b = v0 + 0.5 * (F0 + F1) / mass[:, None] * dt
    np.testing.assert_allclose(v2, v1b, atol=1e-13)
    ek, T = md.kinetic()
    ek_ref = 0.5 * np.sum(mass[:, None] * v1b ** 2)
    assert abs(ek - ek_ref) < 1e-12 * ek_ref
    assert abs(T - 2.0 / 3.0 / (sysm.n - 1) / KB * ek_ref) < 1e-9 * T
    L.check(L.load().tgx_berendsen(ctx.h, 300.0, T, 50.0, dt))
    _, v3 = md.state()
    np.testing.assert_allclose(v3, v1b * np.sqrt(1 + dt / 50.0 * (300.0 / T - 1)), rtol=1e-13)


def _carbon216_block(ctx):
    sysm, h, Qs, al = H.carbon216(200)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(h.rcut_max, fetch=False)
    return sysm, h, ctx.soap_model(h, Qs, al)


def test_surrogate_energies_match_mlp_oracle(ctx):
    """One installed network: the fused inference (SiLU epilogue on the first
    GEMM) gives e_i = shift + scale E(q_i) of the oracle network on the FP64
    descriptors (BF16 operands: 2e-2 of the output range)."""
    sysm, h, mid = _carbon216_block(ctx)
    net = SG.MLP(ctx, h.dim, 128, 256, seed=21)
    net.set_normalization(-0.5, 0.03)
    net.install(mid, 0)
    SG.set_gate(ctx, 1, 1e9)
    ctx.compute("surrogate")
    en, F = ctx.results()
    ctx.soap_backend_begin(mid)
    q, _ = ctx.get_soap(derivatives=False)
    ctx.soap_backend_end()
    f = MO.forward(net.get_weights(), q, emulate_bf16=True)
    e_ref = -0.5 + 0.03 * f["E"]
    assert np.isfinite(F).all()
    assert np.abs(en - e_ref).max() <= 2e-2 * np.abs(e_ref + 0.5).max()


def test_md_neighbour_buffer_matches_rebuild_every_step(ctx):
    """neighbors_buffer = 0.5 A (src/turbogap_md.f90:451-471, turbogap_setup.f90:285): the list is
    built at rcut + 0.5 and refreshed between builds; 30 steps at 1500 K give the energies of the
    run that rebuilds every step to round-off, with fewer builds."""
    sysm, models = H.subset(300, 4096)
    mass = S.masses_for(sysm.species)
    out = []
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    H.load_models(ctx, models)
    for buf in (0.0, 0.5):
        ctx.set_positions(sysm.positions)
        md = MD(ctx, mass, maxwell_velocities(mass, 1500.0, seed=7), 3.7, "fp64_adjoint", neighbors_buffer=buf)
        es = [md.step(1.0)[0] for _ in range(30)]
        x, v = md.state()
        out.append((np.array(es), x, md.n_builds))
    (e0, x0, nb0), (e1, x1, nb1) = out
    assert nb0 == 30 and 1 <= nb1 < 30
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    assert np.abs(x1 - x0).max() <= 1e-9


def test_surrogate_65k_atoms_match_mlp_oracle(ctx):
    """65,536 atoms, one network per block (the benchmark's GEMM row counts):
    site energies of each block against the oracle network on the FP64
    descriptors, as above."""
    _, models = H.subset(500, 4096)
    sysm = S.binary_system(65536, seed=2001)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(3.7, fetch=False)
    ids = H.load_models(ctx, models)
    nets = []
    for k, (mid, (h, _, _)) in enumerate(zip(ids, models)):
        net = SG.MLP(ctx, h.dim, 256, 256, seed=31 + k)
        net.set_normalization(-0.5, 0.03)
        net.install(mid, 0)
        nets.append(net)
    SG.set_gate(ctx, 1, 1e9)
    ctx.compute("surrogate")
    en, F = ctx.results()
    assert np.isfinite(F).all()
    for k, (mid, (h, _, _)) in enumerate(zip(ids, models)):
        ctx.soap_backend_begin(mid)
        q, _ = ctx.get_soap(derivatives=False)
        ctx.soap_backend_end()
        e_ref = -0.5 + 0.03 * MO.forward(nets[k].get_weights(), q, emulate_bf16=True)["E"]
        e_blk = en[sysm.species == h.central_species]
        assert e_blk.shape == e_ref.shape
        assert np.abs(e_blk - e_ref).max() <= 2e-2 * np.abs(e_ref + 0.5).max()


def test_surrogate_block_streams_match_serial(ctx):
    """The two blocks (central species A and B) on streams of their own give
    bit-identical energies, forces and virial to the same step with the blocks
    in turn, also with the ensemble gate sending sites to the FP64 fallback,
    and on the mixed and both FP64 paths."""
    _, models = H.subset(500, 1024)
    sysm = S.binary_system(8192, seed=2001)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(3.7, fetch=False)
    ids = H.load_models(ctx, models)
    for k, (mid, (h, _, _)) in enumerate(zip(ids, models)):
        for i in range(2):
            net = SG.MLP(ctx, h.dim, 128, 256, seed=41 + 7 * k + i)
            net.set_normalization(-0.5, 0.03 + 0.01 * i)
            net.install(mid, i)
    SG.set_gate(ctx, 2, 1e9)
    ctx.compute("surrogate")
    half = 0.5 * SG.gate_stats(ctx)[1]  # gates the sites whose members differ most
    for threshold in (1e9, half):
        SG.set_gate(ctx, 2, threshold)
        out = []
        for on in (True, False):
            ctx.set_block_streams(on)
            E, W = ctx.compute("surrogate")
            en, F = ctx.results()
            out.append((E, W, en, F, SG.gate_stats(ctx)))
        ctx.set_block_streams(True)
        (E1, W1, e1, F1, g1), (E2, W2, e2, F2, g2) = out
        assert E1 == E2 and np.array_equal(W1, W2) and np.array_equal(e1, e2) and np.array_equal(F1, F2)
        assert g1 == g2
    assert 0.0 < g1[0] < 1.0  # the second threshold gates part of the sites
    for path in ("mixed", "fp64_adjoint", "fp64_plain"):  # the other paths (mixed: forced)
        out = []
        for on in (2, 0):
            ctx.set_block_streams(on)
            E, W = ctx.compute(path)
            out.append((E, W) + ctx.results())
        ctx.set_block_streams(True)
        (E1, W1, e1, F1), (E2, W2, e2, F2) = out
        assert E1 == E2 and np.array_equal(W1, W2) and np.array_equal(e1, e2) and np.array_equal(F1, F2), path


def test_surrogate_ensemble_is_member_mean(ctx):
    """Two members with different normalisations: energies are the mean of the
    single-member runs, the spread their difference, and forces (linear in
    dE/dq) the mean of the single-member forces."""
    sysm, h, mid = _carbon216_block(ctx)
    a = SG.MLP(ctx, h.dim, 128, 256, seed=21)
    b = SG.MLP(ctx, h.dim, 128, 256, seed=22)
    a.set_normalization(-0.5, 0.03)
    b.set_normalization(-0.4, 0.05)
    SG.set_gate(ctx, 1, 1e9)
    out = []
    for net in (a, b):
        net.install(mid, 0)
        ctx.compute("surrogate")
        out.append(ctx.results())
    a.install(mid, 0)
    b.install(mid, 1)
    SG.set_gate(ctx, 2, 1e9)
    ctx.compute("surrogate")
    e2, F2 = ctx.results()
    frac, spread = SG.gate_stats(ctx)
    (ea, Fa), (eb, Fb) = out
    np.testing.assert_allclose(e2, 0.5 * (ea + eb), rtol=0, atol=1e-12)
    assert frac == 0.0
    np.testing.assert_allclose(spread, np.abs(ea - eb).max(), rtol=1e-5)
    assert np.abs(F2 - 0.5 * (Fa + Fb)).max() <= 1e-4 * np.abs(F2).max()


def _gate_threshold(ctx, lo_frac, hi_frac):
    """A threshold at which the gate sends between lo_frac and hi_frac of the
    sites at the current positions to the FP64 GAP (bisection on the spread)."""
    SG.set_gate(ctx, 3, 1e9)
    ctx.compute("surrogate")
    lo, hi = 0.0, SG.gate_stats(ctx)[1]
    for _ in range(40):
        t = 0.5 * (lo + hi)
        SG.set_gate(ctx, 3, t)
        ctx.compute("surrogate")
        f = SG.gate_stats(ctx)[0]
        if f > hi_frac:
            lo = t
        elif f < lo_frac:
            hi = t
        else:
            return t, f
    raise AssertionError("no threshold gates the requested share of sites")


def test_surrogate_production_md_with_gate(ctx):
    """Production MD with the surrogate (the 65,536-atom run's loop on the
    4,096-atom two-species box): a 3-member ensemble per block trained on FP64
    GAP MD environments, the gate threshold set so that 10-25 % of the sites go
    to the FP64 GAP at the start, 30 steps of Berendsen NVT at 1500 K then 30
    of NVE (dt = 1 fs, neighbour list at rcut + 0.5 A), an FP64 GAP check every
    20 steps on the same positions as the surrogate evaluation.
      * Gated sites carry the FP64 GAP energies: at every check, the sites whose
        surrogate energy equals the GAP's to 1e-10 eV are exactly the gated
        share the library reports, and that share is not zero.
      * Energies and forces stay finite on every step.
      * Check-step force RMSE against the FP64 GAP below 50 meV/A (the
        benchmark's bound); the surrogate's own error here is a few meV/A.
      * NVE: the total energy's linear drift below 20 meV/atom/ps and its range
        below 2 meV/atom (the check steps, on the GAP's surface, excluded)."""
    sysm, models, ids = _small_models(ctx)
    D = models[0][0].dim
    nets = [[SG.MLP(ctx, D, 256, 1024, seed=51 + 3 * k + i) for i in range(3)] for k in range(2)]
    SG.md_dataset(ctx, ids, sysm, nets, temperatures=(500.0, 1500.0, 2500.0), steps=10)
    n_train = [nets[k][0].add_samples(ids[k], 1 << 30, 1 << 30) for k in range(2)]
    SG.train_ensemble(nets, n_train, epochs=60, batch=1024)
    for k in range(2):
        for i in range(3):
            nets[k][i].install(ids[k], i)
    ctx.set_positions(sysm.positions)
    ctx.build_neighbors_list(3.7 + 0.5, fetch=False)
    thr, f0 = _gate_threshold(ctx, 0.10, 0.25)
    SG.set_gate(ctx, 3, thr)
    mass = S.masses_for(sysm.species)
    md = MD(ctx, mass, maxwell_velocities(mass, 1500.0, seed=8), 3.7, "surrogate", neighbors_buffer=0.5)
    n = sysm.n
    etot, fracs, checks = [], [], []
    for k in range(60):
        T0 = 1500.0 if k < 30 else None
        if (k + 1) % 20 == 0:
            md.forces("surrogate")
            frac = SG.gate_stats(ctx)[0]
            es, Fs = ctx.results()
            E, ek, T = md.step(1.0, T0=T0, tau=100.0, path="fp64_adjoint")
            e64, F64 = ctx.results()
            exact = int(np.sum(np.abs(es - e64) <= 1e-10))
            checks.append((frac, exact, float(np.sqrt(np.mean((Fs - F64) ** 2)))))
            assert np.isfinite(Fs).all() and np.isfinite(es).all()
        else:
            E, ek, T = md.step(1.0, T0=T0, tau=100.0)
            fracs.append(SG.gate_stats(ctx)[0])
            en, F = ctx.results()
            assert np.isfinite(F).all() and np.isfinite(en).all() and np.isfinite(E)
            if k >= 30:
                etot.append((k, (E + ek) / n))
    print(f"gate threshold {1e3 * thr:.3f} meV (start share {f0:.3f}); checks (share, exact sites, "
          f"force RMSE eV/A) {checks}; mean share {np.mean(fracs):.3f}", flush=True)
    for frac, exact, rmse in checks:
        assert exact == int(round(frac * n)) and exact > 0, (frac, exact)
        assert rmse < 0.05, rmse
    assert 0.0 < np.mean(fracs) < 0.6
    t_ps = np.array([k for k, _ in etot]) * 1e-3
    e = np.array([v for _, v in etot])
    drift = np.polyfit(t_ps, e, 1)[0] * 1e3
    print(f"NVE: drift {drift:.4f} meV/atom/ps, range {1e3 * np.ptp(e):.4f} meV/atom", flush=True)
    assert abs(drift) < 20.0 and 1e3 * np.ptp(e) < 2.0


def _block_entries(nl, species, n, h):
    """The list entries of one block in the library's order: centres are the
    atoms of the block's central species in atom order, each with its list
    entries inside the block's cutoff (self first); returns site_beg and the
    0-based atom of every entry."""
    c = h.species_types[h.central_species - 1] if h.central_species > 0 else None
    beg, nbr = [0], []
    off = np.concatenate([[0], np.cumsum(nl["n_neigh"])])
    for i in range(n):
        if c is not None and species[i] != c:
            continue
        for p in range(off[i], off[i + 1]):
            if nl["rjs"][p] < h.rcut_max and int(nl["neighbor_species"][p]) in h.species_types:
                nbr.append((nl["neighbors_list"][p] - 1) % n)
        beg.append(len(nbr))
    return np.array(beg), np.array(nbr)


def test_force_loss_gradients_match_oracle(ctx, orc):
    """Energy + force loss on a whole two-species structure (512 atoms, both
    blocks, D = 324): the library's losses and weight gradients (its FP64 dq/dr,
    pair-force kernel, fixed-order gather and the force-direction kernel, then
    the networks' forward-then-reverse pass) against oracle/mlp_oracle.py's
    force_loss_and_grads fed the same descriptors and dq/dr (BF16 operands
    emulated; the tolerances of the environment-loss test)."""
    rng = np.random.default_rng(11)
    sysm = S.binary_system(512, seed=2024)
    models = []
    for k, h in enumerate(S.soap_hypers([S.SPECIES_A, S.SPECIES_B], central_species=c, n_max=4) for c in (1, 2)):
        Qs, al = S.sparse_set(h, 96, orc.block_descriptors, lambda s: S.binary_system(256, seed=s), seed=501 + k)
        models.append((h, Qs, al))
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    nl = ctx.build_neighbors_list(3.7)
    ids = H.load_models(ctx, models)
    blocks = []
    for mid, (h, Qs, al) in zip(ids, models):
        ctx.soap_backend_begin(mid)
        q, J = ctx.get_soap(derivatives=True)
        ctx.soap_backend_end()
        K = q @ Qs.T
        e_ref = h.delta ** 2 * (K ** h.zeta) @ al
        g_ref = h.delta ** 2 * h.zeta * ((K ** (h.zeta - 1)) * al) @ Qs
        beg, nbr = _block_entries(nl, sysm.species, sysm.n, h)
        assert beg[-1] == J.shape[0] and len(beg) - 1 == q.shape[0]
        blocks.append({"q": q.astype(np.float32).astype(np.float64), "e_ref": e_ref, "g_ref": g_ref, "J": J,
                       "site_beg": beg, "nbr": nbr})
    ctx.compute("fp64_plain")  # the block state the force loss reads (q, E, -Qss, dq/dr)
    nets = []
    D, Hd = models[0][0].dim, 128
    for k, b in enumerate(blocks):
        net = SG.MLP(ctx, D, Hd, 512, seed=5 + k)
        w = {"W1": (rng.normal(size=(Hd, D)) * 0.5).astype(np.float32),
             "b1": (0.1 * rng.normal(size=Hd)).astype(np.float32),
             "W2": (rng.normal(size=(Hd, Hd)) / np.sqrt(Hd)).astype(np.float32),
             "b2": (0.1 * rng.normal(size=Hd)).astype(np.float32),
             "w3": (rng.normal(size=Hd) / np.sqrt(Hd)).astype(np.float32), "b3": np.array([0.02], np.float32)}
        net.set_weights(w)
        shift, scale = float(np.mean(b["e_ref"])), 0.05 + 0.02 * k
        net.set_normalization(shift, scale)
        b.update(w={kk: v.astype(np.float64) for kk, v in w.items()}, shift=shift, scale=scale)
        nets.append(net)
    le, lf = SG.force_gradients(ctx, nets, ids, 0.8, 1.4)
    grads = [net.get_gradients() for net in nets]
    le_o, lf_o, go = MO.force_loss_and_grads(blocks, sysm.n, 0.8, 1.4, emulate_bf16=True)
    assert abs(le - le_o) <= 1e-2 * le_o and abs(lf - lf_o) <= 1e-2 * lf_o, (le, le_o, lf, lf_o)
    for k in range(2):
        for name in ("W1", "b1", "W2", "b2", "w3", "b3"):
            assert _rel(grads[k][name], go[k][name]) < 2e-2, (k, name, _rel(grads[k][name], go[k][name]))
