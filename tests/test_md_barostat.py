"""Fixed atoms, pressure and the Berendsen barostat of the device MD
(src/md.f90:226-283, 384-417; src/turbogap_md.f90:287-295, 413-426).

CPU (oracle pins):
  * the velocity-Verlet restatement, driven the way the MD loop drives it,
    follows the exact trajectory of a harmonic oscillator with second-order
    error (halving dt divides it by four) and keeps fixed components where they
    started, with zero velocity;
  * the pressure's virial term tr(W)/3V is -dE/dV of the 2b oracle under
    isotropic scaling (central differences), and the eV/A^3 -> bar constant is
    e * 1e25 to 1e-8;
  * the iso barostat scales the volume by exactly 1 + dt/tau 4.5e-5/gamma
    (P - P0) and expands the box when P > P0; diag scales each axis by its own
    component.
GPU (parity): the library's velocity Verlet with fixed components against the
oracle fed the library's own forces, four steps, to 1e-12; its pressure and
barostat (iso and diag) against the oracle on the device state; the one-call
MD advance against the separate calls, bit for bit; a short NPT run.
"""
import numpy as np
import pytest

from oracle import md_oracle as MDO
from oracle import oracle as O
from oracle import pair_oracle as PO
from tests import helpers as H
from turbogap_amd import lib as L
from turbogap_amd import structures as S
from turbogap_amd.md import MD, maxwell_velocities

from tests.test_pair_models import silicon_crystal, well_conditioned_2b


def test_vv_follows_a_harmonic_oscillator():
    """Driven as the MD loop drives it (forces at x(t), then the update), the
    positions after t = 20 match the exact solution with the second-order error
    of velocity Verlet: halving dt divides it by four."""
    rng = np.random.default_rng(0)
    n = 5
    x0, v0 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    m, k, t_end = np.full(n, 2.0), 3.0, 20.0
    w = np.sqrt(k / m[0])
    fix = np.zeros((n, 3), bool)
    fix[1, 2] = fix[3, 0] = True
    free = ~fix
    exact = x0 * np.cos(w * t_end) + v0 / w * np.sin(w * t_end)
    err = []
    for dt in (0.02, 0.01):
        x, v, Fp = x0.copy(), v0.copy(), np.zeros((n, 3))
        for s in range(int(round(t_end / dt))):
            x, _, v, Fp = MDO.velocity_verlet(x, v, -k * x, Fp, m, dt, dt, s == 0, fix)
        err.append(np.abs(x[free] - exact[free]).max())
        assert np.array_equal(x[fix], x0[fix]) and not v[fix].any()
    assert err[1] < 1e-3 and 3.6 < err[0] / err[1] < 4.4, err


def test_pressure_virial_term_is_minus_dE_dV():
    kw = well_conditioned_2b()
    sysm = silicon_crystal()

    def energy(s):
        p, b = sysm.positions * s, sysm.box * s
        nl = O.build_neighbors_list(p, sysm.species, b, kw["rcut"])
        return PO.distance_2b_energy_and_forces(nl, sysm.species, sysm.n, kw["species1"], kw["species2"], kw["rcut"],
                                                kw["buffer"], kw["delta"], kw["sigma"], kw["e0"], kw["Qs"],
                                                kw["alphas"], kw["cutoff"])
    E, _, W = energy(1.0)
    V = abs(np.linalg.det(sysm.box))
    h = 1e-5
    dEdV = (energy(1 + h)[0].sum() - energy(1 - h)[0].sum()) / (V * ((1 + h) ** 3 - (1 - h) ** 3))
    p, pt = MDO.pressure(W, 0.0, sysm.n, V)
    assert abs(p / MDO.EV_PER_A3_TO_BAR - (-dEdV)) < 1e-6 * max(1.0, abs(dEdV))
    assert abs(np.trace(pt) / 3 - p) < 1e-9 * max(1.0, abs(p))
    assert abs(MDO.EV_PER_A3_TO_BAR / (1.602176634e-19 * 1e30 / 1e5) - 1) < 1e-8
    # the kinetic term alone: P V = (N - 1) kB T
    p_kin, _ = MDO.pressure(np.zeros((3, 3)), 500.0, 1001, 1.0e4)
    assert abs(p_kin * 1.0e4 / MDO.EV_PER_A3_TO_BAR - 1000 * MDO.KB * 500.0) < 1e-12


def test_barostat_scaling():
    box = np.diag([10.0, 11.0, 12.0])
    P = np.diag([2.0e4, 1.0e4, -5.0e3])
    dt, tau, gamma, P0 = 1.0, 100.0, 2.0, 3.0e3
    iso = MDO.berendsen_barostat(box, P0, P, "iso", tau, gamma, dt)
    ratio = np.linalg.det(iso) / np.linalg.det(box)
    assert abs(ratio - (1 + dt / tau * 4.5e-5 / gamma * (np.trace(P) / 3 - P0))) < 1e-14
    assert ratio > 1.0  # mean pressure above P0: the box grows
    diag = MDO.berendsen_barostat(box, P0, P, "diag", tau, gamma, dt)
    for i in range(3):
        assert abs((diag[i, i] / box[i, i]) ** 3 - (1 + dt / tau * 4.5e-5 / gamma * (P[i, i] - P0))) < 1e-14
    assert diag[2, 2] < box[2, 2] and diag[0, 0] > box[0, 0]


def _md_system(ctx, T=600.0, seed=3):
    sysm, h, Qs, al = H.carbon216(300)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.soap_model(h, Qs, al)
    mass = S.masses_for(sysm.species)
    return sysm, mass, maxwell_velocities(mass, T, seed=seed)


@pytest.mark.gpu
def test_gpu_fixed_atoms_match_oracle(ctx):
    sysm, mass, v0 = _md_system(ctx)
    fix = np.random.default_rng(5).random((sysm.n, 3)) < 0.2
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", fix_atom=fix)
    x, v, Fp = sysm.positions.copy(), v0.copy(), np.zeros((sysm.n, 3))
    dt = 1.0
    for s in range(4):
        md.forces()
        _, F = ctx.results()
        L.check(L.load().tgx_velocity_verlet(ctx.h, dt, int(s == 0)), "velocity_verlet")
        x, _, v, Fp = MDO.velocity_verlet(x, v, F, Fp, mass, dt, dt, s == 0, fix)
        xd, vd = md.state()
        np.testing.assert_allclose(xd, x, rtol=0, atol=1e-12)
        np.testing.assert_allclose(vd, v, rtol=0, atol=1e-13)
    assert np.array_equal(xd[fix], sysm.positions[fix]) and not vd[fix].any()
    assert np.abs(xd[~fix] - sysm.positions[~fix]).max() > 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("sym", ["iso", "diag"])
def test_gpu_pressure_and_barostat_match_oracle(ctx, sym):
    sysm, mass, v0 = _md_system(ctx)
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", box=sysm.box)
    md.forces()
    L.check(L.load().tgx_velocity_verlet(ctx.h, 1.0, 1), "velocity_verlet")
    _, T = md.kinetic()
    p, pt = md.pressure_of(T)
    V = abs(np.linalg.det(sysm.box))
    po, pto = MDO.pressure(md.virial, T, sysm.n, V)
    assert abs(p - po) <= 1e-12 * abs(po) and np.abs(pt - pto).max() <= 1e-12 * np.abs(pto).max()
    x_before, _ = md.state()
    P0 = float(np.trace(pt) / 3) - 2.0e4
    nb = np.empty(9)
    L.check(L.load().tgx_berendsen_barostat(ctx.h, P0, L.dptr(np.ascontiguousarray(pt).reshape(9)), sym.encode(),
                                            500.0, 1.0, 1.0, L.dptr(nb)), "berendsen_barostat")
    x_after, _ = md.state()
    np.testing.assert_allclose(x_after, MDO.berendsen_barostat(x_before, P0, pt, sym, 500.0, 1.0, 1.0), rtol=1e-14,
                               atol=1e-13)
    np.testing.assert_allclose(nb.reshape(3, 3), MDO.berendsen_barostat(sysm.box, P0, pt, sym, 500.0, 1.0, 1.0),
                               rtol=1e-14, atol=0)
    if sym == "iso":
        assert np.linalg.det(nb.reshape(3, 3)) > V  # P above P0: expansion


@pytest.mark.gpu
def test_gpu_md_advance_matches_separate_calls(ctx):
    """tgx_md_advance (one synchronisation per step) against velocity Verlet,
    kinetic energy, Berendsen, wrap and the displacement test called one by one:
    the same positions, velocities, temperatures and rebuild decisions, bit for
    bit, over NVT steps with a neighbour buffer."""
    sysm, mass, v0 = _md_system(ctx, T=900.0, seed=8)
    out = []
    for fused in (True, False):
        ctx.set_positions(sysm.positions)  # same start; the MD object resets velocities and forces_prev
        md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", neighbors_buffer=0.2, fused=fused)
        Ts, builds = [], []
        for _ in range(8):
            E, ek, T = md.step(1.0, T0=600.0, tau=20.0)
            Ts.append((E, ek, T))
            builds.append(md.n_builds)
        x, v = md.state()
        out.append((Ts, builds, x, v))
    (T1, b1, x1, v1), (T2, b2, x2, v2) = out
    assert T1 == T2 and b1 == b2 and b1[-1] > 1
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)


@pytest.mark.gpu
def test_gpu_npt_steps(ctx):
    """Berendsen NPT on the device: the box follows the pressure's sign each
    step, the list is refreshed against the current box, energies stay finite."""
    sysm, mass, v0 = _md_system(ctx, T=300.0)
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", neighbors_buffer=0.3, box=sysm.box)
    vol = abs(np.linalg.det(sysm.box))
    for _ in range(6):
        E, ek, T = md.step(1.0, T0=300.0, tau=100.0, P0=0.0, tau_p=200.0, gamma_p=1.0)
        new = abs(np.linalg.det(md.box))
        assert np.isfinite(E) and np.isfinite(md.pressure)
        assert (new - vol) * md.pressure > 0  # grows under positive pressure, shrinks under negative
        vol = new


def test_box_scaling_oracle():
    """box_scaling (src/md.f90:421-458): one step with a single off-diagonal
    entry moves only the x component of positions, by gamma_xy y / n_steps, and
    the box vectors by f . (a0, b0, c0); after n_steps the box is gamma . box0
    exactly and the positions (1 + (g - 1)/n)^n x0 for gamma = g I; the input's
    one, three or nine numbers fill the matrix as the reference reads them."""
    from turbogap_amd.md import box_scaling_factor
    rng = np.random.default_rng(7)
    x0 = rng.uniform(0, 10, (6, 3))
    box0 = np.array([[10.0, 0.5, 0.0], [0.2, 11.0, 0.3], [0.0, -0.4, 12.0]])
    G = np.eye(3)
    G[0, 1] = 0.06
    n = 4
    x1, b1 = MDO.box_scaling(x0, box0, G, 0, n)
    assert np.array_equal(x1[:, 1:], x0[:, 1:])
    np.testing.assert_allclose(x1[:, 0], x0[:, 0] + 0.06 * x0[:, 1] / n, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(b1[:, 0], box0[:, 0] + 0.06 * box0[:, 1] / n, rtol=1e-15)
    assert np.array_equal(b1[:, 1:], box0[:, 1:])
    Gm = np.array([[1.03, 0.01, -0.02], [0.005, 0.98, 0.0], [0.01, 0.02, 1.05]])
    x = x0
    for i in range(n):
        x, b = MDO.box_scaling(x, box0, Gm, i, n)
    np.testing.assert_allclose(b, box0 @ Gm.T, rtol=1e-14, atol=1e-14)
    x = x0
    for i in range(n):
        x, b = MDO.box_scaling(x, box0, 1.02 * np.eye(3), i, n)
    np.testing.assert_allclose(x, x0 * (1 + 0.02 / n) ** n, rtol=1e-14)
    np.testing.assert_allclose(b, 1.02 * box0, rtol=1e-14)
    assert np.array_equal(box_scaling_factor([1.1]), 1.1 * np.eye(3))
    assert np.array_equal(box_scaling_factor([1.0, 2.0, 3.0]), np.diag([1.0, 2.0, 3.0]))
    nine = box_scaling_factor(np.arange(9.0))
    assert nine[0, 1] == 1.0 and nine[1, 0] == 3.0
    with pytest.raises(ValueError):
        box_scaling_factor([1.0, 2.0])


@pytest.mark.gpu
def test_gpu_box_scaling_matches_oracle(ctx):
    """The library's box_scaling against the oracle on the device positions,
    three steps of a non-symmetric gamma (no dynamics between them), to 1e-13;
    a call at a later step before any step 0 is an error."""
    sysm, mass, v0 = _md_system(ctx)
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint")
    lib = L.load()
    G = np.array([[1.03, 0.01, -0.02], [0.005, 0.98, 0.0], [0.01, 0.02, 1.05]])
    g = np.ascontiguousarray(G.reshape(9))
    nb = np.empty(9)
    assert lib.tgx_box_scaling(ctx.h, L.dptr(g), 1, 3, L.dptr(nb)) != 0
    x = sysm.positions.copy()
    for i in range(3):
        L.check(lib.tgx_box_scaling(ctx.h, L.dptr(g), i, 3, L.dptr(nb)), "box_scaling")
        x, bo = MDO.box_scaling(x, sysm.box, G, i, 3)
        xd, _ = md.state()
        np.testing.assert_allclose(xd, x, rtol=0, atol=1e-13)
        np.testing.assert_allclose(nb.reshape(3, 3), bo, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(nb.reshape(3, 3), sysm.box @ G.T, rtol=1e-14, atol=1e-14)


@pytest.mark.gpu
def test_gpu_md_with_box_scaling(ctx):
    """An MD run that strains the box by 3 % over its steps: the box ends at
    gamma . box0, the energies stay finite and the neighbour list follows."""
    sysm, mass, v0 = _md_system(ctx, T=300.0)
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", neighbors_buffer=0.3, box=sysm.box)
    G = np.diag([1.03, 1.03, 1.03])
    for _ in range(5):
        E, ek, T = md.step(1.0, T0=300.0, tau=100.0, scale_box=(G, 5))
        assert np.isfinite(E) and np.isfinite(T)
    np.testing.assert_allclose(md.box, sysm.box @ G.T, rtol=1e-14)


def test_remove_cm_vel_oracle():
    """The centre-of-mass momentum after remove_cm_vel is zero, relative motion
    is kept, and a uniform drift is removed exactly."""
    rng = np.random.default_rng(1)
    v, m = rng.normal(size=(30, 3)), rng.uniform(1, 50, 30)
    w = MDO.remove_cm_vel(v, m)
    assert np.abs((m[:, None] * w).sum(0)).max() < 1e-12
    np.testing.assert_allclose(w[1:] - w[0], v[1:] - v[0], rtol=0, atol=1e-13)
    u = MDO.remove_cm_vel(np.tile([0.3, -0.2, 0.1], (30, 1)), m)
    assert np.abs(u).max() < 1e-15


@pytest.mark.gpu
def test_gpu_remove_cm_vel_matches_oracle(ctx):
    sysm, mass, v0 = _md_system(ctx)
    drift = v0 + np.array([0.01, -0.02, 0.005])
    md = MD(ctx, mass, drift, 3.7, "fp64_adjoint")
    md.remove_cm()
    _, v = md.state()
    np.testing.assert_allclose(v, MDO.remove_cm_vel(drift, mass), rtol=0, atol=1e-15)
    assert np.abs((mass[:, None] * v).sum(0)).max() < 1e-12


@pytest.mark.gpu
def test_gpu_md_state_belongs_to_its_system(ctx):
    """After tgx_set_system to a system of another size, the MD entry points and
    the result copies fail with an error instead of using buffers sized for the
    previous system; a new tgx_md_set_state and force evaluation restore them, and
    a fixed-atom mask of the previous system is dropped."""
    sysm, mass, v0 = _md_system(ctx)
    fix = np.zeros((sysm.n, 3), bool)
    fix[:5] = True
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", fix_atom=fix)
    md.forces()
    lib = L.load()
    L.check(lib.tgx_velocity_verlet(ctx.h, 1.0, 1), "velocity_verlet")
    big = S.binary_system(1024, seed=9)
    ctx.set_system(big.positions, np.full(big.n, S.CARBON, np.int32), big.box)
    ek, T = L.C.c_double(), L.C.c_double()
    flag = L.C.c_int()
    for call in (lambda: lib.tgx_velocity_verlet(ctx.h, 1.0, 0),
                 lambda: lib.tgx_kinetic(ctx.h, L.C.byref(ek), L.C.byref(T)),
                 lambda: lib.tgx_remove_cm_vel(ctx.h),
                 lambda: lib.tgx_md_displacement(ctx.h, 0.5, 0, L.C.byref(flag)),
                 lambda: lib.tgx_md_advance(ctx.h, 1.0, 0, 1, 0.0, 100.0, 0.5, None, None, L.C.byref(flag))):
        assert call() == 1  # TGX_ERR_ARG
    with pytest.raises(L.TgxError):
        ctx.results()
    m2 = S.masses_for(np.full(big.n, S.CARBON))
    md2 = MD(ctx, m2, maxwell_velocities(m2, 300.0, seed=4), 3.7, "fp64_adjoint")
    md2.forces()
    x0, _ = md2.state()
    L.check(lib.tgx_velocity_verlet(ctx.h, 1.0, 1), "velocity_verlet")
    x1, v1 = md2.state()
    assert np.abs(x1[:5] - x0[:5]).max() > 0.0 and v1[:5].any()  # the old mask no longer holds atoms 0-4
