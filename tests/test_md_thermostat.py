"""The Bussi thermostat and the target-temperature ramp of the device MD
(src/turbogap_md.f90:430-449, src/third_party/bussi_thermostat/resamplekin.f90:50-84,
src/md.f90:1094-1128).

CPU (oracle pins):
  * resample limits: tau/dt huge keeps K (the memory factor is 1); tau/dt <= 0.1
    draws K' = sigma (chi2 + g^2)/ndeg afresh, sigma = ndeg kB T0 / 2;
  * the canonical distribution is stationary: K drawn from Gamma(ndeg/2, kB T0)
    and resampled keeps its mean sigma and variance 2 sigma^2/ndeg (a dropped or
    mis-scaled cross term changes the variance), and from a K far from sigma the
    mean relaxes by exactly (1 - c)(sigma - K), c = exp(-dt/tau);
  * target_temperature: the linear ramp and a t_hold plateau, by hand.
GPU (parity): the library's Bussi step against the oracle's resample of the
device kinetic energy with the same draws: velocities to 1e-14 relative, the
new kinetic energy equal to K'; zero kinetic energy zeroes the velocities;
a short Bussi NVT run.
"""
import numpy as np
import pytest

from oracle import md_oracle as MDO
from tests import helpers as H
from turbogap_amd import lib as L
from turbogap_amd import structures as S
from turbogap_amd.md import MD, maxwell_velocities, target_temperature


def test_bussi_resample_limits():
    ndeg, T0, kk = 297, 500.0, 12.5
    sigma = ndeg * MDO.KB * T0 / 2.0
    assert MDO.bussi_resample(kk, T0, ndeg, 1e300, 1.0, 0.7, 250.0) == kk
    for tau in (0.0, 0.05):
        k = MDO.bussi_resample(kk, T0, ndeg, tau, 1.0, 0.7, 250.0)
        assert abs(k - sigma * (250.0 + 0.49) / ndeg) < 1e-14 * k


def test_bussi_keeps_the_canonical_distribution():
    rng = np.random.default_rng(11)
    ndeg, T0, n = 60, 800.0, 400000
    sigma = ndeg * MDO.KB * T0 / 2.0
    tau, dt = 5.0, 1.0
    kk = rng.gamma(ndeg / 2.0, MDO.KB * T0, n)
    g = rng.standard_normal(n)
    chi2 = rng.chisquare(ndeg - 1, n)
    k2 = MDO.bussi_resample(kk, T0, ndeg, tau, dt, g, chi2)
    se = sigma * np.sqrt(2.0 / ndeg) / np.sqrt(n)
    assert abs(k2.mean() - sigma) < 5 * se
    var = 2.0 * sigma ** 2 / ndeg
    assert abs(k2.var() / var - 1.0) < 0.02
    # from K = 3 sigma the expected K' is K + (1 - c)(sigma - K)
    c = np.exp(-dt / tau)
    k0 = 3.0 * sigma
    k3 = MDO.bussi_resample(k0, T0, ndeg, tau, dt, g, chi2)
    expect = k0 + (1 - c) * (sigma - k0)
    assert abs(k3.mean() - expect) < 5 * k3.std() / np.sqrt(n)


def test_target_temperature_ramp_and_hold():
    assert target_temperature(300.0, 1300.0, 49, 100) == 800.0
    assert target_temperature(300.0, 1300.0, 99, 100) == 1300.0
    hold = [0.0, 20.0, 60.0]
    assert target_temperature(300.0, 1300.0, 9, 100, hold) == 400.0   # before the plateau: the ramp
    assert target_temperature(300.0, 1300.0, 39, 100, hold) == 500.0   # on the plateau: the value at n_start
    assert target_temperature(300.0, 1300.0, 59, 100, hold) == 500.0
    assert abs(target_temperature(300.0, 1300.0, 79, 100, hold) - 900.0) < 1e-12
    assert target_temperature(300.0, 1300.0, 99, 100, hold) == 1300.0
    for i in (9, 39, 79):
        assert target_temperature(300.0, 1300.0, i, 100, hold) == MDO.target_temperature(300.0, 1300.0, i, 100, hold)


def _md_system(ctx, T=600.0, seed=3):
    sysm, h, Qs, al = H.carbon216(300)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.soap_model(h, Qs, al)
    mass = S.masses_for(sysm.species)
    return sysm, mass, maxwell_velocities(mass, T, seed=seed)


@pytest.mark.gpu
def test_gpu_bussi_matches_oracle(ctx):
    sysm, mass, v0 = _md_system(ctx)
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint")
    lib = L.load()
    kk, _ = md.kinetic()
    ndeg = 3 * sysm.n - 3
    g, chi2 = -0.83, float(ndeg - 1) * 1.07
    knew = np.zeros(1)
    L.check(lib.tgx_bussi_thermostat(ctx.h, 300.0, 20.0, 1.0, g, chi2, L.dptr(knew)), "bussi_thermostat")
    ko = MDO.bussi_resample(kk, 300.0, ndeg, 20.0, 1.0, g, chi2)
    assert abs(knew[0] - ko) <= 1e-15 * ko
    _, v = md.state()
    np.testing.assert_allclose(v, v0 * np.sqrt(ko / kk), rtol=1e-14, atol=0)
    k2, _ = md.kinetic()
    assert abs(k2 - ko) < 1e-12 * ko
    md2 = MD(ctx, mass, np.zeros_like(v0), 3.7, "fp64_adjoint")
    L.check(lib.tgx_bussi_thermostat(ctx.h, 300.0, 20.0, 1.0, g, chi2, L.dptr(knew)), "bussi_thermostat")
    _, v = md2.state()
    assert knew[0] == 0.0 and not v.any()


@pytest.mark.gpu
def test_gpu_bussi_nvt_run(ctx):
    """Ten Bussi steps from 900 K towards 300 K with tau = 10 fs: the run stays
    finite and cools."""
    sysm, mass, v0 = _md_system(ctx, T=900.0, seed=6)
    md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", neighbors_buffer=0.3, thermostat="bussi", thermostat_seed=4)
    T = None
    for i in range(10):
        E, ek, T = md.step(1.0, T0=target_temperature(300.0, 300.0, i, 10), tau=10.0)
        assert np.isfinite(E) and np.isfinite(T)
    _, Tend = md.kinetic()
    assert Tend < 700.0
