"""CPU restatement of TurboGAP's MD integrator pieces (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
device MD; the product path never imports it.

  * velocity_verlet follows src/md.f90:226-283: velocities t-dt -> t with the
    mean of the two forces (not on the first step; a fixed component's
    velocity is set to zero instead), positions_prev and forces_prev copied,
    positions t -> t+dt except fixed components.
  * berendsen_barostat follows src/md.f90:384-417: positions (or lattice
    vectors, as columns) scaled by (1 + dt/tau 4.5e-5/gamma (P - P0))^(1/3),
    P the mean diagonal pressure ("iso") or per component ("diag").
  * pressure follows src/turbogap_md.f90:287-295 (bar; eVperA3tobar =
    1602176.6208 as in src/turbogap.f90:95).
  * box_scaling follows src/md.f90:421-458 (indices fixed at 1): positions +=
    (gamma - I) positions / n_steps, box vectors f . box0 vectors with
    f = I + (gamma - I)(i_step + 1)/n_steps.
  * bussi_resample follows resamplekin (src/third_party/bussi_thermostat/
    resamplekin.f90:50-84; Bussi, Donadio and Parrinello, J. Chem. Phys. 126,
    014101 (2007), eq. A7) with its two random draws as arguments: the Gaussian
    of gasdev() and the chi-squared (ndeg - 1) of sumnoises(ndeg - 1).
  * GradientDescent follows gradient_descent (src/md.f90:540-664): the first
    step max_opt_step / max |F| (or the accepted step of an earlier line
    search), backtracking by halves until the Armijo test E <= E0 - gamma/2
    sum F^2 holds, then Barzilai-Borwein steps |s.y|/|y|^2 with positions_prev
    taken to the minimum image (orthorhombic box).
  * variable_time_step follows src/md.f90:464-535: from dt, steps of 0.99
    while the largest |v dt + 0.5 F/m dt^2| exceeds target_pos_step (or 1.01
    while below it, whichever held at dt), then min(new_dt, dt0) at init, else
    the relaxed min(dt tau_dt/(tau_dt + dt - new_dt), dt0).
  * remove_cm_vel follows src/md.f90:316-345.
  * adaptive_time_step follows variable_time_step_adaptive
    (src/stopping/adaptive_time.f90:46-164), its loop over atoms as written:
    the velocity and force times of the last atom that had them carry over,
    and min(dtmin, tmax) overwrites max(dtmin, tmin).
  * target_temperature follows get_target_temp (src/md.f90:1094-1128), the
    t_beg -> t_end ramp with its t_hold plateaus.
Arrays are (n, 3) here; the reference's are (3, n).
"""
from __future__ import annotations

import numpy as np

KB = 8.6173303e-5
EV_PER_A3_TO_BAR = 1602176.6208


def velocity_verlet(positions, velocities, forces, forces_prev, masses, dt, dt_prev, first_step, fix_atom=None):
    """Returns (positions, positions_prev, velocities, forces_prev)."""
    x = np.array(positions, dtype=np.float64)
    v = np.array(velocities, dtype=np.float64)
    F = np.asarray(forces, dtype=np.float64)
    Fp = np.asarray(forces_prev, dtype=np.float64)
    m = np.asarray(masses, dtype=np.float64)[:, None]
    fix = np.zeros(x.shape, bool) if fix_atom is None else np.asarray(fix_atom, bool)
    if not first_step:
        v = np.where(fix, 0.0, v + 0.5 * (F + Fp) / m * dt_prev)
    x_prev = x.copy()
    x = np.where(fix, x, x + v * dt + 0.5 * F / m * dt ** 2)
    return x, x_prev, v, F.copy()


def berendsen_barostat(positions, P0, P, sym, tau, gamma, dt):
    """positions (n, 3) (or the lattice vectors as rows); P (3, 3) in bar."""
    x = np.array(positions, dtype=np.float64)
    P = np.asarray(P, dtype=np.float64)
    if sym[:3] == "iso":
        P_iso = (P[0, 0] + P[1, 1] + P[2, 2]) / 3.0
        return x * (1.0 + dt / tau * 4.5e-5 / gamma * (P_iso - P0)) ** (1.0 / 3.0)
    if sym[:4] == "diag":
        f = np.array([(1.0 + dt / tau * 4.5e-5 / gamma * (P[i, i] - P0)) ** (1.0 / 3.0) for i in range(3)])
        return x * f[None, :]
    raise ValueError(sym)


def pressure(virial, temperature, n_sites, volume):
    """(instant_pressure, instant_pressure_tensor) in bar."""
    W = np.asarray(virial, dtype=np.float64).reshape(3, 3)
    kin = KB * (n_sites - 1) * temperature
    p = (kin + np.trace(W) / 3.0) / volume * EV_PER_A3_TO_BAR
    pt = W / volume * EV_PER_A3_TO_BAR + np.eye(3) * kin / volume * EV_PER_A3_TO_BAR
    return p, pt


def box_scaling(positions, box0, gamma, i_step, n_steps):
    """(positions, box) after step i_step; box0 rows are the vectors a0, b0, c0."""
    x = np.asarray(positions, dtype=np.float64)
    G = np.asarray(gamma, dtype=np.float64).reshape(3, 3)
    I = np.eye(3)
    x = x + (x @ (G - I).T) / float(n_steps)
    f = I + (G - I) * float(i_step + 1) / float(n_steps)
    box = (np.asarray(box0, dtype=np.float64) @ f.T)
    return x, box


def bussi_resample(kk, T0, ndeg, tau, dt, gauss, chi2):
    """The resampled kinetic energy; the operations in the reference's order."""
    sigma = float(np.float32(ndeg)) * KB * T0 / 2.0
    taut = tau / dt
    c = np.exp(-1.0 / taut) if taut > 0.1 else 0.0
    return kk + (1.0 - c) * (sigma * (chi2 + gauss ** 2) / ndeg - kk) + 2.0 * gauss * np.sqrt(
        kk * sigma / ndeg * (1.0 - c) * c)


def target_temperature(t_beg, t_end, md_istep, md_nsteps, t_hold=()):
    """t_hold: flat (T_hold, n_start, n_end) triples; the temperature holds at the
    ramp's value at n_start until n_end, then ramps to t_end by the last step."""
    s = float(md_istep + 1)
    T = t_beg + (t_end - t_beg) * s / float(md_nsteps)
    h = np.asarray(t_hold, dtype=np.float64).reshape(-1, 3)
    for _, n_start, n_end in h:
        if n_start < s <= n_end:
            T = t_beg + (t_end - t_beg) * (n_start / float(md_nsteps))
        elif s >= n_end:
            th = t_beg + (t_end - t_beg) * (n_start / float(md_nsteps))
            T = th + (t_end - th) * (s - n_end) / (float(md_nsteps) - n_end)
    return T


class GradientDescent:
    """State of the reference's save variables; step() mutates the arrays given."""

    def __init__(self):
        self.initialized = False
        self.backtracking = False
        self.energy0 = self.gamma_prev = self.m_prev = self.gamma_back0 = 0.0
        self.x0 = self.f0 = None

    def step(self, x, xprev, F, Fprev, energy, max_opt_step, first_step, box_diag, fix=None):
        """x, xprev, F, Fprev (n, 3) updated in place; returns (gamma, restored)."""
        restored = False
        gamma = 0.0
        if first_step:
            self.backtracking = True
            self.x0, self.f0 = x.copy(), F.copy()
            self.energy0 = energy
            fmax = float(np.sqrt(np.sum(F * F, axis=1)).max())
            if fmax == 0.0:
                gamma = 0.0
            elif self.initialized:
                gamma = self.gamma_back0
            else:
                gamma = max_opt_step / fmax
        elif self.backtracking:
            if energy <= self.energy0 - self.gamma_prev * 0.5 * self.m_prev:
                self.backtracking = False
                self.initialized = True
                self.gamma_back0 = self.gamma_prev
            else:
                gamma = self.gamma_prev * 0.5
                x[:] = self.x0
                F[:] = self.f0
                restored = True
        if not first_step and not self.backtracking:
            L = np.asarray(box_diag, dtype=np.float64)
            d = x - xprev
            d = d - np.rint(d / L) * L
            xprev[:] = x - d
            y = F - Fprev
            syy = float(np.sum(y * y))
            gamma = 0.0 if syy == 0.0 else abs(float(np.sum((x - xprev) * y)) / syy)
        xprev[:] = x
        Fprev[:] = F
        move = np.ones_like(x, dtype=bool) if fix is None else ~np.asarray(fix, dtype=bool)
        x[move] = xprev[move] + gamma * Fprev[move]
        self.gamma_prev = gamma
        self.m_prev = float(np.sum(F * F))
        return gamma, restored


def variable_time_step(init, velocities, forces, masses, target_pos_step, tau_dt, dt0, dt):
    v = np.asarray(velocities, dtype=np.float64)
    a = 0.5 * np.asarray(forces, dtype=np.float64) / np.asarray(masses, dtype=np.float64)[:, None]

    def dmax(c):
        x = v * c + a * (c * c)
        return float(np.sqrt((x[:, 0] * x[:, 0] + x[:, 1] * x[:, 1]) + x[:, 2] * x[:, 2]).max())

    too_large = dmax(dt) > target_pos_step
    new_dt = dt
    while True:
        d = dmax(new_dt)
        if d > target_pos_step and too_large:
            new_dt = new_dt * 0.99
        elif d < target_pos_step and not too_large:
            new_dt = new_dt * 1.01
        else:
            break
    if init:
        return min(new_dt, dt0)
    return min(dt * tau_dt / (tau_dt + dt - new_dt), dt0)


def remove_cm_vel(velocities, masses):
    v = np.array(velocities, dtype=np.float64)
    m = np.asarray(masses, dtype=np.float64)
    return v - (m[:, None] * v).sum(0) / m.sum()


def adaptive_time_step(init, velocities, forces, masses, tmin, tmax, xmax, emax, dt0, dt):
    v = np.asarray(velocities, dtype=np.float64)
    F = np.asarray(forces, dtype=np.float64)
    m = np.asarray(masses, dtype=np.float64)
    dtmin = float(np.float32(1.0e20))
    dtv = dtf = dt0 if init else dt
    for i in range(v.shape[0]):
        vsq = (v[i, 0] * v[i, 0] + v[i, 1] * v[i, 1]) + v[i, 2] * v[i, 2]
        fsq = (F[i, 0] * F[i, 0] + F[i, 1] * F[i, 1]) + F[i, 2] * F[i, 2]
        if vsq > 0.0:
            dtv = xmax / np.sqrt(vsq)
        if fsq > 0.0:
            dtf = np.sqrt(2.0 * xmax * m[i] / np.sqrt(fsq))
        t = min(dtv, dtf)
        if emax > 0.0 and vsq * fsq > 0.0:
            t = min(t, emax / np.sqrt(vsq * fsq))
        x = v[i] * t + 0.5 * F[i] / m[i] * (t * t)
        d = np.sqrt((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2])
        if d > xmax:
            t = t * xmax / d
        dtmin = min(dtmin, t)
    out = min(dtmin, tmax)
    if init:
        out = min(out, dt0)
    return float(out)
