"""Molecular dynamics on the device with the semantics of TurboGAP's md
(src/md.f90:95-345) and its step order (src/turbogap_md.f90:160-470):
forces at x(t) -> velocity_verlet -> E_kin, T -> Berendsen rescale. The
neighbour list is built with rcut + neighbors_buffer (src/turbogap_setup.f90:285)
and rebuilt when an atom has moved more than half the buffer since the last
build (src/turbogap_md.f90:451-471); in between, its distances are recomputed
(rebuild_neighbors_list = .false.). neighbors_buffer = 0 rebuilds every step."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L
from . import structures as S

KB = 8.6173303e-5  # eV/K (src/md.f90:104)


def maxwell_velocities(masses: np.ndarray, T: float, seed: int) -> np.ndarray:
    """randomize_velocities "maxwell" (src/md.f90:126-131): g sqrt(kB T / m), then
    the centre-of-mass velocity removed."""
    n = masses.shape[0]
    g = S.normal(3 * n, 1.0, seed).reshape(n, 3)
    v = g * np.sqrt(KB * T / masses)[:, None]
    cm = (masses[:, None] * v).sum(0) / masses.sum()
    return v - cm


def box_scaling_factor(values) -> np.ndarray:
    """The 3x3 box_scaling_factor from the input's one, three or nine numbers
    (src/read_files.f90:1727-1760): one number on the diagonal, three on the
    diagonal in order, nine row by row."""
    v = np.asarray(values, dtype=np.float64).reshape(-1)
    if v.size == 1:
        return np.eye(3) * v[0]
    if v.size == 3:
        return np.diag(v)
    if v.size == 9:
        return v.reshape(3, 3).copy()
    raise ValueError(f"box_scaling_factor takes 1, 3 or 9 numbers, got {v.size}")


def target_temperature(t_beg: float, t_end: float, md_istep: int, md_nsteps: int, t_hold=()) -> float:
    """The thermostat's target at step md_istep (get_target_temp, src/md.f90:1094-1128):
    the t_beg -> t_end ramp, and for each (T, n_start, n_end) triple of t_hold a
    plateau at the ramp's value at n_start until n_end, after which it ramps to
    t_end by the last step."""
    s = float(md_istep + 1)
    T = t_beg + (t_end - t_beg) * s / float(md_nsteps)
    for _, n_start, n_end in np.asarray(t_hold, dtype=np.float64).reshape(-1, 3):
        start = t_beg + (t_end - t_beg) * (n_start / float(md_nsteps))
        if n_start < s <= n_end:
            T = start
        elif s >= n_end:
            T = start + (t_end - start) * (s - n_end) / (float(md_nsteps) - n_end)
    return T


class MD:
    """fix_atom (n, 3) booleans hold Cartesian components in place
    (velocity_verlet's fix_atom). box (3, 3) rows a, b, c: the system's box, kept
    up to date when a barostat rescales it. thermostat "berendsen" or "bussi"
    (src/turbogap_md.f90:430-449); the Bussi thermostat's random draws come from
    a numpy generator seeded with thermostat_seed."""

    def __init__(self, ctx, masses: np.ndarray, velocities: np.ndarray, rcut: float, path: str = "fp64_plain",
                 neighbors_buffer: float = 0.0, fix_atom=None, box=None, fused: bool = True,
                 thermostat: str = "berendsen", thermostat_seed: int = 0, remove_cm: bool = True):
        if thermostat not in ("berendsen", "bussi"):
            raise ValueError(f"thermostat must be 'berendsen' or 'bussi', got {thermostat!r}")
        self.thermostat = thermostat
        self.remove_cm_each_step = bool(remove_cm)  # remove_cm_vel before each step (src/turbogap_md.f90:172-173)
        self.rng = np.random.default_rng(thermostat_seed)
        self.ctx, self.rcut, self.path = ctx, rcut, path
        self.fused = fused  # steps without a barostat through tgx_md_advance (one synchronisation)
        self.buffer = float(neighbors_buffer)
        self.masses = np.ascontiguousarray(masses, dtype=np.float64)
        v = np.ascontiguousarray(velocities, dtype=np.float64)
        L.check(L.load().tgx_md_set_state(ctx.h, L.dptr(v), L.dptr(self.masses)), "md_set_state")
        if fix_atom is not None:
            fx = np.ascontiguousarray(np.asarray(fix_atom, dtype=bool).reshape(-1, 3), dtype=np.int32)
            if fx.shape[0] != self.masses.shape[0]:
                raise ValueError("fix_atom must be (n_sites, 3)")
            L.check(L.load().tgx_md_set_fixed(ctx.h, L.iptr(fx)), "md_set_fixed")
        else:
            L.check(L.load().tgx_md_set_fixed(ctx.h, None), "md_set_fixed")
        self.box = None if box is None else np.array(box, dtype=np.float64).reshape(3, 3)
        self.step_count = 0
        self.time_step = None
        self.energy = None
        self.virial = None
        self.pressure = None
        self.pressure_tensor = None
        self.rebuild = True
        self.n_builds = 0

    def forces(self, path=None):
        if self.rebuild or self.buffer <= 0.0:
            self.ctx.build_neighbors_list(self.rcut + self.buffer, fetch=False)
            self.n_builds += 1
            if self.buffer > 0.0:
                self._displacement(reset=True)  # totals count from this build
            self.rebuild = False
        else:
            self.ctx.build_neighbors_list(self.rcut + self.buffer, fetch=False, rebuild=False)
        self.energy, self.virial = self.ctx.compute(path or self.path)
        return self.energy

    def _displacement(self, reset=False):
        flag = C.c_int()
        L.check(L.load().tgx_md_displacement(self.ctx.h, self.buffer, int(reset), C.byref(flag)), "md_displacement")
        return bool(flag.value)

    def kinetic(self):
        ek, T = C.c_double(), C.c_double()
        L.check(L.load().tgx_kinetic(self.ctx.h, C.byref(ek), C.byref(T)), "kinetic")
        return ek.value, T.value

    def pressure_of(self, T: float):
        """(pressure, pressure tensor (3, 3)) in bar from the last forces' virial
        and temperature T (src/turbogap_md.f90:287-295)."""
        p = C.c_double()
        pt = np.empty(9)
        w = np.ascontiguousarray(self.virial, dtype=np.float64).reshape(9)
        L.check(L.load().tgx_md_pressure(self.ctx.h, L.dptr(w), float(T), C.byref(p), L.dptr(pt)), "md_pressure")
        return p.value, pt.reshape(3, 3)

    def step(self, dt: float, T0: float = None, tau: float = 100.0, path=None, P0: float = None,
             tau_p: float = 1000.0, gamma_p: float = 1.0, barostat_sym: str = "iso", scale_box=None,
             variable_dt=None, adaptive_dt=None):
        """One step: forces at x(t), velocity Verlet, E_kin, T and pressure, then
        (scale_box = (gamma (3, 3), n_steps) given) box_scaling at this step's index
        or (P0 given) the Berendsen barostat on box and positions, (T0 given) the
        Berendsen or Bussi thermostat, wrap (src/turbogap_md.f90:280-449).
        variable_dt = (target_pos_step, tau_dt): the step is chosen by
        variable_time_step with dt as its ceiling (src/turbogap_md.f90:178-181)
        and kept in self.time_step for the next one. adaptive_dt = (tmin, tmax,
        xmax, emax, interval): variable_time_step_adaptive every interval steps
        (src/turbogap_md.f90:204-210), likewise kept."""
        lib = L.load()
        self.forces(path)
        cm = self.remove_cm_each_step  # still to do
        if variable_dt is not None:
            if cm:  # the step is chosen from the velocities without the CM drift
                L.check(lib.tgx_remove_cm_vel(self.ctx.h), "remove_cm_vel")
                cm = False
            target, tau_dt = variable_dt
            t = C.c_double(float(dt) if self.step_count == 0 or self.time_step is None else self.time_step)
            L.check(lib.tgx_variable_time_step(self.ctx.h, int(self.step_count == 0), float(target), float(tau_dt),
                                               float(dt), C.byref(t)), "variable_time_step")
            self.time_step = t.value
            dt = self.time_step
        if adaptive_dt is not None:
            tmin, tmax, xmax, emax, interval = adaptive_dt
            if cm:
                L.check(lib.tgx_remove_cm_vel(self.ctx.h), "remove_cm_vel")
                cm = False
            if self.step_count % int(interval) == 0 or self.time_step is None:
                t = C.c_double(float(dt) if self.step_count == 0 or self.time_step is None else self.time_step)
                L.check(lib.tgx_adaptive_time_step(self.ctx.h, int(self.step_count == 0), float(tmin), float(tmax),
                                                   float(xmax), float(emax), float(dt), C.byref(t)),
                        "adaptive_time_step")
                self.time_step = t.value
            dt = self.time_step
        bussi = self.thermostat == "bussi" and T0 is not None
        if self.fused and P0 is None and scale_box is None and not bussi and (T0 is None or T0 > 0.0):
            ekc, Tc, rb = C.c_double(), C.c_double(), C.c_int()
            L.check(lib.tgx_md_advance(self.ctx.h, float(dt), int(self.step_count == 0), int(cm),
                                       float(T0) if T0 is not None else 0.0, float(tau), self.buffer, C.byref(ekc),
                                       C.byref(Tc), C.byref(rb)), "md_advance")
            ek, T = ekc.value, Tc.value
            self.pressure, self.pressure_tensor = self.pressure_of(T)
            if self.buffer > 0.0:
                self.rebuild = bool(rb.value)
            self.step_count += 1
            return self.energy, ek, T
        if cm:
            L.check(lib.tgx_remove_cm_vel(self.ctx.h), "remove_cm_vel")
        L.check(lib.tgx_velocity_verlet(self.ctx.h, float(dt), int(self.step_count == 0)), "velocity_verlet")
        ek, T = self.kinetic()
        self.pressure, self.pressure_tensor = self.pressure_of(T)
        if scale_box is not None:
            g, n_steps = scale_box
            gm = np.ascontiguousarray(np.asarray(g, dtype=np.float64).reshape(9))
            nb = np.empty(9)
            L.check(lib.tgx_box_scaling(self.ctx.h, L.dptr(gm), int(self.step_count), int(n_steps), L.dptr(nb)),
                    "box_scaling")
            self.box = nb.reshape(3, 3)
        elif P0 is not None:
            nb = np.empty(9)
            pt = np.ascontiguousarray(self.pressure_tensor).reshape(9)
            L.check(lib.tgx_berendsen_barostat(self.ctx.h, float(P0), L.dptr(pt), barostat_sym.encode(), float(tau_p),
                                               float(gamma_p), float(dt), L.dptr(nb)), "berendsen_barostat")
            self.box = nb.reshape(3, 3)
        if bussi:
            ndeg = 3 * self.masses.shape[0] - 3
            g = float(self.rng.standard_normal())
            chi2 = float(self.rng.chisquare(ndeg - 1)) if ndeg > 1 else 0.0
            L.check(lib.tgx_bussi_thermostat(self.ctx.h, float(T0), float(tau), float(dt), g, chi2, None),
                    "bussi_thermostat")
        elif T0 is not None:
            L.check(lib.tgx_berendsen(self.ctx.h, float(T0), float(T), float(tau), float(dt)), "berendsen")
        L.check(lib.tgx_wrap_pbc(self.ctx.h), "wrap_pbc")
        if self.buffer > 0.0:
            self.rebuild = self._displacement()
        self.step_count += 1
        return self.energy, ek, T

    def gd_step(self, max_opt_step: float = 0.1, path=None):
        """One gradient-descent relaxation step (optimize = "gd",
        src/turbogap_md.f90:224-228, src/md.f90:540-664): forces and energy at
        the current positions, then the step on the device. Returns (energy,
        gamma, restored); restored means the line search went back to the first
        positions, and the neighbour list is rebuilt before the next forces."""
        lib = L.load()
        self.forces(path)
        g, r = C.c_double(), C.c_int()
        L.check(lib.tgx_gradient_descent(self.ctx.h, float(self.energy), float(max_opt_step),
                                         int(self.step_count == 0), C.byref(g), C.byref(r)), "gradient_descent")
        L.check(lib.tgx_wrap_pbc(self.ctx.h), "wrap_pbc")
        if self.buffer > 0.0:
            self.rebuild = self._displacement() or bool(r.value)
        self.step_count += 1
        return self.energy, g.value, bool(r.value)

    def write_frame(self, fh, species_names, cell, dt: float = 1.0):
        """One extended-XYZ frame of the current state (write_extxyz, src/xyz.f90:169-487):
        positions, velocities (A/fs), the forces and local energies of the last
        force evaluation, lattice, temperature, time step, time and step."""
        from . import xyz as X
        x, v = self.state()
        en, F = self.ctx.results()
        _, T = self.kinetic()
        X.write_extxyz(fh, species_names, x, velocities=v, forces=F, local_energies=en, cell=cell, temperature=T,
                       time_step=dt, time=self.step_count * dt, step=self.step_count)

    def state(self):
        n = self.masses.shape[0]
        x = np.empty((n, 3))
        v = np.empty((n, 3))
        L.check(L.load().tgx_md_get_state(self.ctx.h, L.dptr(x), L.dptr(v)), "md_get_state")
        return x, v

    def remove_cm(self):
        L.check(L.load().tgx_remove_cm_vel(self.ctx.h), "remove_cm_vel")
