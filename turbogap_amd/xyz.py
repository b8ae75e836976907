"""Extended-XYZ trajectory frames in TurboGAP's layout (host side).

write_extxyz writes one frame as src/xyz.f90:169-487 (write_extxyz) does:
  * line 1: the atom count as I8;
  * line 2: `` Properties=...`` for the per-atom columns present, then, when
    given, Lattice (F16.10 each, the first left-adjusted in its field), temperature and
    pressure (F16.8), time_step (F16.4), time (F16.6), energy (F16.6 of the
    sum of the local energies) followed by the energy string, virial and stress
    = -virial/volume (F16.8 each, the first left-adjusted in its field), volume (F16.6) and
    step (or i_config for a negative step);
  * one line per atom: species (A8), pos / velocities / forces (3 x F16.8),
    local_energy (F16.8), masses (F16.8, amu), local properties (F16.8 each),
    fix_atoms (3 x L1), each field after one blank.
Values that do not fit a field print as asterisks, as Fortran does.

energy_string builds the per-term string of get_xyz_energy_string
(src/xyz.f90:68-166), whose first term is followed by two blanks.
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

import numpy as np

AMU = 103.6426965268  # src/xyz.f90:463


def _f(x: float, w: int, d: int) -> str:
    s = f"{float(x):{w}.{d}f}"
    return s if len(s) <= w else "*" * w


def _adjl(s: str) -> str:
    """Fortran adjustl then trim."""
    return s.strip()


def energy_string(terms: Sequence[Tuple[str, float]]) -> str:
    """get_xyz_energy_string: terms are (label, summed energy), e.g.
    [("energy_soap", -650.30414244), ("energy_2b", 96.66), ...] in the
    reference's order."""
    parts = [f"{k}={_adjl(_f(v, 16, 8))}" for k, v in terms]
    if not parts:
        return ""
    return parts[0] + ("  " + " ".join(parts[1:]) if len(parts) > 1 else " ")


def _box_triplets(a, b, c, w, d) -> str:
    vals = [a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]]
    s = [_f(v, w, d) for v in vals]
    return _adjl(s[0]).ljust(w) + "".join(s[1:])  # adjustl keeps the field width


def write_extxyz(fh, species: Sequence[str], positions, *, velocities=None, forces=None, local_energies=None,
                 masses=None, local_properties: Optional[Dict[str, np.ndarray]] = None, fix_atoms=None,
                 cell=None, temperature=None, pressure=None, time_step=None, time=None, energy=None,
                 energies: Optional[str] = None, virial=None, stress: bool = False, volume: bool = False,
                 step: Optional[int] = None) -> None:
    """Write one frame to the open text file fh. positions etc. are (n, 3);
    cell is the rows a, b, c; masses in eV fs^2/A^2 (written in amu);
    local_properties maps a label to (n,) values; fix_atoms is (n, 3) bool.
    energy defaults to the sum of local_energies; energies is the energy
    string written after it (energy_string)."""
    pos = np.asarray(positions, dtype=np.float64)
    n = pos.shape[0]
    cols = [("species", "S", 1), ("pos", "R", 3)]
    if velocities is not None:
        cols.append(("velocities", "R", 3))
    if forces is not None:
        cols.append(("forces", "R", 3))
    if local_energies is not None:
        cols.append(("local_energy", "R", 1))
    if masses is not None:
        cols.append(("masses", "R", 1))
    for lab in (local_properties or {}):
        cols.append((lab, "R", 1))
    if fix_atoms is not None:
        cols.append(("fix_atoms", "S", 3))
    head = [" Properties=" + ":".join(f"{k}:{t}:{m}" for k, t, m in cols)]
    if cell is not None:
        a, b, c = (np.asarray(v, dtype=np.float64) for v in cell)
        head.append(' Lattice="' + _box_triplets(a, b, c, 16, 10) + '"')
    if temperature is not None:
        head.append(" temperature=" + _adjl(_f(temperature, 16, 8)))
    if pressure is not None:
        head.append(" pressure=" + _adjl(_f(pressure, 16, 8)))
    if time_step is not None:
        head.append(" time_step=" + _adjl(_f(time_step, 16, 4)))
    if time is not None:
        head.append(" time=" + _adjl(_f(time, 16, 6)))
    if energy is not None or local_energies is not None:
        e = float(np.sum(local_energies)) if energy is None else float(energy)
        head.append(" energy=" + _adjl(_f(e, 16, 6)))
        head.append(" " + (energies or "").strip())
    if virial is not None:
        W = np.asarray(virial, dtype=np.float64)
        head.append(' virial="' + _box_triplets(W[0], W[1], W[2], 16, 8) + '"')
        if stress:
            if cell is None:
                raise ValueError("stress needs the cell")
            vol = float(np.dot(np.cross(a, b), c))
            S = -W / vol
            head.append(' stress="' + _box_triplets(S[0], S[1], S[2], 16, 8) + '"')
    if volume:
        if cell is None:
            raise ValueError("volume needs the cell")
        head.append(" volume=" + _adjl(_f(float(np.dot(np.cross(a, b), c)), 16, 6)))
    if step is not None:
        head.append((" step=" if step >= 0 else " i_config=") + str(abs(int(step))))
    fh.write(f"{n:8d}\n")
    fh.write("".join(head) + "\n")
    vel = None if velocities is None else np.asarray(velocities, dtype=np.float64)
    frc = None if forces is None else np.asarray(forces, dtype=np.float64)
    le = None if local_energies is None else np.asarray(local_energies, dtype=np.float64)
    ms = None if masses is None else np.asarray(masses, dtype=np.float64)
    lp = [np.asarray(v, dtype=np.float64) for v in (local_properties or {}).values()]
    fx = None if fix_atoms is None else np.asarray(fix_atoms, dtype=bool)
    for i in range(n):
        row = [" " + f"{species[i]:<8.8s}"]
        row += [" " + _f(v, 16, 8) for v in pos[i]]
        if vel is not None:
            row += [" " + _f(v, 16, 8) for v in vel[i]]
        if frc is not None:
            row += [" " + _f(v, 16, 8) for v in frc[i]]
        if le is not None:
            row.append(" " + _f(le[i], 16, 8))
        if ms is not None:
            row.append(" " + _f(ms[i] / AMU, 16, 8))
        for v in lp:
            row.append(" " + _f(v[i], 16, 8))
        if fx is not None:
            row += [" " + ("T" if t else "F") for t in fx[i]]
        fh.write("".join(row) + "\n")
