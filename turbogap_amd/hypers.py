"""Hypers of one soap_turbo block: the host mirror of ``type soap_turbo``
(src/types.f90:75-134) and of its ``.gap`` keywords
(src/read_files.f90:4641-5236).

Semantics kept from the reference:
  * ``rcut_soft = rcut - buffer`` (src/read_files.f90:5101-5109); a zero
    buffer means ``rcut_soft = rcut``.
  * ``n_max`` per species is ``alpha_max``; the block's radial size is their
    sum (src/read_files.f90:4769-4780) and the descriptor size is
    ``N (N+1)/2 (l_max+1)`` (src/read_files.f90:5233-5235).
  * ``basis`` is "poly3" or "poly3gauss"; ``scaling_mode`` is "polynomial".
  * Defaults: nf = 4, amplitude_scaling = 1, central_weight = 1,
    global_scaling = 1, scalings 0 (src/read_files.f90:4650-4668).
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

from . import lib as L

BASES = {"poly3": 0, "poly3gauss": 1}


@dataclasses.dataclass
class SoapTurbo:
    species_types: List[int]
    alpha_max: List[int]
    l_max: int
    rcut_hard: List[float]
    rcut_soft: List[float]
    atom_sigma_r: List[float]
    atom_sigma_t: List[float]
    atom_sigma_r_scaling: Optional[List[float]] = None
    atom_sigma_t_scaling: Optional[List[float]] = None
    amplitude_scaling: Optional[List[float]] = None
    central_weight: Optional[List[float]] = None
    global_scaling: Optional[List[float]] = None
    nf: Optional[List[float]] = None
    basis: str = "poly3gauss"
    radial_enhancement: int = 0
    scaling_mode: str = "polynomial"
    central_species: int = 0
    zeta: float = 2.0
    delta: float = 1.0
    n_quad_soft: int = 32
    n_quad_buffer: int = 16

    def __post_init__(self):
        n = len(self.species_types)
        for name, default in (
            ("atom_sigma_r_scaling", 0.0),
            ("atom_sigma_t_scaling", 0.0),
            ("amplitude_scaling", 1.0),
            ("central_weight", 1.0),
            ("global_scaling", 1.0),
            ("nf", 4.0),
        ):
            if getattr(self, name) is None:
                setattr(self, name, [default] * n)
        for name in ("alpha_max", "rcut_hard", "rcut_soft", "atom_sigma_r", "atom_sigma_t", "atom_sigma_r_scaling",
                     "atom_sigma_t_scaling", "amplitude_scaling", "central_weight", "global_scaling", "nf"):
            if len(getattr(self, name)) != n:
                raise ValueError(f"{name} must have one entry per species ({n})")
        if self.basis not in BASES:
            raise ValueError(f"basis must be one of {sorted(BASES)}, got {self.basis!r}")
        if self.scaling_mode != "polynomial":
            raise ValueError("scaling_mode must be 'polynomial'")

    @property
    def n_species(self) -> int:
        return len(self.species_types)

    @property
    def n_max(self) -> int:
        return int(sum(self.alpha_max))

    @property
    def dim(self) -> int:
        N = self.n_max
        return N * (N + 1) // 2 * (self.l_max + 1)

    @property
    def rcut_max(self) -> float:
        return float(max(self.rcut_hard))

    @classmethod
    def from_buffer(cls, rcut: List[float], buffer: List[float], **kw) -> "SoapTurbo":
        """Build from the .gap keywords rcut and buffer (rcut_soft = rcut - buffer)."""
        soft = [r if b == 0.0 else r - b for r, b in zip(rcut, buffer)]
        return cls(rcut_hard=list(rcut), rcut_soft=soft, **kw)

    def c_struct(self):
        """tgx_soap_hypers plus the arrays it points into (keep both alive)."""
        keep = {
            "alpha_max": np.ascontiguousarray(self.alpha_max, dtype=np.int32),
            "rcut_hard": np.ascontiguousarray(self.rcut_hard, dtype=np.float64),
            "rcut_soft": np.ascontiguousarray(self.rcut_soft, dtype=np.float64),
            "nf": np.ascontiguousarray(self.nf, dtype=np.float64),
            "global_scaling": np.ascontiguousarray(self.global_scaling, dtype=np.float64),
            "atom_sigma_r": np.ascontiguousarray(self.atom_sigma_r, dtype=np.float64),
            "atom_sigma_r_scaling": np.ascontiguousarray(self.atom_sigma_r_scaling, dtype=np.float64),
            "atom_sigma_t": np.ascontiguousarray(self.atom_sigma_t, dtype=np.float64),
            "atom_sigma_t_scaling": np.ascontiguousarray(self.atom_sigma_t_scaling, dtype=np.float64),
            "amplitude_scaling": np.ascontiguousarray(self.amplitude_scaling, dtype=np.float64),
            "central_weight": np.ascontiguousarray(self.central_weight, dtype=np.float64),
        }
        s = L.SoapHypersC()
        s.n_species = self.n_species
        s.l_max = self.l_max
        s.basis = BASES[self.basis]
        s.radial_enhancement = self.radial_enhancement
        s.n_quad_soft = self.n_quad_soft
        s.n_quad_buffer = self.n_quad_buffer
        s.alpha_max = L.iptr(keep["alpha_max"])
        for k in keep:
            if k != "alpha_max":
                setattr(s, k, L.dptr(keep[k]))
        return s, keep
