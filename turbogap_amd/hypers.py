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
    # compress_soap: the compressed size and the matrix entries (1-based), as
    # src/read_files.f90:5162-5211 stores them; None for an uncompressed block
    compress_dim: Optional[int] = None
    compress_P_i: Optional[List[int]] = None
    compress_P_j: Optional[List[int]] = None
    compress_P_el: Optional[List[float]] = None

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
    def full_dim(self) -> int:
        """Size of the uncompressed power spectrum."""
        N = self.n_max
        return N * (N + 1) // 2 * (self.l_max + 1)

    @property
    def compressed(self) -> bool:
        return self.compress_P_i is not None

    @property
    def dim(self) -> int:
        """Descriptor size the sparse set has: the compressed size when compressed."""
        return int(self.compress_dim) if self.compressed else self.full_dim

    def with_compression(self, dim: int, P_i, P_j, P_el) -> "SoapTurbo":
        """A copy carrying the compression matrix (entries 1-based)."""
        P_i, P_j, P_el = (np.asarray(P_i, dtype=np.int64).reshape(-1), np.asarray(P_j, dtype=np.int64).reshape(-1),
                          np.asarray(P_el, dtype=np.float64).reshape(-1))
        if not (P_i.shape == P_j.shape == P_el.shape) or P_i.size == 0:
            raise ValueError("compression needs matching, non-empty P_i, P_j, P_el")
        if P_i.min() < 1 or P_i.max() > dim or P_j.min() < 1 or P_j.max() > self.full_dim:
            raise ValueError(f"compression entries outside 1..{dim} x 1..{self.full_dim}")
        return dataclasses.replace(self, compress_dim=int(dim), compress_P_i=P_i.tolist(), compress_P_j=P_j.tolist(),
                                   compress_P_el=P_el.tolist())

    def compression_arrays(self):
        """(dim, P_i int32, P_j int32, P_el float64) of a compressed block."""
        return (int(self.compress_dim), np.ascontiguousarray(self.compress_P_i, dtype=np.int32),
                np.ascontiguousarray(self.compress_P_j, dtype=np.int32),
                np.ascontiguousarray(self.compress_P_el, dtype=np.float64))

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


def compress_trivial(alpha_max: List[int], l_max: int):
    """compress_mode "trivial": keep the components p(n, n', l) whose n or n' is
    the first radial function of a species block, in descriptor order, as the
    reference's tools/compress_indices/compress_indices.py:39-57 lists them.
    Returns (dim, P_i, P_j, P_el) with 1-based entries and unit weights."""
    N = int(sum(alpha_max))
    pivots = set(int(x) for x in np.cumsum([1] + list(alpha_max[:-1])))
    keep = []
    k = 1
    for n in range(1, N + 1):
        for m in range(n, N + 1):
            for _ in range(l_max + 1):
                if n in pivots or m in pivots:
                    keep.append(k)
                k += 1
    dim = len(keep)
    return dim, list(range(1, dim + 1)), keep, [1.0] * dim


def read_compress_file(path: str, n_species: int):
    """A file_compress_soap file (src/read_files.f90:5166-5211): a header of
    n_species alpha_max values, l_max and the compressed size, then either a
    P_transformation ... end_transformation block of "i j el" lines or (the
    older form) one kept component index per line. Returns (dim, P_i, P_j, P_el,
    alpha_max, l_max) with 1-based entries."""
    with open(path) as f:
        lines = f.read().splitlines()
    toks = []
    k = 0
    while len(toks) < n_species + 2:
        if k >= len(lines):
            raise ValueError(f"{path}: header needs {n_species} alpha_max values, l_max and the compressed size")
        toks += lines[k].replace(",", " ").split()
        k += 1
    head = [int(t) for t in toks[:n_species + 2]]
    alpha_max, l_max, dim = head[:n_species], head[n_species], head[n_species + 1]
    body = lines[k:]
    if body and body[0].strip() == "P_transformation":
        P_i, P_j, P_el = [], [], []
        for ln in body[1:]:
            if ln.strip() == "end_transformation":
                break
            t = ln.replace(",", " ").split()
            P_i.append(int(t[0]))
            P_j.append(int(t[1]))
            P_el.append(float(t[2].replace("d", "e").replace("D", "e")))
        else:
            raise ValueError(f"{path}: P_transformation without end_transformation")
    else:
        if len(body) < dim:
            raise ValueError(f"{path}: {dim} component indices expected, {len(body)} found")
        P_j = [int(ln.split()[0]) for ln in body[:dim]]
        P_i = list(range(1, dim + 1))
        P_el = [1.0] * dim
    return dim, P_i, P_j, P_el, alpha_max, l_max
