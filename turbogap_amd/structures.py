"""Seeded synthetic structures and models for the benchmark and test workloads.

Every random draw comes from libtgx's host generators (xoshiro256** seeded by
splitmix64), so the same seeds give the same structures on any machine; no GPU
is needed to generate them.

  * ``rsa``: random sequential addition in a periodic orthorhombic box with a
    minimum pair distance.
  * ``carbon216_system``: 216 C atoms, L = 12.91 Å, d_min = 1.2 Å.
  * ``binary_system``: 65,536 atoms, L = 86.8 Å, species A/B 50:50 by a seeded
    permutation; ``subset_system`` is the 4,096-atom box at the same density.
  * ``soap_carbon`` / ``soap_binary``: poly3gauss, n_max 8 per species,
    l_max 8, rcut 3.7 Å, buffer 0.5 Å (rcut_soft 3.2 Å), sigma_r = sigma_t =
    0.5 Å, central_weight 1, zeta 4, delta 1 eV.
  * ``sparse_set``: sparse descriptors are the descriptors of the first
    n_sparse centres of seeded second structures; alphas ~ N(0, (1e-3 eV)^2).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Callable, List

import numpy as np

from . import lib as L
from .hypers import SoapTurbo

CARBON = 6
SPECIES_A = 1
SPECIES_B = 2
AMU = 103.6426965268  # amu -> eV fs^2 / Å^2 (src/read_files.f90:241)


@dataclass
class System:
    positions: np.ndarray  # (N, 3)
    species: np.ndarray    # (N,) int32 codes
    box: np.ndarray        # (3, 3) rows a, b, c

    @property
    def n(self) -> int:
        return self.positions.shape[0]


def rsa(n: int, L_box, dmin: float, seed: int) -> np.ndarray:
    Lb = np.ascontiguousarray(np.broadcast_to(np.asarray(L_box, float), (3,)), dtype=np.float64)
    pos = np.empty((n, 3))
    L.check(L.load().tgx_generate_rsa(int(n), L.dptr(Lb), float(dmin), C.c_uint64(seed), L.dptr(pos)),
            "generate_rsa")
    return pos


def permutation(n: int, seed: int) -> np.ndarray:
    p = np.empty(n, np.int32)
    L.check(L.load().tgx_generate_permutation(int(n), C.c_uint64(seed), L.iptr(p)), "generate_permutation")
    return p


def normal(n: int, std: float, seed: int) -> np.ndarray:
    out = np.empty(n)
    L.check(L.load().tgx_generate_normal(int(n), float(std), C.c_uint64(seed), L.dptr(out)), "generate_normal")
    return out


def cubic_box(Lc: float) -> np.ndarray:
    return np.diag([Lc, Lc, Lc]).astype(np.float64)


def carbon216_system(seed: int = 1001) -> System:
    Lc = 12.91
    return System(rsa(216, Lc, 1.2, seed), np.full(216, CARBON, np.int32), cubic_box(Lc))


def two_species_system(n: int, Lc: float, seed: int) -> System:
    pos = rsa(n, Lc, 1.2, seed)
    perm = permutation(n, seed + 7919)
    sp = np.full(n, SPECIES_A, np.int32)
    sp[perm[n // 2:]] = SPECIES_B
    return System(pos, sp, cubic_box(Lc))


def binary_system(n: int = 65536, seed: int = 2001) -> System:
    # L = 86.8 Å at 65,536 atoms; other sizes keep the density
    Lc = 86.8 * (n / 65536.0) ** (1.0 / 3.0)
    return two_species_system(n, Lc, seed)


def subset_system(seed: int = 2002) -> System:
    """The 4,096-atom box at the density of the 65,536-atom box (L = 86.8 (1/16)^(1/3) Å)."""
    return binary_system(4096, seed)


def soap_hypers(species: List[int], central_species: int = 0, n_max: int = 8, l_max: int = 8) -> SoapTurbo:
    ns = len(species)
    return SoapTurbo.from_buffer(
        rcut=[3.7] * ns, buffer=[0.5] * ns, species_types=list(species), alpha_max=[n_max] * ns, l_max=l_max,
        atom_sigma_r=[0.5] * ns, atom_sigma_t=[0.5] * ns, central_weight=[1.0] * ns, basis="poly3gauss",
        central_species=central_species, zeta=4.0, delta=1.0)


def soap_carbon() -> SoapTurbo:
    return soap_hypers([CARBON], central_species=0)


def soap_binary() -> List[SoapTurbo]:
    return [soap_hypers([SPECIES_A, SPECIES_B], central_species=c) for c in (1, 2)]


DescriptorFn = Callable[[SoapTurbo, System], np.ndarray]


def sparse_set(hypers: SoapTurbo, n_sparse: int, describe: DescriptorFn, make_structure: Callable[[int], System],
               seed: int, alpha_std: float = 1.0e-3):
    """Qs (n_sparse, D) from the centres of seeded structures make_structure(seed + k),
    k = 0, 1, ..., taken in order; alphas ~ N(0, alpha_std^2) with seed + 104729."""
    rows = []
    have = 0
    k = 0
    while have < n_sparse:
        q = describe(hypers, make_structure(seed + k))
        rows.append(q)
        have += q.shape[0]
        k += 1
        if k > 1000:
            raise RuntimeError("sparse_set: structures yield no centres")
    Qs = np.concatenate(rows, axis=0)[:n_sparse].copy()
    alphas = normal(n_sparse, alpha_std, seed + 104729)
    return Qs, alphas


def masses_for(species: np.ndarray, amu_by_code=None) -> np.ndarray:
    """Masses in eV fs^2/Å^2; carbon (12.011 amu) for every code by default."""
    amu_by_code = amu_by_code or {}
    return np.array([amu_by_code.get(int(s), 12.011) for s in species]) * AMU


def gpu_describe(ctx_factory=None):
    """A DescriptorFn that computes a block's centre descriptors on the GPU."""
    from .context import Context

    def describe(hypers: SoapTurbo, system: System) -> np.ndarray:
        with (ctx_factory() if ctx_factory else Context(0)) as ctx:
            ctx.set_system(system.positions, system.species, system.box)
            ctx.build_neighbors_list(hypers.rcut_max, fetch=False)
            mid = ctx.soap_model(hypers, np.zeros((1, hypers.dim)), np.zeros(1))
            ctx.soap_backend_begin(mid)
            q, _ = ctx.get_soap(derivatives=False)
            ctx.soap_backend_end()
        return q

    return describe


def binary_models(describe: DescriptorFn, n_sparse: int = 4000, seed: int = 9001):
    """Both central-species blocks of the two-species box with their sparse sets, taken from
    seeded 16,384-atom structures at the same density."""
    models = []
    for k, h in enumerate(soap_binary()):
        Qs, al = sparse_set(h, n_sparse, describe, lambda s: binary_system(16384, seed=s), seed=seed + 100 * k)
        models.append((h, Qs, al))
    return models
