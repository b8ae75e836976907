"""Python loader for the FP64 CPU oracle (oracle/tgx_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
check the GPU result, and by bench.py's cpu_baseline leg. The product path
(turbogap_amd) never imports it.

Each function names the reference routine it restates (paths relative to the
TurboGAP tree).
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "_build" / "libtgx_oracle.so"

_lib = None
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)
_PL = C.POINTER(C.c_long)


class OrcHypers(C.Structure):
    _fields_ = [
        ("n_species", C.c_int), ("l_max", C.c_int), ("basis", C.c_int), ("radial_enhancement", C.c_int),
        ("n_quad_soft", C.c_int), ("n_quad_buffer", C.c_int), ("alpha_max", _PI), ("rcut_hard", _PD),
        ("rcut_soft", _PD), ("nf", _PD), ("global_scaling", _PD), ("atom_sigma_r", _PD),
        ("atom_sigma_r_scaling", _PD), ("atom_sigma_t", _PD), ("atom_sigma_t_scaling", _PD),
        ("amplitude_scaling", _PD), ("central_weight", _PD),
    ]


def build():
    subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    lib.orc_build_neighbors_list.restype = C.c_long
    lib.orc_build_neighbors_list.argtypes = [C.c_int, C.c_int, _PD, _PD, _PD, _PD, _PI, C.c_double, _PI, _PI, _PI,
                                             C.c_long, _PI, _PD, _PD, _PD, _PD, _PI]
    lib.orc_get_distance.argtypes = [_PD, _PD, _PD, _PD, _PD, _PD, _PD]
    lib.orc_number_of_unit_cells.argtypes = [_PD, _PD, _PD, C.c_double, _PI]
    lib.orc_build_supercell.argtypes = [C.c_int, _PD, _PI, _PD, _PD, _PD, _PI, _PD, _PI]
    lib.orc_get_soap.restype = C.c_int
    lib.orc_get_soap.argtypes = [C.POINTER(OrcHypers), C.c_int, _PI, _PI, _PD, _PD, _PD, _PD]
    lib.orc_soap_dim.argtypes = [C.POINTER(OrcHypers)]
    lib.orc_radial_table.restype = C.c_int
    lib.orc_radial_table.argtypes = [C.POINTER(OrcHypers), C.c_int, _PD, _PD, _PD, _PD, _PD]
    lib.orc_real_sph_harm.argtypes = [C.c_int, _PD, _PD, _PD]
    lib.orc_get_soap_energy_and_forces.argtypes = [C.c_int, C.c_int, C.c_int, _PD, _PD, _PD, _PD, C.c_double,
                                                   C.c_double, C.c_double, _PI, _PI, _PD, C.c_int, C.c_int, _PD,
                                                   _PD, _PD]
    lib.orc_gap_soap.restype = C.c_long
    lib.orc_gap_soap.argtypes = [C.POINTER(OrcHypers), C.c_int, C.c_int, _PI, _PI, _PD, _PD, _PI, C.c_int, _PD, _PD,
                                 C.c_int, C.c_double, C.c_double, C.c_int, _PD, _PD, _PD, _PD, _PI]
    lib.orc_get_number_of_atom_pairs.restype = C.c_int
    lib.orc_get_number_of_atom_pairs.argtypes = [C.c_int, _PI, C.c_long, _PD, C.c_double, C.c_int, C.c_int,
                                                 C.c_int, C.c_int, C.c_double, C.c_int, _PI, _PI, _PL, _PL]
    lib.orc_soap_batch_memory_model.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _PD, _PD]
    lib.orc_get_soap_compressed.restype = C.c_int
    lib.orc_get_soap_compressed.argtypes = [C.POINTER(OrcHypers), C.POINTER(OrcCompress), C.c_int, _PI, _PI, _PD,
                                            _PD, _PD, _PD]
    lib.orc_gap_soap_compressed.restype = C.c_long
    lib.orc_gap_soap_compressed.argtypes = [C.POINTER(OrcHypers), C.POINTER(OrcCompress), C.c_int, C.c_int, _PI, _PI,
                                            _PD, _PD, _PI, C.c_int, _PD, _PD, C.c_int, C.c_double, C.c_double,
                                            C.c_int, _PD, _PD, _PD, _PD, _PI]
    lib.orc_num_threads.restype = C.c_int
    lib.orc_set_num_threads.argtypes = [C.c_int]
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_PD) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_PI) if a is not None else None


class OrcCompress(C.Structure):
    _fields_ = [("dim", C.c_int), ("nnz", C.c_int), ("P_i", _PI), ("P_j", _PI), ("P_el", _PD)]


def compress_struct(h):
    """OrcCompress of a compressed block (arrays kept alive), or (None, None)."""
    if not getattr(h, "compressed", False):
        return None, None
    keep = (np.ascontiguousarray(h.compress_P_i, dtype=np.int32), np.ascontiguousarray(h.compress_P_j, dtype=np.int32),
            np.ascontiguousarray(h.compress_P_el, dtype=np.float64))
    c = OrcCompress(int(h.compress_dim), keep[0].shape[0], _i(keep[0]), _i(keep[1]), _d(keep[2]))
    return c, keep


def hypers_struct(h):
    """OrcHypers for a turbogap_amd.hypers.SoapTurbo-like object (arrays kept alive)."""
    keep = {k: np.ascontiguousarray(getattr(h, k), dtype=np.float64) for k in (
        "rcut_hard", "rcut_soft", "nf", "global_scaling", "atom_sigma_r", "atom_sigma_r_scaling", "atom_sigma_t",
        "atom_sigma_t_scaling", "amplitude_scaling", "central_weight")}
    keep["alpha_max"] = np.ascontiguousarray(h.alpha_max, dtype=np.int32)
    s = OrcHypers()
    s.n_species = len(h.alpha_max)
    s.l_max = h.l_max
    s.basis = {"poly3": 0, "poly3gauss": 1}[h.basis]
    s.radial_enhancement = h.radial_enhancement
    s.n_quad_soft = h.n_quad_soft
    s.n_quad_buffer = h.n_quad_buffer
    s.alpha_max = _i(keep["alpha_max"])
    for k, v in keep.items():
        if k != "alpha_max":
            setattr(s, k, _d(v))
    return s, keep


def get_distance(posi, posj, box):
    """src/neighbors.f90:47-163 (PBC on)."""
    lib = load()
    b = np.ascontiguousarray(box, dtype=np.float64)
    pi = np.ascontiguousarray(posi, dtype=np.float64)
    pj = np.ascontiguousarray(posj, dtype=np.float64)
    dist = np.empty(3)
    d = np.empty(1)
    lib.orc_get_distance(_d(pi), _d(pj), _d(b[0]), _d(b[1]), _d(b[2]), _d(dist), _d(d))
    return dist, d[0]


def number_of_unit_cells(box, rcut):
    """src/neighbors.f90:175-236."""
    b = np.ascontiguousarray(box, dtype=np.float64)
    ind = np.empty(3, np.int32)
    load().orc_number_of_unit_cells(_d(b[0]), _d(b[1]), _d(b[2]), float(rcut), _i(ind))
    return ind


def build_supercell(positions, species, box, indices):
    """src/read_files.f90:320-367. Returns positions_sc, species_sc, box_sc."""
    n = positions.shape[0]
    ind = np.ascontiguousarray(indices, dtype=np.int32)
    m = int(np.prod(ind))
    b = np.array(box, dtype=np.float64, copy=True)
    psc = np.empty((n * m, 3))
    ssc = np.empty(n * m, np.int32)
    load().orc_build_supercell(n, _d(np.ascontiguousarray(positions, dtype=np.float64)),
                               _i(np.ascontiguousarray(species, dtype=np.int32)), _d(b[0]), _d(b[1]), _d(b[2]),
                               _i(ind), _d(psc), _i(ssc))
    return psc, ssc, b


def build_neighbors_list(positions_sc, species_sc, box, rcut_max, n_sites=None, indices=(1, 1, 1), do_list=None):
    """src/neighbors.f90:243-605. positions_sc (Nsc, 3); returns the reference's arrays."""
    lib = load()
    pos = np.ascontiguousarray(positions_sc, dtype=np.float64)
    nsc = pos.shape[0]
    n = nsc if n_sites is None else int(n_sites)
    sp = np.ascontiguousarray(species_sc, dtype=np.int32)
    b = np.array(box, dtype=np.float64, copy=True)
    ind = np.ascontiguousarray(indices, dtype=np.int32)
    dl = None if do_list is None else np.ascontiguousarray(do_list, dtype=np.int32)
    nn = np.zeros(n, np.int32)
    args = lambda cap, arrs: lib.orc_build_neighbors_list(
        n, nsc, _d(pos), _d(b[0]), _d(b[1]), _d(b[2]), _i(sp), float(rcut_max), _i(dl), _i(ind), _i(nn), cap,
        *[(_d(a) if a.dtype == np.float64 else _i(a)) if a is not None else None for a in arrs])
    P = args(0, [None] * 6)
    out = {
        "neighbors_list": np.empty(P, np.int32), "rjs": np.empty(P), "thetas": np.empty(P), "phis": np.empty(P),
        "xyz": np.empty((P, 3)), "neighbor_species": np.empty(P, np.int32)}
    b = np.array(box, dtype=np.float64, copy=True)
    args(P, [out["neighbors_list"], out["rjs"], out["thetas"], out["phis"], out["xyz"], out["neighbor_species"]])
    out["n_neigh"] = nn
    out["n_atom_pairs"] = P
    return out


def soap_dim(h):
    s, keep = hypers_struct(h)
    return load().orc_soap_dim(C.byref(s))


def radial_table(h, species_index):
    """Quadrature nodes, weights, basis values psi (nb, nq), B (nb, nq), central coefficients."""
    s, keep = hypers_struct(h)
    nb = h.alpha_max[species_index]
    xq = np.empty(64)
    wq = np.empty(64)
    psi = np.empty(nb * 64)
    B = np.empty(nb * 64)
    cen = np.empty(nb)
    nq = load().orc_radial_table(C.byref(s), int(species_index), _d(xq), _d(wq), _d(psi), _d(B), _d(cen))
    return xq[:nq], wq[:nq], psi[:nb * nq].reshape(nb, nq), B[:nb * nq].reshape(nb, nq), cen


def real_sph_harm(lmax, u):
    """Real spherical harmonics (l*l + l + m order) and the gradient of r^l Y_lm at unit u."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    nlm = (lmax + 1) ** 2
    Y = np.empty(nlm)
    dY = np.empty((nlm, 3))
    load().orc_real_sph_harm(int(lmax), _d(u), _d(Y), _d(dY))
    return Y, dY


def get_soap(h, n_neigh, pair_species, rjs, xyz, derivatives=True):
    """soap_turbo get_soap restated (conventions in tgx_oracle.cpp). Returns
    soap (n, D) and soap_cart_der (P, D, 3)."""
    s, keep = hypers_struct(h)
    comp, ckeep = compress_struct(h)
    lib = load()
    D = comp.dim if comp is not None else lib.orc_soap_dim(C.byref(s))
    nn = np.ascontiguousarray(n_neigh, dtype=np.int32)
    ps = np.ascontiguousarray(pair_species, dtype=np.int32)
    r = np.ascontiguousarray(rjs, dtype=np.float64)
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    n = nn.shape[0]
    P = int(nn.sum())
    soap = np.empty((n, D))
    der = np.empty((P, D, 3)) if derivatives else None
    if comp is None:
        lib.orc_get_soap(C.byref(s), n, _i(nn), _i(ps), _d(r), _d(x), _d(soap), _d(der))
    elif lib.orc_get_soap_compressed(C.byref(s), C.byref(comp), n, _i(nn), _i(ps), _d(r), _d(x), _d(soap),
                                     _d(der)) < 0:
        raise ValueError("compression entries outside the descriptor")
    return soap, der


def get_soap_energy_and_forces(soap, soap_der, Qs, alphas, delta, zeta, e0, n_neigh, neighbors_list, xyz, n_sites0,
                               do_forces=True):
    """src/gap.f90:66-259."""
    soap = np.ascontiguousarray(soap, dtype=np.float64)
    n, D = soap.shape
    Qs = np.ascontiguousarray(Qs, dtype=np.float64)
    al = np.ascontiguousarray(alphas, dtype=np.float64)
    en = np.empty(n)
    f = np.zeros((n_sites0, 3))
    v = np.zeros(9)
    sd = np.ascontiguousarray(soap_der, dtype=np.float64) if do_forces else None
    load().orc_get_soap_energy_and_forces(
        n, D, Qs.shape[0], _d(soap), _d(sd), _d(Qs), _d(al), float(delta), float(zeta), float(e0),
        _i(np.ascontiguousarray(n_neigh, dtype=np.int32)), _i(np.ascontiguousarray(neighbors_list, dtype=np.int32)),
        _d(np.ascontiguousarray(xyz, dtype=np.float64)), int(n_sites0), int(do_forces), _d(en), _d(f), _d(v))
    return en, f, v.reshape(3, 3).T.copy()


def species_index(codes, species_types):
    idx = {int(c): i for i, c in enumerate(species_types)}
    return np.array([idx.get(int(c), -1) for c in codes], np.int32)


def gap_soap(h, nl, species_sc, Qs, alphas, n_total_sites, do_forces=True):
    """src/gap_interface.f90:180-431 for one block (marshal, get_soap,
    prediction, fold). energies0/forces0 are zero where the block does not touch."""
    s, keep = hypers_struct(h)
    comp, ckeep = compress_struct(h)
    lib = load()
    n0 = nl["n_neigh"].shape[0]
    spi = species_index(species_sc, h.species_types)
    Qs = np.ascontiguousarray(Qs, dtype=np.float64)
    al = np.ascontiguousarray(alphas, dtype=np.float64)
    e0 = np.zeros(n_total_sites)
    f0 = np.zeros((n_total_sites, 3))
    v = np.zeros(9)
    D = comp.dim if comp is not None else lib.orc_soap_dim(C.byref(s))
    soap = np.empty((n0, D))
    nsb = C.c_int()
    args = (int(n_total_sites), n0, _i(nl["n_neigh"]), _i(nl["neighbors_list"]), _d(nl["rjs"]),
            _d(np.ascontiguousarray(nl["xyz"])), _i(spi), int(h.central_species), _d(Qs), _d(al), Qs.shape[0],
            float(h.delta), float(h.zeta), int(do_forces), _d(e0), _d(f0), _d(v), _d(soap), C.byref(nsb))
    if comp is None:
        P = lib.orc_gap_soap(C.byref(s), *args)
    else:
        P = lib.orc_gap_soap_compressed(C.byref(s), C.byref(comp), *args)
        if P < 0:
            raise ValueError("compression entries outside the descriptor")
    return e0, f0, v.reshape(3, 3).T.copy(), soap[:nsb.value].copy(), int(P)


def compute(system, models, nl=None):
    """All blocks on one structure (src/turbogap.f90:1271-1380,1941-1954).
    models: list of (hypers, Qs, alphas). Returns E, energies, forces, virial, nl."""
    rc = max(h.rcut_max for h, _, _ in models)
    if nl is None:
        nl = build_neighbors_list(system.positions, system.species, system.box, rc)
    n = system.n
    en = np.zeros(n)
    f = np.zeros((n, 3))
    w = np.zeros((3, 3))
    for h, Qs, al in models:
        e0, f0, v, _, _ = gap_soap(h, nl, system.species, Qs, al, n)
        en += e0
        f += f0
        w += v
    return en.sum(), en, f, w, nl


def block_descriptors(h, system):
    """Descriptors of one block's centres (for building sparse sets)."""
    rc = h.rcut_max
    nl = build_neighbors_list(system.positions, system.species, system.box, rc)
    e0, f0, v, soap, P = gap_soap(h, nl, system.species, np.zeros((1, h.dim)), np.zeros(1), system.n,
                                  do_forces=False)
    return soap


def set_threads(n):
    load().orc_set_num_threads(int(n))


def num_threads():
    return load().orc_num_threads()
