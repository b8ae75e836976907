"""Host-side mirror of TurboGAP's SOAP -> GAP call surface over libtgx.

``Context`` owns one device context (``gpu_context_init``,
src/gpu_context.f90:41-58). Its methods keep the reference's names, argument
meaning and array conventions:

  build_neighbors_list       src/neighbors.f90:243-605
  number_of_unit_cells       src/neighbors.f90:175-236 (module function, host only)
  build_supercell            src/read_files.f90:320-367 (module function, host only)
  soap_model / backend_begin src/gap.f90:49-64 (one resident sparse set per block)
  get_gap_soap               src/gap_interface.f90:41-480
  get_soap                   soap_turbo get_soap (call shape src/gap_interface.f90:320-325)
  get_soap_energy_and_forces src/gap.f90:66-259
  compute                    src/turbogap.f90:1253-1380,1941-1954 (all blocks, totals)

Arrays follow the Fortran layouts: positions(3, N) is a C array of shape
(N, 3); soap(D, n) is (n, D); soap_cart_der(3, D, P) is (P, D, 3); neighbour
indices are 1-based. Errors raise ``TgxError`` (the reference stops instead).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import lib as L
from .hypers import SoapTurbo


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def number_of_unit_cells(box, rcut: float) -> np.ndarray:
    """number_of_unit_cells_for_given_cutoff (src/neighbors.f90:175-236): the
    repetitions along a, b, c (rows of box) that put the supercell's opposite
    faces at least 2 rcut apart. Host only."""
    b = _f64(box, (3, 3))
    ind = np.empty(3, np.int32)
    L.check(L.load().tgx_number_of_unit_cells(L.dptr(b), float(rcut), L.iptr(ind)), "number_of_unit_cells")
    return ind


def build_supercell(positions, species, box, indices):
    """The supercell of read_xyz (src/read_files.f90:320-367): positions (N, 3)
    and species repeated indices times (copies in a-slowest order, the cell
    first); returns (positions_sc, species_sc, box_sc). Host only."""
    pos = _f64(positions)
    n = pos.shape[0]
    ind = _i32(indices).reshape(3)
    m = int(np.prod(ind))
    psc = np.empty((n * m, 3))
    ssc = np.empty(n * m, np.int32)
    bsc = np.empty((3, 3))
    L.check(L.load().tgx_build_supercell(n, L.dptr(pos), L.iptr(_i32(species)), L.dptr(_f64(box, (3, 3))),
                                         L.iptr(ind), L.dptr(psc), L.iptr(ssc), L.dptr(bsc)), "build_supercell")
    return psc, ssc, bsc


class Context:
    def __init__(self, device: int = 0):
        self._lib = L.load()
        h = C.c_void_p()
        L.check(self._lib.tgx_init(device, C.byref(h)), "gpu_context_init")
        self.h = h
        self._children = []  # objects (networks) that hold device state of this context
        self.models: Dict[int, SoapTurbo] = {}
        self.n_sites = 0
        self.n_sites_sc = 0

    # -- lifetime --------------------------------------------------------------
    def close(self):
        for c in self._children:
            c.close()
        self._children = []
        if self.h is not None and self.h.value:
            self._lib.tgx_finalize(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        L.check(self._lib.tgx_device_name(self.h, buf, 256))
        return buf.value.decode()

    def memory_budget(self):
        f, t = C.c_double(), C.c_double()
        L.check(self._lib.tgx_memory_budget(self.h, C.byref(f), C.byref(t)), "gpu_memory_budget_init")
        return f.value, t.value

    # -- system and neighbours ---------------------------------------------------
    def set_system(self, positions, species, box, n_sites: Optional[int] = None):
        """positions (Nsc, 3), species codes (Nsc,), box (3, 3) rows a, b, c."""
        pos = _f64(positions)
        sp = _i32(species)
        bx = _f64(box, (9,))
        nsc = pos.shape[0]
        n = nsc if n_sites is None else int(n_sites)
        L.check(self._lib.tgx_set_system(self.h, n, nsc, L.dptr(pos), L.iptr(sp), L.dptr(bx)), "set_system")
        self.n_sites, self.n_sites_sc = n, nsc
        self._box = bx.copy()

    def set_positions(self, positions):
        pos = _f64(positions)
        L.check(self._lib.tgx_set_positions(self.h, L.dptr(pos)), "set_positions")

    def build_neighbors_list(self, rcut_max: float, indices=(1, 1, 1), do_list=None, fetch: bool = True,
                             rebuild: bool = True):
        """build_neighbors_list (src/neighbors.f90:243-605) on the device.

        Returns a dict with n_atom_pairs, n_neigh, neighbors_list (1-based),
        rjs, xyz (P, 3), thetas, phis, neighbor_species; with fetch=False only
        the pair count (the list stays resident for get_gap_soap/compute).
        rebuild=False is rebuild_neighbors_list = .false.: the resident list keeps
        its pairs and order and gets distances from the current positions.
        """
        if rebuild:
            ind = _i32(indices)
            dl = None if do_list is None else _i32(np.asarray(do_list).astype(np.int32))
            npairs = C.c_int64()
            L.check(self._lib.tgx_build_neighbors_list(self.h, float(rcut_max), L.iptr(ind), L.iptr(dl),
                                                       C.byref(npairs)), "build_neighbors_list")
            P = npairs.value
        else:
            L.check(self._lib.tgx_refresh_neighbors_list(self.h), "refresh_neighbors_list")
            P = self.n_atom_pairs
        self.n_atom_pairs = P
        if not fetch:
            return {"n_atom_pairs": P}
        out = {
            "n_atom_pairs": P,
            "n_neigh": np.empty(self.n_sites, np.int32),
            "neighbors_list": np.empty(P, np.int32),
            "rjs": np.empty(P),
            "thetas": np.empty(P),
            "phis": np.empty(P),
            "xyz": np.empty((P, 3)),
            "neighbor_species": np.empty(P, np.int32),
        }
        L.check(self._lib.tgx_get_neighbors_list(self.h, L.iptr(out["n_neigh"]), L.iptr(out["neighbors_list"]),
                                                 L.dptr(out["rjs"]), L.dptr(out["thetas"]), L.dptr(out["phis"]),
                                                 L.dptr(out["xyz"]), L.iptr(out["neighbor_species"])),
                "get_neighbors_list")
        return out

    # -- models ----------------------------------------------------------------------
    def soap_model(self, hypers: SoapTurbo, Qs, alphas, e0: float = 0.0) -> int:
        """Upload one soap_turbo block and its sparse set; Qs is (n_sparse, dim)."""
        Qs = _f64(Qs)
        al = _f64(alphas)
        if Qs.ndim != 2 or Qs.shape[1] != hypers.dim or Qs.shape[0] != al.shape[0]:
            raise ValueError(f"Qs must be (n_sparse, {hypers.dim}) with one alpha per row")
        s, keep = hypers.c_struct()
        st = _i32(hypers.species_types)
        mid = C.c_int()
        if hypers.compressed:
            dim, P_i, P_j, P_el = hypers.compression_arrays()
            L.check(self._lib.tgx_soap_model_create_compressed(
                self.h, C.byref(s), L.iptr(st), int(hypers.central_species), float(hypers.delta), float(hypers.zeta),
                float(e0), Qs.shape[0], Qs.shape[1], L.dptr(Qs), L.dptr(al), P_i.shape[0], L.iptr(P_i), L.iptr(P_j),
                L.dptr(P_el), C.byref(mid)), "soap_model_create_compressed")
        else:
            L.check(self._lib.tgx_soap_model_create(self.h, C.byref(s), L.iptr(st), int(hypers.central_species),
                                                    float(hypers.delta), float(hypers.zeta), float(e0), Qs.shape[0],
                                                    Qs.shape[1], L.dptr(Qs), L.dptr(al), C.byref(mid)),
                    "soap_model_create")
        self.models[mid.value] = hypers
        return mid.value

    def distance_2b_model(self, species1: int, species2: int, rcut: float, buffer: float, delta: float,
                          sigma: float, Qs, alphas, cutoff, e0: float = 0.0) -> int:
        """A distance_2b block (src/gap.f90:400-563): Qs (n_sparse,) distances,
        alphas and cutoff (n_sparse,), species as the system's codes."""
        q, a, c = _f64(Qs).reshape(-1), _f64(alphas).reshape(-1), _f64(cutoff).reshape(-1)
        if not (q.shape == a.shape == c.shape):
            raise ValueError("Qs, alphas and cutoff must have one entry per sparse point")
        mid = C.c_int()
        L.check(self._lib.tgx_distance_2b_create(self.h, int(species1), int(species2), float(rcut), float(buffer),
                                                 float(delta), float(sigma), float(e0), q.shape[0], L.dptr(q),
                                                 L.dptr(a), L.dptr(c), C.byref(mid)), "distance_2b_create")
        return mid.value

    def core_pot_model(self, species1: int, species2: int, x, V, yp1: float, ypn: float, cutoff: float = 1.0e10,
                       buffer: float = 1.0) -> int:
        """A core_pot block (src/gap.f90:565-721): the table V(x), end slopes
        (>= 0.99e30 for a natural end), core_pot_cutoff / core_pot_buffer."""
        xx, vv = _f64(x).reshape(-1), _f64(V).reshape(-1)
        if xx.shape != vv.shape:
            raise ValueError("x and V must have the same length")
        mid = C.c_int()
        L.check(self._lib.tgx_core_pot_create(self.h, int(species1), int(species2), xx.shape[0], L.dptr(xx),
                                              L.dptr(vv), float(yp1), float(ypn), float(cutoff), float(buffer),
                                              C.byref(mid)), "core_pot_create")
        return mid.value

    def angle_3b_model(self, species_center: int, species1: int, species2: int, rcut: float, buffer: float,
                       delta: float, sigma, Qs, alphas, cutoff, e0: float = 0.0, kernel_type: str = "exp") -> int:
        """An angle_3b block (src/gap.f90:724-1112): Qs (n_sparse, 3) triplet
        descriptors, alphas and cutoff (n_sparse,), sigma (3,), kernel_type
        "exp" or "pol", species as the system's codes."""
        q, a, c = _f64(Qs).reshape(-1, 3), _f64(alphas).reshape(-1), _f64(cutoff).reshape(-1)
        sg = _f64(sigma).reshape(-1)
        if not (q.shape[0] == a.shape[0] == c.shape[0]) or sg.shape != (3,):
            raise ValueError("Qs (n, 3), alphas and cutoff (n,) per sparse point and three sigmas")
        mid = C.c_int()
        L.check(self._lib.tgx_angle_3b_create(self.h, int(species_center), int(species1), int(species2), float(rcut),
                                              float(buffer), float(delta), L.dptr(sg), float(e0),
                                              str(kernel_type).encode(), q.shape[0], L.dptr(q), L.dptr(a), L.dptr(c),
                                              C.byref(mid)), "angle_3b_create")
        return mid.value

    def pair_energy_and_forces(self, model_id: int):
        """(energies (n,), forces (n, 3), virial (3, 3)) of a pair or triplet model on the resident list."""
        n = self.n_sites
        e, f, w = np.empty(n), np.empty((n, 3)), np.empty((3, 3))
        L.check(self._lib.tgx_pair_energy_and_forces(self.h, int(model_id), L.dptr(e), L.dptr(f), L.dptr(w)),
                "pair_energy_and_forces")
        return e, f, w

    def soap_backend_begin(self, model_id: int):
        L.check(self._lib.tgx_soap_backend_begin(self.h, int(model_id)), "soap_backend_begin")
        self._current = model_id

    def soap_backend_end(self):
        L.check(self._lib.tgx_soap_backend_end(self.h), "soap_backend_end")
        self._current = None

    # -- one block -------------------------------------------------------------------
    def get_gap_soap(self, path: str = "fp64_plain", do_forces: bool = True, energies=None, forces=None,
                     sites=None):
        """get_gap_soap for the current block: energies/forces are ASSIGNED for the
        atoms the block touches (src/gap_interface.f90:422-431). sites = (i_beg, i_end),
        1-based inclusive, restricts the centres to one batch of the reference's
        this_i_beg:this_i_end loop."""
        n = self.n_sites
        energies = np.zeros(n) if energies is None else energies
        forces = np.zeros((n, 3)) if forces is None else forces
        virial = np.zeros((3, 3))
        ns = C.c_int()
        npb = C.c_int64()
        if sites is None:
            L.check(self._lib.tgx_get_gap_soap(self.h, L.PATHS[path], int(do_forces), L.dptr(energies),
                                               L.dptr(forces), L.dptr(virial), C.byref(ns), C.byref(npb)),
                    "get_gap_soap")
        else:
            L.check(self._lib.tgx_get_gap_soap_sites(self.h, L.PATHS[path], int(do_forces), int(sites[0]),
                                                     int(sites[1]), L.dptr(energies), L.dptr(forces),
                                                     L.dptr(virial), C.byref(ns), C.byref(npb)), "get_gap_soap")
        return energies, forces, virial.T.copy(), ns.value, npb.value

    def set_memory_budget(self, gib: float):
        """Device memory (GiB) the FP64 plain path may hold for one block before it
        splits the block's centres into batches (0: 80% of the device)."""
        L.check(self._lib.tgx_set_memory_budget(self.h, float(gib)), "set_memory_budget")

    def get_soap(self, derivatives: bool = True):
        """Descriptors of the current block: soap (n_sites_block, D) and
        soap_cart_der (n_pairs_block, D, 3)."""
        ns = C.c_int()
        npb = C.c_int64()
        L.check(self._lib.tgx_soap_block_size(self.h, C.byref(ns), C.byref(npb)), "soap_block_size")
        D = self.models[self._current].dim
        soap = np.empty((ns.value, D))
        der = np.empty((npb.value, D, 3)) if derivatives else None
        L.check(self._lib.tgx_get_soap(self.h, L.dptr(soap), L.dptr(der)), "get_soap")
        return soap, der

    def local_property_predict(self, derivatives: bool = True):
        """local_property_predict (src/local_properties.f90:174-297) with the current
        block as the property model: V (n_sites_block,) and V_der (n_pairs_block, 3)."""
        ns = C.c_int()
        npb = C.c_int64()
        L.check(self._lib.tgx_soap_block_size(self.h, C.byref(ns), C.byref(npb)), "soap_block_size")
        V = np.empty(ns.value)
        Vd = np.empty((npb.value, 3)) if derivatives else None
        L.check(self._lib.tgx_local_property_predict(self.h, int(derivatives), L.dptr(V), L.dptr(Vd), None, None),
                "local_property_predict")
        return V, Vd

    def get_soap_energy_and_forces(self, soap, soap_der, n_neigh, neighbors_list, xyz, n_sites0, e0=0.0,
                                   do_forces=True, delta=None, zeta0=None):
        """src/gap.f90:66-259 with host arrays: soap (n, D), soap_der (P, D, 3),
        neighbors_list 1-based local indices, xyz (P, 3). delta and zeta0 default
        to the current block's and must equal them (TgxError otherwise)."""
        if delta is None or zeta0 is None:
            if self._current is None:
                raise L.TgxError("get_soap_energy_and_forces: no block selected (soap_backend_begin)")
            h = self.models[self._current]
            delta = h.delta if delta is None else delta
            zeta0 = h.zeta if zeta0 is None else zeta0
        soap = _f64(soap)
        n, D = soap.shape
        energies = np.zeros(n)
        forces = np.zeros((n_sites0, 3))
        virial = np.zeros((3, 3))
        P = 0
        args = [None] * 4
        if do_forces:
            sd = _f64(soap_der)
            P = sd.shape[0]
            args = [sd, _i32(n_neigh), _i32(neighbors_list), _f64(xyz)]
        L.check(self._lib.tgx_get_soap_energy_and_forces(
            self.h, n, D, P, L.dptr(soap), L.dptr(args[0]), float(delta), float(zeta0), float(e0),
            L.iptr(args[1]), L.iptr(args[2]),
            L.dptr(args[3]), int(n_sites0), int(do_forces), L.dptr(energies), L.dptr(forces), L.dptr(virial)),
            "get_soap_energy_and_forces")
        return energies, forces, virial.T.copy()

    # -- whole step --------------------------------------------------------------------
    def compute(self, path: str = "fp64_plain"):
        """E, F, W of every block on the resident list. Returns (energy, virial)."""
        e = C.c_double()
        v = np.zeros(9)
        L.check(self._lib.tgx_compute(self.h, L.PATHS[path], C.byref(e), L.dptr(v)), "compute")
        return e.value, v.reshape(3, 3).T.copy()

    def results(self):
        n = self.n_sites
        en = np.empty(n)
        f = np.empty((n, 3))
        L.check(self._lib.tgx_get_results(self.h, L.dptr(en), L.dptr(f)), "get_results")
        return en, f

    # -- profiling ---------------------------------------------------------------------
    def set_profiling(self, on: bool):
        L.check(self._lib.tgx_set_profiling(self.h, int(on)))

    def set_block_streams(self, mode):
        """Blocks of distinct central species on streams of their own: True / 1 (the
        default) on the surrogate and FP64 paths with the mixed path's blocks in
        turn, 2 on every path, False / 0 every path in turn."""
        L.check(self._lib.tgx_set_block_streams(self.h, int(mode)), "set_block_streams")

    def reset_profile(self):
        L.check(self._lib.tgx_reset_profile(self.h))

    def kernel_time(self, name: str = ""):
        c = C.c_int()
        t = C.c_double()
        L.check(self._lib.tgx_kernel_time(self.h, name.encode(), C.byref(c), C.byref(t)))
        return c.value, t.value

    def synchronize(self):
        L.check(self._lib.tgx_synchronize(self.h))

    def hbm_probe(self, nbytes: int = 1 << 30, reps: int = 5) -> float:
        """Read + write GB/s of the library's streaming copy kernel (best of reps)."""
        g = C.c_double()
        L.check(self._lib.tgx_hbm_probe(self.h, int(nbytes), int(reps), C.byref(g)), "hbm_probe")
        return g.value
