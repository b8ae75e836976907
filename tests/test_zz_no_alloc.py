"""GPU: the allocation-free get_soap_energy_and_forces (device arrays, work
buffers in the library's static device array) against the allocating entry
point on the same inputs, two calls in a row, the second with fewer sites and
pairs: energies and forces bit for bit, the virial to 1e-12 relative (a
fixed-order device sum against the host's sequential one)."""
import ctypes as C

import numpy as np
import pytest

from tests import helpers as H
from turbogap_amd import lib as L

pytestmark = pytest.mark.gpu


class _Device:
    """Device buffers for the test through the HIP runtime libtgx itself links
    (the test allocates; the entry point under test does not)."""

    H2D, D2H = 1, 2

    def __init__(self):
        L.load()
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, self.H2D) == 0
        return p

    def get(self, p, shape, dtype=np.float64):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _noalloc(ctx, dev, h, q, der, n_neigh, nlist, xyz, n0):
    args = [dev.put(x) for x in (q, der, n_neigh.astype(np.int32), nlist.astype(np.int32), xyz)]
    en, F, W = dev.put(np.zeros(q.shape[0])), dev.put(np.zeros((n0, 3))), dev.put(np.zeros((3, 3)))
    L.check(L.load().tgx_get_soap_energy_and_forces_noalloc(
        ctx.h, q.shape[0], q.shape[1], der.shape[0], args[0], args[1], h.delta, h.zeta, 0.0, args[2], args[3],
        args[4], n0, 1, en, F, W), "get_soap_energy_and_forces_noalloc")
    L.check(L.load().tgx_synchronize(ctx.h), "synchronize")
    return dev.get(en, q.shape[0]), dev.get(F, (n0, 3)), dev.get(W, (3, 3)).T.copy()


def test_noalloc_matches_allocating_entry(ctx, orc):
    sysm, h, Qs, al = H.carbon216(2000)
    nl = orc.build_neighbors_list(sysm.positions, sysm.species, sysm.box, h.rcut_max)
    spi = orc.species_index(nl["neighbor_species"], h.species_types)
    q, der = orc.get_soap(h, nl["n_neigh"], spi, nl["rjs"], nl["xyz"])
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.soap_backend_begin(ctx.soap_model(h, Qs, al))
    ns2 = 100
    p2 = int(nl["n_neigh"][:ns2].sum())
    cases = [(q, der, nl["n_neigh"], nl["neighbors_list"], nl["xyz"]),
             (q[:ns2], der[:p2], nl["n_neigh"][:ns2], nl["neighbors_list"][:p2], nl["xyz"][:p2])]
    dev = _Device()
    for qc, dc, nn, nlc, xc in cases:
        e_ref, f_ref, w_ref = ctx.get_soap_energy_and_forces(qc, dc, nn, nlc, xc, sysm.n)
        e, f, w = _noalloc(ctx, dev, h, qc, dc, nn, nlc, xc, sysm.n)
        assert np.array_equal(e, e_ref), np.abs(e - e_ref).max()
        assert np.array_equal(f, f_ref), np.abs(f - f_ref).max()
        assert np.abs(w - w_ref).max() <= 1e-12 * np.abs(w_ref).max()
        assert np.abs(f).max() > 0.0
    dev.free()
    # sizes beyond the work array are refused before any launch
    assert L.load().tgx_get_soap_energy_and_forces_noalloc(
        ctx.h, 20000, q.shape[1], 0, None, None, h.delta, h.zeta, 0.0, None, None, None, sysm.n, 0, None, None,
        None) == 1
    assert b"16384" in L.load().tgx_last_error() or b"16,384" in L.load().tgx_last_error()
    ctx.soap_backend_end()
