"""local_property_predict (src/local_properties.f90:174-297).

CPU: the oracle's V_der of a pair (i, j) is dV_i/dr_j by central differences
of its own V (h = 1e-5 A, through the oracle's SOAP), on the 216-atom carbon
box with V0 set so that part of the centres clamp to zero.
GPU: the library's V and V_der against the oracle on the same box and sparse
set: V to 1e-10 relative, V_der to 1e-9.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import local_property as LP
from tests import helpers as H
from turbogap_amd import structures as S


def _setup():
    sysm, h, Qs, al = H.carbon216(200)
    return sysm, h, Qs, al


def _oracle_V(h, Qs, al, V0, sysm, derivatives):
    nl = O.build_neighbors_list(sysm.positions, sysm.species, sysm.box, h.rcut_max)
    spi = O.species_index(nl["neighbor_species"], h.species_types)
    q, der = O.get_soap(h, nl["n_neigh"], spi, nl["rjs"], nl["xyz"], derivatives=derivatives)
    V, Vd = LP.local_property_predict(q, Qs, al, V0, h.delta, h.zeta, der, nl["n_neigh"])
    return V, Vd, nl


def _v0(h, Qs, al, sysm):
    V, _, _ = _oracle_V(h, Qs, al, 0.0, sysm, False)
    return -float(np.median(V))  # about half of the centres clamp


def test_oracle_vder_is_the_gradient_of_v():
    sysm, h, Qs, al = _setup()
    V0 = _v0(h, Qs, al, sysm)
    V, Vd, nl = _oracle_V(h, Qs, al, V0, sysm, True)
    assert 0 < np.sum(V == 0.0) < sysm.n
    off = np.concatenate([[0], np.cumsum(nl["n_neigh"])])
    rng = np.random.default_rng(0)
    hstep = 1e-5
    checked = 0
    for i in rng.permutation(sysm.n):
        if V[i] == 0.0 or checked >= 3:
            continue
        checked += 1
        p = off[i] + 1 + int(rng.integers(nl["n_neigh"][i] - 1))  # a non-self pair of i
        j = int(nl["neighbors_list"][p]) - 1
        for c in range(3):
            vals = []
            for s in (1, -1):
                x = sysm.positions.copy()
                x[j, c] += s * hstep
                vals.append(_oracle_V(h, Qs, al, V0, S.System(x, sysm.species, sysm.box), False)[0][i])
            fd = (vals[0] - vals[1]) / (2 * hstep)
            assert abs(fd - Vd[p, c]) < 1e-6 * max(1.0, abs(Vd[p, c])), (i, j, c, fd, Vd[p, c])
    assert checked == 3


@pytest.mark.gpu
def test_gpu_local_property_matches_oracle(ctx):
    sysm, h, Qs, al = _setup()
    V0 = _v0(h, Qs, al, sysm)
    ctx.set_system(sysm.positions, sysm.species, sysm.box)
    ctx.build_neighbors_list(h.rcut_max, fetch=False)
    mid = ctx.soap_model(h, Qs, al, e0=V0)
    ctx.soap_backend_begin(mid)
    V, Vd = ctx.local_property_predict(derivatives=True)
    ctx.soap_backend_end()
    Vo, Vdo, _ = _oracle_V(h, Qs, al, V0, sysm, True)
    np.testing.assert_allclose(V, Vo, rtol=1e-10, atol=1e-12)
    assert np.array_equal(V == 0.0, Vo == 0.0)
    assert np.abs(Vd - Vdo).max() <= 1e-9
