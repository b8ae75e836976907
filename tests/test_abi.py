"""CPU checks of the C ABI: libtgx.so loads, exports every function include/tgx.h
declares, rejects bad arguments with a status and message (no GPU needed for
these), and the seeded generators are deterministic."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from turbogap_amd import lib as L
from turbogap_amd import structures as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "tgx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(tgx_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    names = set(declared()) - {"tgx_last_error"}
    assert names <= set(L._SIGS), sorted(names - set(L._SIGS))


def test_error_paths_without_gpu():
    lib = L.load()
    assert lib.tgx_generate_rsa(-1, None, 1.0, C.c_uint64(0), None) != 0
    assert b"bad arguments" in lib.tgx_last_error()
    assert lib.tgx_set_system(None, 1, 1, None, None, None) != 0


def test_generators_deterministic_and_valid():
    a = S.rsa(500, 18.0, 1.2, 42)
    b = S.rsa(500, 18.0, 1.2, 42)
    assert np.array_equal(a, b)
    d = a[:, None, :] - a[None, :, :]
    d -= 18.0 * np.round(d / 18.0)
    r = np.linalg.norm(d, axis=2) + np.eye(500) * 99
    assert r.min() >= 1.2
    p = S.permutation(1000, 3)
    assert sorted(p.tolist()) == list(range(1000)) and np.array_equal(p, S.permutation(1000, 3))
    x = S.normal(20000, 1e-3, 5)
    assert abs(x.std() - 1e-3) < 3e-5 and abs(x.mean()) < 3e-5


def test_config_shapes():
    s1 = S.binary_system(4096, seed=1)
    assert s1.n == 4096 and abs(s1.box[0, 0] - 86.8 * (1 / 16) ** (1 / 3)) < 1e-12
    assert (s1.species == S.SPECIES_A).sum() == 2048
    h0 = S.soap_carbon()
    assert h0.dim == 324 and h0.rcut_soft == [3.2]
    assert S.soap_binary()[0].dim == 1224


def test_fortran_shim_binds_declared_symbols():
    """Every bind(C) name in the Fortran shim is a libtgx export the header declares."""
    src = open(os.path.join(ROOT, "turbogap_amd", "fortran", "tgx_turbogap.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(tgx_\w+)"\)', src))
    assert {"tgx_init", "tgx_build_neighbors_list", "tgx_get_gap_soap", "tgx_soap_backend_begin"} <= bound
    assert bound <= set(declared()), bound - set(declared())
