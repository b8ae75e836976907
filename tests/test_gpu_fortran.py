"""A Fortran host drives the device path through the iso_c_binding shim
(turbogap_amd/fortran/tgx_turbogap.f90): the TurboGAP call sequence
gpu_context_init -> build_neighbors_list -> soap_backend_begin ->
get_gap_soap -> soap_backend_end on the 216-atom carbon box, with the sparse set
read from the reference's alphas / desc_sparse file format. Results must equal
the committed fixture (energy 1e-8 relative, forces 1e-6 eV/Å)."""
import os
import pathlib
import subprocess

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu
ROOT = pathlib.Path(__file__).resolve().parents[1]
EXE = ROOT / "turbogap_amd" / "fortran" / "_build" / "tgx_example"
GOLDEN = ROOT / "tests" / "golden" / "carbon216_golden.npz"


def _write_inputs(d: pathlib.Path):
    sysm, h, Qs, al = H.carbon216(200)
    L = float(sysm.box[0, 0])
    with open(d / "positions.dat", "w") as f:
        f.write(f"{len(sysm.positions)} {L:.17g}\n")
        for p, s in zip(sysm.positions, sysm.species):
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g} {int(s)}\n")
    np.savetxt(d / "alphas.dat", al, fmt="%.17g")
    np.savetxt(d / "desc_sparse.dat", Qs.reshape(-1), fmt="%.17g")


def test_fortran_host_matches_fixture(tmp_path):
    if not EXE.exists():
        pytest.fail(f"{EXE} not built (make -C turbogap_amd/fortran)")
    _write_inputs(tmp_path)
    out = subprocess.run([str(EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {}
    for line in out.stdout.splitlines():
        k, *v = line.split()
        vals[k] = [float(x) for x in v]
    ref = np.load(GOLDEN)
    assert int(vals["pairs"][0]) == len(ref["rjs"])
    E = vals["energy"][0]
    assert abs(E - ref["energies"].sum()) <= 1e-8 * abs(ref["energies"].sum())
    assert np.abs(np.array(vals["force1"]) - ref["forces"][0]).max() <= 1e-6
    assert abs(vals["virial_trace"][0] - np.trace(ref["virial"])) <= 3e-6
