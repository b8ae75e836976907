"""A Fortran host drives the device path through the iso_c_binding shim
(turbogap_amd/fortran/tgx_turbogap.f90): the TurboGAP call sequence
gpu_context_init -> build_neighbors_list -> soap_backend_begin ->
get_gap_soap -> soap_backend_end on the 216-atom carbon box, with the sparse set
read from the reference's alphas / desc_sparse file format. Results must equal
the committed fixture (energy 1e-8 relative, forces 1e-6 eV/Å). A second host
(tgx_pair_example) creates and evaluates the distance_2b, angle_3b and core_pot
families through the same shim, checked against the oracle. A third
(tgx_md_example) runs the device MD loop through the shim's MD bindings and
must match the Python driver making the same calls. A fourth
(tgx_ref_example) makes the reference's own calls with its argument lists
(get_gap_soap with 66 arguments over site batches, get_soap)."""
import os
import pathlib
import shutil
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


PAIR_EXE = ROOT / "turbogap_amd" / "fortran" / "_build" / "tgx_pair_example"


def test_fortran_pair_families_match_oracle(tmp_path):
    """distance_2b, angle_3b and core_pot blocks of the reference suite's silicon
    potential created and evaluated from Fortran (tgx_distance_2b_model /
    tgx_angle_3b_model / tgx_core_pot_model, add_pair_contribution) on a
    216-atom silicon crystal, against the oracle at the tolerances of the
    Python-side parity tests."""
    from oracle import angle_3b as A3
    from oracle import oracle as O
    from oracle import pair_oracle as PO
    from tests.test_angle_3b import silicon_3b
    from tests.test_pair_models import cancel_scale, silicon_crystal, silicon_models

    if not PAIR_EXE.exists():
        pytest.fail(f"{PAIR_EXE} not built (make -C turbogap_amd/fortran)")
    sysm = silicon_crystal()
    m = silicon_models()
    b2, cp, b3 = m["distance_2b"], m["core_pot"], silicon_3b()
    with open(tmp_path / "positions.dat", "w") as f:
        f.write(f"{sysm.n} {float(sysm.box[0, 0]):.17g}\n")
        for p, s in zip(sysm.positions, sysm.species):
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g} {int(s)}\n")
    with open(tmp_path / "2b.dat", "w") as f:
        f.write(f"{b2['species1']} {b2['species2']} {b2['rcut']:.17g} {b2['delta']:.17g} {b2['sigma']:.17g} "
                f"{len(b2['Qs'])}\n")
        for q, a, c in zip(b2["Qs"], b2["alphas"], b2["cutoff"]):
            f.write(f"{q:.17g} {a:.17g} {c:.17g}\n")
    with open(tmp_path / "3b.dat", "w") as f:
        sg = " ".join(f"{x:.17g}" for x in b3["sigma"])
        f.write(f"{b3['species_center']} {b3['species1']} {b3['species2']} {b3['rcut']:.17g} {b3['delta']:.17g} {sg} "
                f"{b3['kernel_type']} {len(b3['alphas'])}\n")
        for q, a, c in zip(b3["Qs"], b3["alphas"], b3["cutoff"]):
            f.write(f"{q[0]:.17g} {q[1]:.17g} {q[2]:.17g} {a:.17g} {c:.17g}\n")
    with open(tmp_path / "corepot.dat", "w") as f:
        f.write(f"{cp['species1']} {cp['species2']} {len(cp['x'])} {cp['yp1']:.17g} {cp['ypn']:.17g}\n")
        for x, v in zip(cp["x"], cp["V"]):
            f.write(f"{x:.17g} {v:.17g}\n")
    out = subprocess.run([str(PAIR_EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr

    def read(kind):
        rows = (tmp_path / f"out_{kind}.dat").read_text().split("\n")
        W = np.array([float(t) for t in rows[0].split()]).reshape(3, 3).T
        body = np.array([[float(t) for t in r.split()] for r in rows[1:1 + sysm.n]])
        return body[:, 0], body[:, 1:], W

    def nl(rc):
        return O.build_neighbors_list(sysm.positions, sysm.species, sysm.box, rc)

    ref = {
        "distance_2b": PO.distance_2b_energy_and_forces(nl(b2["rcut"]), sysm.species, sysm.n, b2["species1"],
                                                        b2["species2"], b2["rcut"], 0.5, b2["delta"], b2["sigma"], 0.0,
                                                        b2["Qs"], b2["alphas"], b2["cutoff"]),
        "angle_3b": A3.angle_3b_energy_and_forces(nl(b3["rcut"]), sysm.species, sysm.n, b3["species_center"],
                                                  b3["species1"], b3["species2"], b3["rcut"], 0.5, b3["delta"],
                                                  b3["sigma"], 0.0, b3["Qs"], b3["alphas"], b3["cutoff"],
                                                  b3["kernel_type"]),
        "core_pot": PO.core_pot_energy_and_forces(nl(float(cp["x"].max())), sysm.species, sysm.n, cp["species1"],
                                                  cp["species2"], PO.core_pot_model(cp["x"], cp["V"], cp["yp1"],
                                                                                    cp["ypn"])),
    }
    scale = {"distance_2b": cancel_scale("distance_2b", b2), "core_pot": 1.0,
             "angle_3b": max(1.0, 2 * b3["delta"] ** 2 * np.abs(b3["alphas"] * b3["cutoff"]).sum())}
    for kind, (Eo, Fo, Wo) in ref.items():
        e, F, W = read(kind)
        s = scale[kind]
        assert np.abs(e - Eo).max() <= 1e-12 * s, kind
        assert np.abs(F - Fo).max() <= 1e-11 * max(s, np.abs(Fo).max()), kind
        assert np.abs(W - Wo).max() <= 1e-11 * max(s, np.abs(Wo).max()), kind


MD_EXE = ROOT / "turbogap_amd" / "fortran" / "_build" / "tgx_md_example"


def test_fortran_md_loop_matches_python_driver(tmp_path):
    """Five NVT steps of the carbon box driven from Fortran through the device
    MD bindings (md_build_list, md_forces, remove_cm_vel_device,
    velocity_verlet_device, md_kinetic, md_pressure, berendsen_thermostat_device,
    wrap_pbc_device) against the Python driver making the same library calls in
    the same order: identical energies, temperatures, pressures, positions and
    velocities."""
    from turbogap_amd import Context
    from turbogap_amd import structures as S
    from turbogap_amd.md import MD, maxwell_velocities

    if not MD_EXE.exists():
        pytest.fail(f"{MD_EXE} not built (make -C turbogap_amd/fortran)")
    _write_inputs(tmp_path)
    sysm, h, Qs, al = H.carbon216(200)
    mass = S.masses_for(sysm.species)
    v0 = maxwell_velocities(mass, 900.0, seed=12)
    with open(tmp_path / "md_state.dat", "w") as f:
        for v, m in zip(v0, mass):
            f.write(f"{v[0]:.17g} {v[1]:.17g} {v[2]:.17g} {m:.17g}\n")
    out = subprocess.run([str(MD_EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    steps, atoms = [], []
    for line in out.stdout.splitlines():
        k, *v = line.split()
        if k == "step":
            steps.append([float(x) for x in v[1:]])
        elif k == "atom":
            atoms.append([float(x) for x in v])
    steps, atoms = np.array(steps), np.array(atoms)
    assert steps.shape == (5, 3) and atoms.shape == (sysm.n, 6)
    ctx = Context(0)
    try:
        ctx.set_system(sysm.positions, sysm.species, sysm.box)
        ctx.soap_model(h, Qs, al)
        md = MD(ctx, mass, v0, 3.7, "fp64_adjoint", fused=False)
        ref = []
        for _ in range(5):
            E, ek, T = md.step(1.0, T0=600.0, tau=50.0)
            ref.append([E, T, md.pressure])
        x, v = md.state()
    finally:
        ctx.close()
    np.testing.assert_allclose(steps, np.array(ref), rtol=1e-15, atol=0)
    np.testing.assert_allclose(atoms[:, :3], x, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(atoms[:, 3:], v, rtol=1e-15, atol=1e-18)


REF_EXE = ROOT / "turbogap_amd" / "fortran" / "_build" / "tgx_ref_example"


def _run_ref_example(d, sysm, models, names, n_batches, compress_file="none"):
    with open(d / "positions.dat", "w") as f:
        f.write(f"{sysm.n} {float(sysm.box[0, 0]):.17g}\n")
        for p, s in zip(sysm.positions, sysm.species):
            f.write(f"{p[0]:.17g} {p[1]:.17g} {p[2]:.17g} {int(s)} {names[int(s)]}\n")
    h0 = models[0][0]
    with open(d / "blocks.dat", "w") as f:
        f.write(f"{len(models)} {n_batches} {h0.rcut_hard[0]:.17g} {h0.rcut_soft[0]:.17g} {h0.zeta:.17g} "
                f"{h0.radial_enhancement}\n")
        for k, (h, Qs, al) in enumerate(models):
            np.savetxt(d / f"alphas_{k}.dat", al, fmt="%.17g")
            np.savetxt(d / f"desc_{k}.dat", Qs.reshape(-1), fmt="%.17g")
            f.write(f"alphas_{k}.dat desc_{k}.dat {h.central_species} {len(h.species_types)} {compress_file}\n")
            f.write(" ".join(names[int(c)] for c in h.species_types) + "\n")
    out = subprocess.run([str(REF_EXE)], cwd=d, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    vals, F = {}, np.zeros((sysm.n, 3))
    for line in (d / "out.dat").read_text().splitlines():
        k, *v = line.split()
        if k == "force":
            F[int(v[0]) - 1] = [float(x) for x in v[1:]]
        else:
            vals[k] = np.array([float(x) for x in v])
    return vals, F


@pytest.mark.parametrize("system,n_batches", [("carbon216", 1), ("carbon216", 3), ("binary4096", 3),
                                              ("silicon_compressed", 2)])
def test_fortran_reference_call_surface(tmp_path, orc, system, n_batches):
    """tgx_ref_example makes TurboGAP's own calls with the reference's argument
    lists (gpu_context_init(params, rank, n_omp), build_neighbors_list with 20
    arguments, soap_backend_begin(hypers), get_gap_soap with 66 arguments over
    this_i_beg:this_i_end batches, soap_backend_end, per-batch energies and
    forces zeroed and added as src/turbogap.f90:1281-1373 does): the totals
    equal the oracle's (energy 1e-8 relative, forces and virial 1e-6 in the
    potential's force scale), also for the compressed silicon block with the
    reference suite's compress.dat. On one-block systems it also calls get_soap
    on the whole list, whose descriptor and derivative sums equal the oracle's
    to 1e-10 relative."""
    if not REF_EXE.exists():
        pytest.fail(f"{REF_EXE} not built (make -C turbogap_amd/fortran)")
    cfile = "none"
    if system == "carbon216":
        sysm, h, Qs, al = H.carbon216(200)
        models = [(h, Qs, al)]
        names = {6: "C"}
    elif system == "silicon_compressed":
        sysm, h, Qs, al = H.silicon_compressed()
        models = [(h, Qs, al)]
        names = {14: "Si"}
        # list-directed reads end at a slash: the file goes beside the inputs
        shutil.copy(pathlib.Path(H.GOLD) / "gap_files" / "compress.dat", tmp_path / "compress.dat")
        cfile = "compress.dat"
    else:
        sysm, models = H.subset(300, 4096)
        names = {1: "A", 2: "B"}
    vals, F = _run_ref_example(tmp_path, sysm, models, names, n_batches, cfile)
    Eo, eno, Fo, Wo, nl = orc.compute(sysm, models)
    fs, ws = max(1.0, np.abs(Fo).max()), max(1.0, np.abs(Wo).max())
    assert abs(vals["energy"][0] - Eo) <= 1e-8 * abs(Eo)
    assert np.abs(F - Fo).max() <= 1e-6 * fs
    assert np.abs(vals["virial"].reshape(3, 3).T - Wo).max() <= 1e-6 * ws
    if len(models) == 1:
        h = models[0][0]
        ps = np.zeros(nl["n_atom_pairs"], np.int32)
        soap, der = orc.get_soap(h, nl["n_neigh"], ps, nl["rjs"], nl["xyz"])
        assert abs(vals["soap_sum"][0] - soap.sum()) <= 1e-10 * np.abs(soap).sum()
        assert abs(vals["soap_der_abs_sum"][0] - np.abs(der).sum()) <= 1e-10 * np.abs(der).sum()
