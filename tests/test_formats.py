"""Model and trajectory formats (host side, no GPU).

  * .gap files (src/read_files.f90:4641-5236): the reference test suite's
    silicon.gap fixture parses into its four blocks with the soap_turbo
    defaults and buffer -> rcut_soft conversion; a written two-species file loads back into the hypers
    and sparse sets it came from, bit for bit.
  * Extended XYZ (src/xyz.f90:68-487): the first frame of a trajectory the
    reference wrote (tests/regression/cases/xrd_mad/expected) is re-written
    from its own parsed values, line for line (the stress, computed from the
    printed virial, to its last printed digit).
"""
import io
import os
import shlex

import numpy as np
import pytest

from turbogap_amd import gap_file as G
from turbogap_amd import structures as S
from turbogap_amd import xyz as X
from turbogap_amd.sparse_io import write_alphas_and_descriptors

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_silicon_gap_blocks():
    blocks = G.read_gap_file(os.path.join(GOLD, "silicon.gap"))
    assert [b.kind for b in blocks] == ["distance_2b", "angle_3b", "soap_turbo", "core_pot"]
    b2, b3, bs, bc = blocks
    assert b2.values["rcut"] == ["5.0"] and b2.values["species1"] == ["Si"]
    assert b3.values["sigma"] == ["4.", "4.", "4."] and b3.values["kernel_type"] == ["pol"]
    assert bc.values["core_pot_file"] == ["gap_files/corepot.dat"]
    h = bs.hypers
    assert bs.species_names == ["Si"] and h.species_types == [14]
    assert h.alpha_max == [8] and h.l_max == 8 and h.dim == 8 * 9 // 2 * 9
    assert h.rcut_hard == [5.0] and h.rcut_soft == [4.0]  # buffer 1.0 is a width
    assert h.atom_sigma_r == [0.5] and h.atom_sigma_t == [0.5]
    assert h.atom_sigma_r_scaling == [0.0] and h.amplitude_scaling == [1.0] and h.nf == [4.0]
    assert h.central_weight == [1.0] and h.global_scaling == [1.0]
    assert h.basis == "poly3gauss" and h.radial_enhancement == 1
    assert h.zeta == 6.0 and h.delta == 1.0 and h.central_species == 1
    assert bs.compress and bs.file_alphas == "gap_files/alphas_mb.dat"
    # the block's desc_sparse file is not in the reference tree (the block
    # loads with a synthesised one in test_compression.py)
    with pytest.raises(FileNotFoundError):
        G.load_soap_blocks(os.path.join(GOLD, "silicon.gap"), base_dir=GOLD)


def test_gap_file_defaults_and_fallbacks(tmp_path):
    p = tmp_path / "m.gap"
    p.write_text("gap_beg soap_turbo\nn_species = 2\nspecies = C H\nrcut = 4.0 3.0\nbuffer = 0.0 0.5\n"
                 "atom_sigma_r = 0.4 0.3\natom_sigma_t = 0.4 0.3\nn_max = 6 4\nl_max = 5\n"
                 "basis = \"unknown\"\nradial_enhancement = 7\ngap_end\n")
    (b,) = G.read_gap_file(str(p))
    h = b.hypers
    assert h.species_types == [6, 1] and h.alpha_max == [6, 4] and h.n_max == 10
    assert h.rcut_soft == [4.0, 2.5]  # zero buffer: rcut_soft = rcut
    assert h.basis == "poly3" and h.radial_enhancement == 0
    assert h.nf == [4.0, 4.0] and h.central_weight == [1.0, 1.0] and h.zeta == 2.0


def test_angle_3b_block_defaults(tmp_path):
    """An angle_3b block without delta, sigma or kernel_type takes the defaults of
    src/types.f90:154-170 (1, 1 1 1, "exp"); Z_center / Z1 / Z2 name the species;
    a descriptor file without three values per sparse point is rejected."""
    (tmp_path / "a.dat").write_text("0.5 1.0\n-0.25 0.5\n")
    (tmp_path / "d.dat").write_text("\n".join(["4.0", "1.0", "3.0", "5.0", "0.5", "2.5"]) + "\n")
    (tmp_path / "bad.dat").write_text("4.0\n1.0\n3.0\n5.0\n")
    p = tmp_path / "m.gap"
    p.write_text("gap_beg angle_3b\nZ_center = C\nZ1 = C\nZ2 = H\nrcut = 3.5\n"
                 "desc_sparse = 'd.dat'\nalphas_sparse = 'a.dat'\ngap_end\n")
    ((kind, kw),) = G.load_pair_blocks(str(p), base_dir=str(tmp_path))
    assert kind == "angle_3b"
    assert (kw["species_center"], kw["species1"], kw["species2"]) == (6, 6, 1)
    assert kw["delta"] == 1.0 and kw["sigma"] == [1.0, 1.0, 1.0] and kw["kernel_type"] == "exp"
    assert kw["buffer"] == 0.5 and kw["e0"] == 0.0 and kw["rcut"] == 3.5
    np.testing.assert_array_equal(kw["Qs"], [[4.0, 1.0, 3.0], [5.0, 0.5, 2.5]])
    np.testing.assert_array_equal(kw["alphas"], [0.5, -0.25])
    np.testing.assert_array_equal(kw["cutoff"], [1.0, 0.5])
    p.write_text(p.read_text().replace("'d.dat'", "'bad.dat'"))
    with pytest.raises(ValueError):
        G.load_pair_blocks(str(p), base_dir=str(tmp_path))


def test_gap_file_round_trip(tmp_path):
    models = []
    lines = []
    rng = np.random.default_rng(5)
    for k, h in enumerate(S.soap_binary()):
        Qs = rng.normal(size=(7, h.dim))
        al = rng.normal(scale=1e-3, size=7)
        write_alphas_and_descriptors(str(tmp_path / f"d{k}"), str(tmp_path / f"a{k}"), al, Qs)
        models.append((h, Qs, al))
        lines += ["gap_beg soap_turbo", "n_species = 2", "species = A B", f"central_species = {h.central_species}",
                  "rcut = 3.7 3.7", "buffer = 0.5 0.5", "atom_sigma_r = 0.5 0.5", "atom_sigma_t = 0.5 0.5",
                  "n_max = 8 8", "l_max = 8", "basis = \"poly3gauss\"", "zeta = 4", "delta = 1.0",
                  f"desc_sparse = 'd{k}'", f"alphas_sparse = 'a{k}'", "compress_soap = .false.", "gap_end", ""]
    (tmp_path / "binary.gap").write_text("\n".join(lines))
    got = G.load_soap_blocks(str(tmp_path / "binary.gap"), species_map={"A": S.SPECIES_A, "B": S.SPECIES_B},
                             base_dir=str(tmp_path))
    assert len(got) == 2
    for (h, Qs, al), (h2, Qs2, al2) in zip(models, got):
        assert h2 == h
        assert np.array_equal(Qs2, Qs) and np.array_equal(al2, al)


def _parse_frame(path):
    lines = open(path).read().splitlines()
    n = int(lines[0])
    kv = {}
    for tok in shlex.split(lines[1]):
        k, _, v = tok.partition("=")
        kv[k] = v
    props = kv["Properties"].split(":")
    cols = [(props[i], props[i + 1], int(props[i + 2])) for i in range(0, len(props), 3)]
    rows = [ln.split() for ln in lines[2:2 + n]]
    data, c = {}, 0
    for name, typ, m in cols:
        data[name] = [r[c:c + m] for r in rows]
        c += m
    return lines, n, kv, data


def test_extxyz_frame_matches_reference_writer():
    path = os.path.join(GOLD, "xrd_mad_frame0.xyz")
    lines, n, kv, data = _parse_frame(path)
    f = lambda name: np.array(data[name], dtype=np.float64)
    head = lines[1]
    # the energy string: the terms between energy=... and virial=...
    estr = head[head.index(" energy_soap="):head.index(' virial="')].strip()
    terms = [(t.split("=")[0], float(t.split("=")[1])) for t in estr.split()]
    assert X.energy_string(terms) == estr  # two blanks after the first term, as get_xyz_energy_string writes
    lat = np.array(kv["Lattice"].split(), dtype=np.float64).reshape(3, 3)
    W = np.array(kv["virial"].split(), dtype=np.float64).reshape(3, 3)
    le = f("local_energy")[:, 0]
    assert abs(le.sum() - float(kv["energy"])) < 1e-5
    out = io.StringIO()
    X.write_extxyz(out, [r[0] for r in data["species"]], f("pos"), velocities=f("velocities"), forces=f("forces"),
                   local_energies=le, local_properties={"core_electron_be": f("core_electron_be")[:, 0]},
                   fix_atoms=np.array(data["fix_atoms"]) == "T", cell=lat, temperature=float(kv["temperature"]),
                   pressure=1e30, time_step=float(kv["time_step"]), time=float(kv["time"]),
                   energy=float(kv["energy"]), energies=X.energy_string(terms), virial=W, stress=True, volume=True,
                   step=int(kv["step"]))
    mine = out.getvalue().splitlines()
    assert len(mine) == len(lines)
    assert mine[0] == lines[0]
    assert mine[2:] == lines[2:]
    # header: identical apart from the stress, recomputed from the rounded virial
    cut = lambda s: s[:s.index(' stress="')] + s[s.index('" volume='):]
    assert cut(mine[1]) == cut(lines[1])
    st = lambda s: np.array(s[s.index(' stress="') + 9:s.index('" volume=')].split(), dtype=np.float64)
    np.testing.assert_allclose(st(mine[1]), st(lines[1]), rtol=0, atol=2e-8)
