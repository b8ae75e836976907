"""TurboGAP ``.gap`` potential files (host side).

A file is a sequence of blocks ``gap_beg <type>`` ... ``gap_end``; each line
inside is ``keyword = value [value ...]``, read list-directed the way
src/read_files.f90:4641-5236 reads them (quotes around strings optional,
Fortran logicals ``.true.``/``.false.``/``T``/``F``, reals such as ``4.``).

read_gap_file returns every block with its raw values. soap_turbo blocks also
get their hypers as a SoapTurbo, with the reference's semantics:

  * per-species lists are ``n_species`` values wide, in the order of
    ``species``; ``n_max`` gives alpha_max per species (read_files.f90:4769-4780);
  * defaults nf = 4, buffer = 0.5 (as a width), scalings 0, amplitude_scaling 1,
    central_weight 1, global_scaling 1 (read_files.f90:4650-4668); basis
    poly3 and scaling_mode polynomial when absent or not understood;
  * ``buffer`` is a width: rcut_soft = rcut - buffer, and a zero buffer means
    rcut_soft = rcut (read_files.f90:5101-5109);
  * radial_enhancement outside 0..2 falls back to 0 (read_files.f90:4785-4795).

Species names map to the integer codes the caller uses for atoms through
``species_map`` (element symbols default to their atomic numbers).

load_soap_blocks reads the sparse sets of the soap_turbo blocks
(alphas_sparse, desc_sparse; sparse_io) into (hypers, Qs, alphas) triples for
Context.soap_model. A compressed block (compress_soap = .true.) takes its
matrix from file_compress_soap (src/read_files.f90:5166-5211; the file takes
precedence) or from compress_mode "trivial" (the recipe of the reference's
tools/compress_indices/compress_indices.py); other modes are rejected.

soap_turbo blocks describe atoms with this library's SOAP conventions (radial
basis, quadrature, normalisation). A sparse set written by soap_turbo itself
was made with soap_turbo's, which this library does not reproduce bit for bit,
so load_soap_blocks warns (SparseSetWarning) that the energies of such a
potential are not those soap_turbo would give. load_pair_blocks reads distance_2b blocks (alphas and cutoff per
line, one distance per line; src/read_files.f90:722-742, 5238-5305) and
core_pot blocks (a table headed by n, yp1, ypn; src/read_files.f90:5392-5427)
and angle_3b blocks (alphas and cutoff per line, three descriptor components
per sparse point, one per line; src/read_files.f90:743-770, 5305-5384;
defaults delta 1, sigma 1 1 1, kernel_type "exp", src/types.f90:154-170)
into keyword arguments of Context.distance_2b_model / core_pot_model /
angle_3b_model. The other block types are returned by read_gap_file but not
evaluated.
"""
from __future__ import annotations

import dataclasses
import os
import shlex
import warnings
from typing import Dict, List, Optional, Tuple

import numpy as np

from .hypers import SoapTurbo, compress_trivial, read_compress_file
from .sparse_io import read_alphas_and_descriptors


class SparseSetWarning(UserWarning):
    """A sparse set made by another SOAP implementation is being evaluated."""

_ELEMENTS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
             "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
             "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
             "Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og").split()
ATOMIC_NUMBER = {s: z for z, s in enumerate(_ELEMENTS, start=1)}


@dataclasses.dataclass
class GapBlock:
    kind: str                                  # soap_turbo, distance_2b, angle_3b, core_pot, ...
    values: Dict[str, List[str]]               # keyword -> raw value tokens, in file order
    hypers: Optional[SoapTurbo] = None         # soap_turbo blocks
    species_names: Optional[List[str]] = None  # soap_turbo: the block's species, in order
    file_desc: Optional[str] = None
    file_alphas: Optional[str] = None
    compress: bool = False


def _tokens(line: str) -> List[str]:
    lex = shlex.shlex(line, posix=True)
    lex.whitespace += ",="
    lex.whitespace_split = True
    lex.commenters = "!#"
    return list(lex)


def _logical(tok: str) -> bool:
    t = tok.strip().lower().lstrip(".")
    if t.startswith("t"):
        return True
    if t.startswith("f"):
        return False
    raise ValueError(f"not a Fortran logical: {tok!r}")


def _floats(vals: List[str], n: int, key: str) -> List[float]:
    if len(vals) < n:
        raise ValueError(f"{key}: expected {n} values, got {len(vals)}")
    return [float(v.replace("d", "e").replace("D", "e")) for v in vals[:n]]


def _species_code(name: str, species_map: Optional[Dict[str, int]]) -> int:
    if species_map is not None and name in species_map:
        return int(species_map[name])
    if name in ATOMIC_NUMBER:
        return ATOMIC_NUMBER[name]
    raise ValueError(f"species {name!r} has no code: pass species_map")


def _soap_hypers(values: Dict[str, List[str]], species_map) -> Tuple[SoapTurbo, List[str]]:
    if "n_species" not in values:
        raise ValueError("soap_turbo block without n_species")
    n = int(values["n_species"][0])

    def per_species(key, default=None):
        if key in values:
            return _floats(values[key], n, key)
        if default is None:
            raise ValueError(f"soap_turbo block without {key}")
        return [default] * n

    names = values.get("species", [])[:n]
    if len(names) != n:
        raise ValueError(f"species: expected {n} names")
    rcut = per_species("rcut")
    buffer = per_species("buffer", 0.5)
    basis = values.get("basis", ["poly3"])[0]
    if basis not in ("poly3", "poly3gauss", "poly3operator"):
        basis = "poly3"
    if basis == "poly3operator":
        raise NotImplementedError("soap_turbo basis poly3operator is not implemented")
    re = int(values.get("radial_enhancement", ["0"])[0])
    if re < 0 or re > 2:
        re = 0
    alpha_max = [int(float(v)) for v in values["n_max"][:n]] if "n_max" in values else None
    if alpha_max is None or len(alpha_max) != n:
        raise ValueError(f"n_max: expected {n} values")
    h = SoapTurbo.from_buffer(
        rcut=rcut, buffer=buffer,
        species_types=[_species_code(s, species_map) for s in names],
        alpha_max=alpha_max,
        l_max=int(values["l_max"][0]),
        atom_sigma_r=per_species("atom_sigma_r"),
        atom_sigma_t=per_species("atom_sigma_t"),
        atom_sigma_r_scaling=per_species("atom_sigma_r_scaling", 0.0),
        atom_sigma_t_scaling=per_species("atom_sigma_t_scaling", 0.0),
        amplitude_scaling=per_species("amplitude_scaling", 1.0),
        central_weight=per_species("central_weight", 1.0),
        global_scaling=per_species("global_scaling", 1.0),
        nf=per_species("nf", 4.0),
        basis=basis,
        radial_enhancement=re,
        scaling_mode="polynomial",
        central_species=int(values.get("central_species", ["0"])[0]),
        zeta=float(values.get("zeta", ["2.0"])[0]),
        delta=float(values.get("delta", ["1.0"])[0]),
    )
    return h, list(names)


def read_gap_file(path: str, species_map: Optional[Dict[str, int]] = None) -> List[GapBlock]:
    """Every block of a .gap file, in file order."""
    blocks: List[GapBlock] = []
    cur: Optional[GapBlock] = None
    with open(path) as f:
        for raw in f:
            toks = _tokens(raw)
            if not toks:
                continue
            key = toks[0]
            if key == "gap_beg":
                if cur is not None:
                    raise ValueError(f"{path}: gap_beg inside a {cur.kind} block")
                cur = GapBlock(kind=toks[1] if len(toks) > 1 else "", values={})
            elif key == "gap_end":
                if cur is None:
                    raise ValueError(f"{path}: gap_end outside a block")
                if cur.kind == "soap_turbo":
                    cur.hypers, cur.species_names = _soap_hypers(cur.values, species_map)
                    cur.file_desc = cur.values.get("desc_sparse", [None])[0]
                    cur.file_alphas = cur.values.get("alphas_sparse", [None])[0]
                    cur.compress = _logical(cur.values["compress_soap"][0]) if "compress_soap" in cur.values else False
                blocks.append(cur)
                cur = None
            elif cur is not None:
                cur.values[key] = toks[1:]
    if cur is not None:
        raise ValueError(f"{path}: block {cur.kind} not closed")
    return blocks


def load_soap_blocks(path: str, species_map: Optional[Dict[str, int]] = None,
                     base_dir: Optional[str] = None) -> List[Tuple[SoapTurbo, np.ndarray, np.ndarray]]:
    """(hypers, Qs (n_sparse, dim), alphas) of every soap_turbo block. Sparse-set
    paths are relative to base_dir (default: the working directory, as for
    turbogap)."""
    base = os.getcwd() if base_dir is None else base_dir
    out = []
    for b in read_gap_file(path, species_map):
        if b.kind != "soap_turbo":
            continue
        hyp = compressed_hypers(b, base) if b.compress else b.hypers
        if not b.file_desc or not b.file_alphas:
            raise ValueError("soap_turbo block without desc_sparse / alphas_sparse")
        alphas, Qs = read_alphas_and_descriptors(os.path.join(base, b.file_desc), os.path.join(base, b.file_alphas))
        if Qs.shape[1] != hyp.dim:
            raise ValueError(f"{b.file_desc}: descriptor size {Qs.shape[1]} does not match the hypers ({hyp.dim})")
        warnings.warn(f"{b.file_desc}: soap_turbo sparse sets are evaluated with this library's SOAP descriptor, "
                      "which does not reproduce soap_turbo's bit for bit; energies differ from soap_turbo's unless the "
                      "sparse set was made with this library", SparseSetWarning, stacklevel=2)
        out.append((hyp, Qs, alphas))
    return out


def compressed_hypers(b: GapBlock, base_dir: str) -> SoapTurbo:
    """The hypers of a compress_soap block with its matrix: from
    file_compress_soap when given (it takes precedence), else from
    compress_mode (src/read_files.f90:5166-5226)."""
    fname = b.values.get("file_compress_soap", ["none"])[0]
    mode = b.values.get("compress_mode", ["none"])[0]
    h = b.hypers
    if fname != "none":
        dim, P_i, P_j, P_el, _, _ = read_compress_file(os.path.join(base_dir, fname), h.n_species)
    elif mode == "trivial":
        dim, P_i, P_j, P_el = compress_trivial(h.alpha_max, h.l_max)
    elif mode != "none":
        raise NotImplementedError(f'compress_mode "{mode}" is not implemented (file_compress_soap, or "trivial")')
    else:
        raise ValueError("compress_soap = .true. needs file_compress_soap or compress_mode")
    return h.with_compression(dim, P_i, P_j, P_el)


def load_pair_blocks(path: str, species_map: Optional[Dict[str, int]] = None, base_dir: Optional[str] = None,
                     core_pot_cutoff: float = 1.0e10, core_pot_buffer: float = 1.0) -> List[Tuple[str, dict]]:
    """[(kind, kwargs)] of the distance_2b, angle_3b and core_pot blocks, in file order."""
    base = os.getcwd() if base_dir is None else base_dir
    out = []
    for b in read_gap_file(path, species_map):
        v = b.values
        if b.kind == "distance_2b":
            arec = [ln.split() for ln in open(os.path.join(base, v["alphas_sparse"][0])).read().splitlines()]
            drec = [ln.split() for ln in open(os.path.join(base, v["desc_sparse"][0])).read().splitlines()]
            if len(drec) != len(arec):
                raise ValueError("distance_2b: descriptor and alphas files differ in length")
            f = lambda t: float(t.replace("d", "e").replace("D", "e"))
            out.append(("distance_2b", dict(
                species1=_species_code(v["species1"][0], species_map),
                species2=_species_code(v["species2"][0], species_map),
                # the driver passes buffer 0.5 and e0 = 0 (src/gap_backend_cpu.f90:79-139); delta and sigma
                # default to 1 (src/types.f90:137-151)
                rcut=f(v["rcut"][0]), buffer=0.5, delta=f(v.get("delta", ["1.0"])[0]),
                sigma=f(v.get("sigma", ["1.0"])[0]), e0=0.0,
                Qs=np.array([f(r[0]) for r in drec]), alphas=np.array([f(r[0]) for r in arec]),
                cutoff=np.array([f(r[1]) for r in arec]))))
        elif b.kind == "angle_3b":
            f = lambda t: float(t.replace("d", "e").replace("D", "e"))
            arec = [ln.split() for ln in open(os.path.join(base, v["alphas_sparse"][0])).read().splitlines() if ln.strip()]
            desc = [f(ln.split()[0]) for ln in open(os.path.join(base, v["desc_sparse"][0])).read().splitlines()
                    if ln.strip()]
            if len(desc) != 3 * len(arec):
                raise ValueError("angle_3b: the descriptor file must hold three values per sparse point")
            sigma = [f(t) for t in v["sigma"][:3]] if "sigma" in v else [1.0, 1.0, 1.0]
            if len(sigma) != 3:
                raise ValueError("angle_3b: sigma needs three values")

            def code(*keys):
                for key in keys:
                    if key in v:
                        return _species_code(v[key][0], species_map)
                raise ValueError(f"angle_3b block without {keys[-1]}")
            out.append(("angle_3b", dict(
                species_center=code("Z_center", "z_center", "species_center"),
                species1=code("Z1", "z1", "species1"), species2=code("Z2", "z2", "species2"),
                # the driver passes buffer 0.5 and e0 = 0 (src/gap_backend_cpu.f90:250-258)
                rcut=f(v["rcut"][0]), buffer=0.5, delta=f(v.get("delta", ["1.0"])[0]), sigma=sigma, e0=0.0,
                kernel_type=v.get("kernel_type", ["exp"])[0],
                Qs=np.array(desc).reshape(-1, 3), alphas=np.array([f(r[0]) for r in arec]),
                cutoff=np.array([f(r[1]) for r in arec]))))
        elif b.kind == "core_pot":
            lines = open(os.path.join(base, v["core_pot_file"][0])).read().split("\n")
            head = lines[0].split()
            n, yp1, ypn = int(head[0]), float(head[1]), float(head[2])
            tab = np.array([[float(t) for t in ln.split()[:2]] for ln in lines[1:1 + n]])
            out.append(("core_pot", dict(
                species1=_species_code(v["species1"][0], species_map),
                species2=_species_code(v["species2"][0], species_map),
                x=tab[:, 0], V=tab[:, 1], yp1=yp1, ypn=ypn, cutoff=core_pot_cutoff, buffer=core_pot_buffer)))
    return out
