"""Sparse-set files of a soap_turbo block (host side).

read_alphas_and_descriptors follows src/read_files.f90:660-724 for
descriptor_type "soap_turbo": n_sparse is the number of records in the alphas
file, dim = (records in the descriptor file) // n_sparse, and the descriptor
file holds Qs(dim, n_sparse) one value per record with the component index
fastest. Records are counted the way Fortran list-directed reads count them
(every line is a record, blank lines included). write_alphas_and_descriptors
writes the same layout (17 significant digits, so a read returns the values
bit for bit).
"""
from __future__ import annotations

from typing import Tuple

import numpy as np


def _records(path: str):
    with open(path) as f:
        return f.read().splitlines()


def read_alphas_and_descriptors(file_desc: str, file_alphas: str) -> Tuple[np.ndarray, np.ndarray]:
    """Returns (alphas (n_sparse,), Qs (n_sparse, dim)); Qs row s is the
    reference's Qs(:, s)."""
    arec = _records(file_alphas)
    drec = _records(file_desc)
    n_sparse = len(arec)
    if n_sparse == 0:
        raise ValueError(f"{file_alphas}: no alphas")
    dim = len(drec) // n_sparse
    if dim == 0:
        raise ValueError(f"{file_desc}: fewer descriptor records than alphas")
    alphas = np.array([float(r.split()[0]) for r in arec])
    vals = np.array([float(r.split()[0]) for r in drec[: n_sparse * dim]])
    return alphas, vals.reshape(n_sparse, dim)


def write_alphas_and_descriptors(file_desc: str, file_alphas: str, alphas, Qs) -> None:
    alphas = np.asarray(alphas, dtype=np.float64)
    Qs = np.asarray(Qs, dtype=np.float64)
    if Qs.ndim != 2 or Qs.shape[0] != alphas.shape[0]:
        raise ValueError("Qs must be (n_sparse, dim) with one alpha per row")
    np.savetxt(file_alphas, alphas, fmt="%.17g")
    np.savetxt(file_desc, Qs.reshape(-1), fmt="%.17g")
