"""Writes tests/golden/carbon216_golden.npz: the FP64 oracle's outputs for the
216-atom configuration (seeded structure, 200 sparse points), used as a
regression fixture by tests/test_golden.py (oracle) and tests/test_gpu_golden.py
(GPU). Inputs are not stored: they are regenerated from the same seeds.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from turbogap_amd import structures as S  # noqa: E402


def build():
    sysm = S.carbon216_system()
    h = S.soap_carbon()
    Qs, al = S.sparse_set(h, 200, O.block_descriptors, S.carbon216_system, seed=5001)
    E, en, F, W, nl = O.compute(sysm, [(h, Qs, al)])
    spi = O.species_index(nl["neighbor_species"], h.species_types)
    q, _ = O.get_soap(h, nl["n_neigh"], spi, nl["rjs"], nl["xyz"], derivatives=False)
    return dict(n_neigh=nl["n_neigh"], neighbors_list=nl["neighbors_list"], rjs=nl["rjs"], xyz=nl["xyz"],
                energies=en, forces=F, virial=W, energy=np.array([E]), soap_sites=q[[0, 17, 100, 215]],
                Qs_checksum=np.array([Qs.sum(), al.sum()]))


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "carbon216_golden.npz")
    np.savez_compressed(out, **build())
    print("wrote", out, os.path.getsize(out), "bytes")
