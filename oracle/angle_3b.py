"""CPU restatement of TurboGAP's three-body GAP family, angle_3b (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
angle_3b kernel; the product path never imports it.

  * cov_pp / cov_pp_der follow src/gap.f90:1049-1112 (d = 3, q = 1 only:
    cov = (1 - r)^4 (4 r + 1) on 0 < r < 1, 1 at r <= 0, 0 at r >= 1; the
    derivative is divided by r and is 0 outside 0 < r < 1).
  * angle_3b_energy_and_forces follows get_3b_energy_and_forces
    (src/gap.f90:724-1042): centres of species sc; neighbour pairs j < k (list
    order) with species (s1, s2) in either order and r12, r13 < rcut; the
    descriptor q = (r12 + r13, (r12 - r13)^2, r23) and the scaled distance
    r_s = |(q - Qs_s) / sigma| to each sparse point; E_i = e0 + fcut12 fcut13
    sum_s 2 delta^2 alpha_s cutoff_s k(r_s) with k = exp(-r^2/2) ("exp") or
    cov_pp ("pol"); the forces on the centre and on both neighbours (folded to
    their home atoms, mod(index - 1, n_sites) + 1), and the virial
    sum 1/2 (f2 x12^T + x12 f2^T) + 1/2 (f3 x13^T + x13 f3^T).
  * angle_3b_energy_brute is an independent statement of the same energy from
    positions (all periodic images within rcut, every unordered neighbour pair
    at once), without the neighbour list: the check on the list-driven loops.
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def cov_pp(r):
    r = np.asarray(r, dtype=np.float64)
    j = 3  # d/2 + q + 1 with d = 3, q = 1
    out = np.where(r >= 1.0, 0.0, 1.0)
    m = (r > 0.0) & (r < 1.0)
    out[m] = (1.0 - r[m]) ** (j + 1) * ((j + 1.0) * r[m] + 1.0)
    return out


def cov_pp_der(r):
    r = np.asarray(r, dtype=np.float64)
    j = 3
    out = np.zeros_like(r)
    m = (r > 0.0) & (r < 1.0)
    rm = r[m]
    out[m] = (-(j + 1.0) * (1.0 - rm) ** j * ((j + 1.0) * rm + 1.0) + (1.0 - rm) ** (j + 1) * (j + 1.0)) / rm
    return out


def angle_3b_energy_and_forces(nl, species_sc, n_sites, sc, s1, s2, rcut, buffer, delta, sigma, e0, Qs, alphas,
                               cutoff, kernel_type="exp"):
    """nl holds the reference's list arrays (n_neigh, neighbors_list (1-based),
    neighbor_species, rjs, xyz); species codes compare directly."""
    if kernel_type not in ("exp", "pol"):
        raise ValueError(kernel_type)
    Qs = np.asarray(Qs, dtype=np.float64).reshape(-1, 3)
    sigma = np.asarray(sigma, dtype=np.float64)
    pref = 2.0 * delta ** 2 * np.asarray(alphas, dtype=np.float64) * np.asarray(cutoff, dtype=np.float64)
    n_neigh = np.asarray(nl["n_neigh"])
    nlist = np.asarray(nl["neighbors_list"])
    nspec = np.asarray(nl["neighbor_species"])
    rjs, xyz = np.asarray(nl["rjs"]), np.asarray(nl["xyz"]).reshape(-1, 3)
    pi = math.acos(-1.0)
    E = np.full(n_sites, float(e0))
    F = np.zeros((n_sites, 3))
    W = np.zeros((3, 3))

    def cut(r):
        if r < rcut - buffer:
            return 1.0, 0.0
        a = pi * (r - rcut + buffer) / buffer
        return (math.cos(a) + 1.0) / 2.0, math.sin(a) / 2.0 * pi / buffer

    k = 0
    for i in range(n_neigh.shape[0]):
        if species_sc[i] != sc:
            k += n_neigh[i]
            continue
        i3 = (nlist[k] - 1) % n_sites
        k += 1
        for j in range(1, n_neigh[i]):
            r12 = rjs[k]
            if (nspec[k] != s1 and nspec[k] != s2) or r12 > rcut:
                k += 1
                continue
            j3 = (nlist[k] - 1) % n_sites
            x12 = xyz[k]
            u12 = x12 / r12
            for k2 in range(k + 1, k + n_neigh[i] - j):
                r13 = rjs[k2]
                pair_ok = (nspec[k] == s1 and nspec[k2] == s2) or (nspec[k] == s2 and nspec[k2] == s1)
                if not (pair_ok and r13 < rcut and r12 < rcut):
                    continue
                k3 = (nlist[k2] - 1) % n_sites
                x13 = xyz[k2]
                u13 = x13 / r13
                x23 = x13 - x12
                r23 = math.sqrt(float(np.sum(x23 ** 2)))
                u23 = x23 / r23
                f12, g12 = cut(r12)
                f13, g13 = cut(r13)
                dfc12, dfc13 = g12 * u12, g13 * u13
                fcut = f12 * f13
                dfcut = f12 * dfc13 + f13 * dfc12
                q = np.array([r12 + r13, (r12 - r13) ** 2, r23])
                rr = np.sqrt(np.sum((q[None, :] - Qs) ** 2 / sigma[None, :] ** 2, axis=1))
                drdq = (q[None, :] - Qs) / sigma[None, :] ** 2
                # the reference's force sums read -sum(kernel*(dfcut + fcut*drdx)) for "exp" and
                # -sum(kernel*dfcut - kernel_der*fcut*drdx) for "pol": kd is the factor of fcut*drdx
                if kernel_type == "exp":
                    kern = pref * np.exp(-0.5 * rr ** 2)
                    kd = kern
                else:
                    kern = pref * cov_pp(rr)
                    kd = -pref * cov_pp_der(rr)
                E[i] += fcut * np.sum(kern)
                f1, f2, f3 = np.zeros(3), np.zeros(3), np.zeros(3)
                for l in range(3):
                    d1 = drdq[:, 0] * (u12[l] + u13[l]) + drdq[:, 1] * 2.0 * (r12 - r13) * (u12[l] - u13[l])
                    d2 = -drdq[:, 0] * u12[l] - drdq[:, 1] * 2.0 * (r12 - r13) * u12[l] + drdq[:, 2] * u23[l]
                    d3 = -drdq[:, 0] * u13[l] - drdq[:, 1] * 2.0 * (r13 - r12) * u13[l] - drdq[:, 2] * u23[l]
                    f1[l] = -np.sum(kern * dfcut[l] + kd * fcut * d1)
                    f2[l] = -np.sum(-kern * f13 * dfc12[l] + kd * fcut * d2)
                    f3[l] = -np.sum(-kern * f12 * dfc13[l] + kd * fcut * d3)
                F[i3] += f1
                F[j3] += f2
                F[k3] += f3
                W += 0.5 * (np.outer(f2, x12) + np.outer(x12, f2))
                W += 0.5 * (np.outer(f3, x13) + np.outer(x13, f3))
            k += 1
    return E, F, W


def angle_3b_energy_brute(positions, species, box, sc, s1, s2, rcut, buffer, delta, sigma, e0, Qs, alphas, cutoff,
                          kernel_type="exp"):
    """Per-atom energies from positions: every image of every atom within rcut
    of a centre, all unordered pairs of them at once."""
    pos = np.asarray(positions, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64)
    n = pos.shape[0]
    Qs = np.asarray(Qs, dtype=np.float64).reshape(-1, 3)
    w = 1.0 / np.asarray(sigma, dtype=np.float64) ** 2
    pref = 2.0 * delta ** 2 * np.asarray(alphas) * np.asarray(cutoff)
    reach = [int(math.ceil(rcut / (abs(np.linalg.det(box)) / np.linalg.norm(np.cross(box[(c + 1) % 3], box[(c + 2) % 3])))))
             for c in range(3)]
    shifts = np.array(list(itertools.product(*[range(-m, m + 1) for m in reach]))) @ box
    E = np.full(n, float(e0))
    for i in range(n):
        if species[i] != sc:
            continue
        d = (pos[None, :, :] + shifts[:, None, :] - pos[i]).reshape(-1, 3)
        spc = np.tile(species, shifts.shape[0])
        r = np.linalg.norm(d, axis=1)
        keep = (r > 0.0) & (r < rcut) & ((spc == s1) | (spc == s2))
        d, r, spc = d[keep], r[keep], spc[keep]
        a, b = np.triu_indices(r.shape[0], 1)
        ok = ((spc[a] == s1) & (spc[b] == s2)) | ((spc[a] == s2) & (spc[b] == s1))
        a, b = a[ok], b[ok]
        if a.size == 0:
            continue

        def fc(x):
            return np.where(x < rcut - buffer, 1.0, (np.cos(np.pi * (x - rcut + buffer) / buffer) + 1.0) / 2.0)

        q = np.stack([r[a] + r[b], (r[a] - r[b]) ** 2, np.linalg.norm(d[b] - d[a], axis=1)], axis=1)
        rr2 = np.sum((q[:, None, :] - Qs[None, :, :]) ** 2 * w, axis=2)
        if kernel_type == "exp":
            kern = np.exp(-0.5 * rr2)
        else:
            x = np.sqrt(rr2)
            kern = np.where(x < 1.0, (1.0 - np.minimum(x, 1.0)) ** 4 * (4.0 * x + 1.0), 0.0)
        E[i] += np.sum(fc(r[a]) * fc(r[b]) * (kern @ pref))
    return E
