"""CPU restatement of TurboGAP's pair-potential GAP families (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
distance_2b and core_pot kernels; the product path never imports it.

  * distance_2b_energy_and_forces follows get_2b_energy_and_forces
    (src/gap.f90:400-563): centres of species s1 or s2 with a neighbour of the
    other, r < rcut, fcut = (cos(pi (r - rcut + buffer)/buffer) + 1)/2 over the
    last `buffer`, E_i = e0 + delta^2 sum alpha_s cutoff_s fcut
    exp(-(r - Qs_s)^2 / (2 sigma^2)), the pair force on the centre and the
    half-symmetrised virial times 1/2.
  * core_pot_model follows src/read_files.f90:5424-5510 (cutoff damping of the
    table, the tridiagonal sweep for the spline's second derivatives, natural
    ends' slopes from the spline); spline / spline_der follow
    src/splines.f90:35-190; core_pot_energy_and_forces follows
    get_core_pot_energy_and_forces (src/gap.f90:565-721), E_i = 1/2 sum V(r).
    The slope of a natural upper end is taken on the last interval (the
    reference's search for the interval of r = x(n) runs past the table).
The loops run in the reference's order (centres, then their pairs).
"""
from __future__ import annotations

import math

import numpy as np


def _pairs(nl, species_sc, s1, s2):
    """(i, k) of the selected pairs in the reference's loop order; nl holds the
    reference's arrays (n_neigh, neighbor_species, rjs, xyz)."""
    n_neigh = np.asarray(nl["n_neigh"])
    nspec = np.asarray(nl["neighbor_species"])
    k = 0
    for i in range(n_neigh.shape[0]):
        si = species_sc[i]
        if si != s1 and si != s2:
            k += n_neigh[i]
            continue
        k += 1
        for _ in range(1, n_neigh[i]):
            if (si == s1 and nspec[k] == s2) or (si == s2 and nspec[k] == s1):
                yield i, k
            k += 1


def distance_2b_energy_and_forces(nl, species_sc, n_sites, s1, s2, rcut, buffer, delta, sigma, e0, Qs, alphas,
                                  cutoff):
    Qs, alphas, cutoff = (np.asarray(a, dtype=np.float64) for a in (Qs, alphas, cutoff))
    rjs, xyz = np.asarray(nl["rjs"]), np.asarray(nl["xyz"]).reshape(-1, 3)
    pi = math.acos(-1.0)
    E = np.full(n_sites, float(e0))
    F = np.zeros((n_sites, 3))
    W = np.zeros((3, 3))
    for i, k in _pairs(nl, species_sc, s1, s2):
        r = rjs[k]
        if not r < rcut:
            continue
        if r < rcut - buffer:
            fc, dfc = 1.0, 0.0
        else:
            fc = (math.cos(pi * (r - rcut + buffer) / buffer) + 1.0) / 2.0
            dfc = pi / 2.0 / buffer * math.sin(pi * (r - rcut + buffer) / buffer)
        ex = np.exp(-0.5 * (r - Qs) ** 2 / sigma ** 2)
        E[i] += np.sum(delta ** 2 * alphas * cutoff * fc * ex)
        for s in range(Qs.shape[0]):
            f = -2.0 * delta ** 2 * alphas[s] * cutoff[s] * ex[s] * xyz[k] / r * ((r - Qs[s]) / sigma ** 2 * fc + dfc)
            F[i] += f
            W -= 0.5 * (np.outer(f, xyz[k]) + np.outer(xyz[k], f))
    return E, F, 0.5 * W


def core_pot_model(x, V, yp1, ypn, cutoff=1.0e10, buffer=1.0):
    """(x, V, d2V/dx2, yp1, ypn) of a core_pot table after the cutoff damping."""
    pi = math.acos(-1.0)
    xs, vs = [], []
    for xi, vi in zip(x, V):
        if xi <= cutoff:
            xs.append(float(xi))
            vs.append(float(vi) if xi <= cutoff - buffer else
                      float(vi) * 0.5 * (math.cos(pi / buffer * (xi - cutoff + buffer)) + 1.0))
    x, V = np.array(xs), np.array(vs)
    n = x.shape[0]
    y2, u = np.zeros(n), np.zeros(n)
    if yp1 > 0.99e30:
        y2[0], u[0] = 0.0, 0.0
    else:
        y2[0] = -0.5
        u[0] = (3.0 / (x[1] - x[0])) * ((V[1] - V[0]) / (x[1] - x[0]) - yp1)
    for i in range(1, n - 1):
        sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1])
        p = sig * y2[i - 1] + 2.0
        y2[i] = (sig - 1.0) / p
        u[i] = (V[i + 1] - V[i]) / (x[i + 1] - x[i]) - (V[i] - V[i - 1]) / (x[i] - x[i - 1])
        u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p
    if ypn > 0.99e30:
        qn = un = 0.0
    else:
        qn = 0.5
        un = (3.0 / (x[n - 1] - x[n - 2])) * (ypn - (V[n - 1] - V[n - 2]) / (x[n - 1] - x[n - 2]))
    y2[n - 1] = (un - qn * u[n - 2]) / (qn * y2[n - 2] + 1.0)
    for i in range(n - 2, -1, -1):
        y2[i] = y2[i] * y2[i + 1] + u[i]
    if yp1 > 0.99e30:
        yp1 = _interval_der(x, V, y2, 0, x[0])
    if ypn > 0.99e30:
        ypn = _interval_der(x, V, y2, n - 2, x[n - 1])
    return x, V, y2, yp1, ypn


def _interval(x, r):
    j = 0
    while j < x.shape[0] - 2 and not r < x[j + 1]:
        j += 1
    return j


def _interval_der(x, y, y2, j, r):
    h = x[j + 1] - x[j]
    A = (x[j + 1] - r) / h
    B = 1.0 - A
    return (-1.0 / h) * y[j] + (1.0 / h) * y[j + 1] + (1.0 - 3.0 * A ** 2) * (h / 6.0) * y2[j] + \
        (3.0 * B ** 2 - 1.0) * (h / 6.0) * y2[j + 1]


def spline(model, r):
    x, y, y2, yp1, ypn = model
    if r < x[0]:
        return y[0] + (r - x[0]) * yp1
    if r > x[-1]:
        return y[-1] + (r - x[-1]) * ypn
    j = _interval(x, r)
    h = x[j + 1] - x[j]
    A = (x[j + 1] - r) / h
    B = 1.0 - A
    h26 = h ** 2 / 6.0
    return A * y[j] + B * y[j + 1] + (A ** 3 - A) * h26 * y2[j] + (B ** 3 - B) * h26 * y2[j + 1]


def spline_der(model, r):
    x, y, y2, yp1, ypn = model
    if r < x[0]:
        return yp1
    if r > x[-1]:
        return ypn
    return _interval_der(x, y, y2, _interval(x, r), r)


def core_pot_energy_and_forces(nl, species_sc, n_sites, s1, s2, model):
    rjs, xyz = np.asarray(nl["rjs"]), np.asarray(nl["xyz"]).reshape(-1, 3)
    rcut = float(np.max(model[0]))
    E = np.zeros(n_sites)
    F = np.zeros((n_sites, 3))
    W = np.zeros((3, 3))
    for i, k in _pairs(nl, species_sc, s1, s2):
        r = rjs[k]
        if not r < rcut:
            continue
        E[i] += 0.5 * spline(model, r)
        f = spline_der(model, r) * xyz[k] / r
        F[i] += f
        W -= 0.5 * (np.outer(f, xyz[k]) + np.outer(xyz[k], f))
    return E, F, 0.5 * W
