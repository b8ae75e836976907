"""CPU restatement of local_property_predict (src/local_properties.f90:174-297), numpy FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
tgx_local_property_predict; the product path never imports it.

K = soap . Qs (dgemm 't','n'); with an integer zeta (|nint(zeta) - zeta| < 1e-10)
K^zeta by integer power; V = delta^2 K^zeta alphas + V0, clamped at 0;
Qss = delta^2 (zeta K^(zeta-1) o alphas) Qs^T; V_der(:, p) = Qss(i, :) .
soap_cart_der(:, :, p) for the pairs p of centre i, zero where V_i = 0.
"""
from __future__ import annotations

import numpy as np


def local_property_predict(soap, Qs, alphas, V0, delta, zeta, soap_cart_der=None, n_neigh=None):
    """soap (n_sites, D), Qs (n_sparse, D), soap_cart_der (P, D, 3) in the
    reference's pair order, n_neigh (n_sites,). Returns V (n_sites,) and
    V_der (P, 3) (None without derivatives)."""
    soap, Qs, alphas = (np.asarray(a, dtype=np.float64) for a in (soap, Qs, alphas))
    K = soap @ Qs.T
    zi = int(round(zeta))
    z = zi if abs(float(zi) - zeta) < 1e-10 else zeta
    V = delta ** 2 * (K ** z) @ alphas + V0
    V = np.where(V < 0.0, 0.0, V)
    if soap_cart_der is None:
        return V, None
    Qss = delta ** 2 * ((z * K ** (z - 1)) * alphas[None, :]) @ Qs
    der = np.asarray(soap_cart_der, dtype=np.float64)
    Vd = np.zeros((der.shape[0], 3))
    j = 0
    for i, nn in enumerate(np.asarray(n_neigh)):
        for _ in range(nn):
            if V[i] != 0.0:
                Vd[j] = Qss[i] @ der[j]
            j += 1
    return V, Vd
