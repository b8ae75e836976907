"""CPU pins of the oracle's SOAP descriptor (the conventions stated at the top
of oracle/tgx_oracle.cpp; soap_turbo itself is not available to compare with).

Independent checks:
  * real spherical harmonics: orthonormality by sphere quadrature, the
    addition theorem sum_m Y_lm(u) Y_lm(v) = (2l+1)/(4 pi) P_l(u.v) with scipy's
    Legendre polynomials, and finite differences of the solid-harmonic gradient;
  * radial basis: the poly3 basis equals an exact rational Gram-Schmidt of
    (1-x)^3, (1-x)^4, ... evaluated at the nodes;
  * closed forms: one neighbour (central weight 0) gives
    p(n,n',l) = w R_n R_n' A_l^2 (2l+1)/(4 pi), and a lone centre gives
    p(n,n',0) = w c_n c_n', with R_n and c_n from scipy quadrature of the
    density definition and the exact basis;
  * invariance under rotations and neighbour permutations, the translation sum
    rule (self pair = -sum of the others), and central finite differences of q.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy import integrate, special

from turbogap_amd.hypers import SoapTurbo


def hyp(basis="poly3", n_max=3, l_max=4, cw=0.0, buffer=0.5, re=0):
    return SoapTurbo.from_buffer(rcut=[3.7], buffer=[buffer], species_types=[6], alpha_max=[n_max], l_max=l_max,
                                 atom_sigma_r=[0.5], atom_sigma_t=[0.5], central_weight=[cw], basis=basis,
                                 radial_enhancement=re)


def exact_poly3_basis(n):
    """Ordered Gram-Schmidt of t^3..t^(n+2) on [0,1] (t = 1 - x), exact rationals.
    Returns coefficient rows c[k][a] for psi_k = sum_a c[k][a] t^(a+3) / norm_k, and norms^2."""
    ip = lambda a, b: sum(Fraction(ca * cb, 1) / (i + j + 7) for i, ca in enumerate(a) for j, cb in enumerate(b))
    rows, nrm2 = [], []
    for k in range(n):
        v = [Fraction(0)] * n
        v[k] = Fraction(1)
        for r, n2 in zip(rows, nrm2):
            c = ip(v, r) / n2
            v = [vi - c * ri for vi, ri in zip(v, r)]
        rows.append(v)
        nrm2.append(ip(v, v))
    return rows, nrm2


def eval_exact_basis(x, n):
    rows, nrm2 = exact_poly3_basis(n)
    t = 1.0 - np.asarray(x)
    out = []
    for r, n2 in zip(rows, nrm2):
        out.append(sum(float(c) * t ** (i + 3) for i, c in enumerate(r)) / np.sqrt(float(n2)))
    return np.array(out)


def test_poly3_basis_matches_exact_gram_schmidt(orc):
    h = hyp("poly3", n_max=6)
    xq, wq, psi, B, cen = orc.radial_table(h, 0)
    np.testing.assert_allclose(psi, eval_exact_basis(xq, 6), rtol=0, atol=1e-9)
    # orthonormal under the node rule
    np.testing.assert_allclose((psi * wq) @ psi.T, np.eye(6), atol=1e-12)


def test_real_harmonics_addition_theorem_and_orthonormality(orc):
    rng = np.random.default_rng(0)
    L = 8
    for _ in range(5):
        u, v = rng.normal(size=3), rng.normal(size=3)
        u /= np.linalg.norm(u)
        v /= np.linalg.norm(v)
        Yu, _ = orc.real_sph_harm(L, u)
        Yv, _ = orc.real_sph_harm(L, v)
        for l in range(L + 1):
            s = Yu[l * l:(l + 1) * (l + 1)] @ Yv[l * l:(l + 1) * (l + 1)]
            assert abs(s - (2 * l + 1) / (4 * np.pi) * special.eval_legendre(l, u @ v)) < 1e-13
    # orthonormality by product Gauss quadrature on the sphere
    ct, wt = np.polynomial.legendre.leggauss(20)
    ph = np.arange(40) * 2 * np.pi / 40
    G = np.zeros(((L + 1) ** 2, (L + 1) ** 2))
    for c, w in zip(ct, wt):
        st = np.sqrt(1 - c * c)
        for p in ph:
            Y, _ = orc.real_sph_harm(L, np.array([st * np.cos(p), st * np.sin(p), c]))
            G += w * (2 * np.pi / 40) * np.outer(Y, Y)
    np.testing.assert_allclose(G, np.eye((L + 1) ** 2), atol=1e-12)


def test_solid_harmonic_gradient_finite_difference(orc):
    rng = np.random.default_rng(1)
    L, hh = 6, 1e-6
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    Y, dY = orc.real_sph_harm(L, u)

    def solid(x):
        r = np.linalg.norm(x)
        Yx, _ = orc.real_sph_harm(L, x / r)
        ls = np.concatenate([[l] * (2 * l + 1) for l in range(L + 1)])
        return r ** ls * Yx

    for k in range(3):
        e = np.zeros(3)
        e[k] = hh
        fd = (solid(u + e) - solid(u - e)) / (2 * hh)
        np.testing.assert_allclose(dY[:, k], fd, atol=1e-8)


def density_radial(h, r, n_list_psi):
    """R_n(r) by adaptive quadrature of the density definition with the exact basis."""
    rc, rs = h.rcut_hard[0], h.rcut_soft[0]
    xs, nf = rs / rc, h.nf[0]
    x0 = r / rc
    sig = h.atom_sigma_r[0] * (1 + h.atom_sigma_r_scaling[0] * r) / rc
    P = 1 + 2 * x0 ** 3 - 3 * x0 ** 2
    amp = h.global_scaling[0] / sig * P ** h.amplitude_scaling[0]
    F = lambda x: 1.0 if x <= xs else np.exp(-nf * (x - xs) ** 2 / (1 - xs) ** 2)
    out = []
    for psi in n_list_psi:
        f = lambda x: psi(x) * np.exp(-(x - x0) ** 2 / (2 * sig * sig)) * F(x) * x ** h.radial_enhancement
        v = integrate.quad(f, 0, xs, epsabs=1e-15, epsrel=1e-13, limit=200)[0]
        v += integrate.quad(f, xs, 1, epsabs=1e-15, epsrel=1e-13, limit=200)[0]
        out.append(amp * v)
    return np.array(out)


def closed_form_q(p):
    return p / np.linalg.norm(p)


def spectrum(c_nl, L, N):
    """p(n,n',l) in the oracle's order from per-(n, l) amplitudes c_nl[n][l]
    whose angular parts are orthonormal (sum_m Y^2 = (2l+1)/(4 pi) folded in)."""
    out = []
    for n in range(N):
        for m in range(n, N):
            w = 1.0 if n == m else np.sqrt(2.0)
            for l in range(L + 1):
                out.append(w * c_nl[n][l] * c_nl[m][l])
    return np.array(out)


@pytest.mark.parametrize("r", [2.1, 3.45])
def test_single_neighbour_closed_form(orc, r):
    h = hyp("poly3", n_max=3, l_max=4, cw=0.0)
    N, L = 3, 4
    psis = [lambda x, k=k: eval_exact_basis(np.array([x]), N)[k][0] for k in range(N)]
    R = density_radial(h, r, psis)
    st = h.atom_sigma_t[0]
    A = np.exp(-np.arange(L + 1) * (np.arange(L + 1) + 1) * st ** 2 / (2 * r * r))
    c_nl = [[R[n] * A[l] * np.sqrt((2 * l + 1) / (4 * np.pi)) for l in range(L + 1)] for n in range(N)]
    q_ref = closed_form_q(spectrum(c_nl, L, N))
    u = np.array([0.3, -0.5, 0.81])
    u /= np.linalg.norm(u)
    q, _ = orc.get_soap(h, [2], [0, 0], [0.0, r], np.array([[0, 0, 0], r * u]), derivatives=False)
    np.testing.assert_allclose(q[0], q_ref, rtol=0, atol=1e-9)


def test_lone_centre_closed_form_poly3gauss(orc):
    h = hyp("poly3gauss", n_max=3, l_max=3, cw=1.0)
    N = 3
    s0 = h.atom_sigma_r[0] / h.rcut_hard[0]
    xs = h.rcut_soft[0] / h.rcut_hard[0]
    g = lambda x: np.exp(-x * x / (2 * s0 * s0))
    poly = [lambda x, k=k: eval_exact_basis(np.array([x]), N - 1)[k][0] for k in range(N - 1)]
    funcs = [g] + poly
    ip = lambda f1, f2: integrate.quad(lambda x: f1(x) * f2(x), 0, 1, epsabs=1e-15, epsrel=1e-13, limit=200,
                                       points=[xs])[0]
    G = np.array([[ip(a, b) for b in funcs] for a in funcs])
    Lc = np.linalg.cholesky(G)
    Linv = np.linalg.inv(Lc)
    basis = [lambda x, k=k: sum(Linv[k, j] * funcs[j](x) for j in range(N)) for k in range(N)]
    nf = h.nf[0]
    F = lambda x: 1.0 if x <= xs else np.exp(-nf * (x - xs) ** 2 / (1 - xs) ** 2)
    cn = [integrate.quad(lambda x, b=b: b(x) * g(x) * F(x), 0, 1, points=[xs], epsabs=1e-15, epsrel=1e-13)[0] / s0
          for b in basis]
    c_nl = [[cn[n] * (1.0 if l == 0 else 0.0) for l in range(4)] for n in range(N)]
    q_ref = closed_form_q(spectrum(c_nl, 3, N))
    q, _ = orc.get_soap(h, [1], [0], [0.0], np.zeros((1, 3)), derivatives=False)
    np.testing.assert_allclose(q[0], q_ref, atol=1e-8)


def _env(seed, n=14, rmin=1.3, rmax=3.5):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(rmin, rmax, n)
    xyz = np.vstack([np.zeros(3), d * r[:, None]])
    return xyz, np.concatenate([[0.0], r])


def test_rotation_and_permutation_invariance(orc):
    h = hyp("poly3gauss", n_max=8, l_max=8, cw=1.0)
    xyz, r = _env(3)
    n = len(r)
    q0, _ = orc.get_soap(h, [n], [0] * n, r, xyz, derivatives=False)
    from scipy.spatial.transform import Rotation
    Rm = Rotation.random(random_state=4).as_matrix()
    q1, _ = orc.get_soap(h, [n], [0] * n, r, xyz @ Rm.T, derivatives=False)
    perm = np.concatenate([[0], 1 + np.random.default_rng(5).permutation(n - 1)])
    q2, _ = orc.get_soap(h, [n], [0] * n, r[perm], xyz[perm], derivatives=False)
    np.testing.assert_allclose(q1, q0, atol=1e-12)
    np.testing.assert_allclose(q2, q0, atol=1e-13)
    assert abs(np.linalg.norm(q0) - 1.0) < 1e-14


def test_derivatives_finite_difference_and_sum_rule(orc):
    h = hyp("poly3gauss", n_max=8, l_max=8, cw=1.0)
    xyz, r = _env(7, n=10, rmax=3.3)
    n = len(r)
    q, der = orc.get_soap(h, [n], [0] * n, r, xyz, derivatives=True)
    np.testing.assert_allclose(der[0], -der[1:].sum(0), atol=1e-14)
    hh = 1e-5
    for j in (1, 4, 9):
        for k in range(3):
            xp, xm = xyz.copy(), xyz.copy()
            xp[j, k] += hh
            xm[j, k] -= hh
            qp, _ = orc.get_soap(h, [n], [0] * n, np.linalg.norm(xp, axis=1), xp, derivatives=False)
            qm, _ = orc.get_soap(h, [n], [0] * n, np.linalg.norm(xm, axis=1), xm, derivatives=False)
            np.testing.assert_allclose(der[j, :, k], (qp[0] - qm[0]) / (2 * hh), atol=2e-8)


def test_contribution_vanishes_at_hard_cutoff(orc):
    h = hyp("poly3gauss", n_max=8, l_max=8, cw=1.0)
    xyz, r = _env(9, n=8, rmax=3.0)
    n = len(r)
    q0, _ = orc.get_soap(h, [n], [0] * n, r, xyz, derivatives=False)
    u = np.array([0.0, 0.6, 0.8])
    for eps in (1e-3, 1e-5):
        rr = 3.7 - eps
        q1, d1 = orc.get_soap(h, [n + 1], [0] * (n + 1), np.append(r, rr), np.vstack([xyz, rr * u]))
        assert np.abs(q1 - q0).max() < 50 * eps ** 2
        assert np.abs(d1[-1]).max() < 50 * eps


def test_l_max_zero_is_radial_only(orc):
    h = hyp("poly3", n_max=4, l_max=0, cw=0.0)
    assert orc.soap_dim(h) == 4 * 5 // 2
    xyz, r = _env(11)
    q, _ = orc.get_soap(h, [len(r)], [0] * len(r), r, xyz, derivatives=False)
    Y0 = 0.5 / np.sqrt(np.pi)
    psis = [lambda x, k=k: eval_exact_basis(np.array([x]), 4)[k][0] for k in range(4)]
    cn = sum(density_radial(h, rj, psis) for rj in r[1:]) * Y0
    q_ref = closed_form_q(spectrum([[c] for c in cn], 0, 4) / (Y0 * Y0) * (1 / (4 * np.pi)))
    np.testing.assert_allclose(q[0], q_ref, atol=1e-9)
