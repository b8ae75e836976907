// Host-side pieces of libtgx: error state, synthetic-input generators and the
// radial quadrature tables the SOAP kernels read.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <vector>
#include <algorithm>

#include "tgx_internal.h"

static thread_local char g_err[1024] = "no error";

void tgx_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char* tgx_last_error(void) { return g_err; }

// ---------------------------------------------------------------------------
// Deterministic RNG: xoshiro256** seeded by splitmix64.
namespace {
struct rng {
  uint64_t s[4];
  explicit rng(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 4; ++i) {
      z += 0x9e3779b97f4a7c15ULL;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
      s[i] = x ^ (x >> 31);
    }
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};
}  // namespace

extern "C" int tgx_generate_rsa(int n, const double* L, double dmin, uint64_t seed, double* pos) {
  TGX_REQUIRE(n >= 0 && L && pos && L[0] > 0 && L[1] > 0 && L[2] > 0 && dmin >= 0, "tgx_generate_rsa: bad arguments");
  rng g(seed);
  int m[3];
  for (int k = 0; k < 3; ++k) {
    m[k] = dmin > 0 ? (int)(L[k] / dmin) : 1;
    if (m[k] < 1) m[k] = 1;
    if (m[k] > 1024) m[k] = 1024;
  }
  std::vector<std::vector<int>> cells((size_t)m[0] * m[1] * m[2]);
  const double d2min = dmin * dmin;
  long tries = 0, max_tries = 2000L * (long)n + 100000;
  for (int i = 0; i < n; ++i) {
    for (;;) {
      if (++tries > max_tries) {
        tgx_set_error("tgx_generate_rsa: packing failed after %ld trials at atom %d", tries, i);
        return TGX_ERR_ARG;
      }
      double p[3] = {g.uniform() * L[0], g.uniform() * L[1], g.uniform() * L[2]};
      int c[3];
      for (int k = 0; k < 3; ++k) { c[k] = (int)(p[k] / L[k] * m[k]); if (c[k] >= m[k]) c[k] = m[k] - 1; }
      bool ok = true;
      for (int dz = -1; dz <= 1 && ok; ++dz)
        for (int dy = -1; dy <= 1 && ok; ++dy)
          for (int dx = -1; dx <= 1 && ok; ++dx) {
            int cx = ((c[0] + dx) % m[0] + m[0]) % m[0], cy = ((c[1] + dy) % m[1] + m[1]) % m[1],
                cz = ((c[2] + dz) % m[2] + m[2]) % m[2];
            // with fewer than 3 cells in a dimension the loop revisits cells; harmless
            for (int j : cells[((size_t)cz * m[1] + cy) * m[0] + cx]) {
              double d2 = 0.0;
              for (int k = 0; k < 3; ++k) {
                double d = pos[3 * (size_t)j + k] - p[k];
                d -= L[k] * std::round(d / L[k]);
                d2 += d * d;
              }
              if (d2 < d2min) { ok = false; break; }
            }
          }
      if (!ok) continue;
      for (int k = 0; k < 3; ++k) pos[3 * (size_t)i + k] = p[k];
      cells[((size_t)c[2] * m[1] + c[1]) * m[0] + c[0]].push_back(i);
      break;
    }
  }
  return TGX_OK;
}

extern "C" int tgx_generate_permutation(int n, uint64_t seed, int* perm) {
  TGX_REQUIRE(n >= 0 && perm, "tgx_generate_permutation: bad arguments");
  rng g(seed);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int i = n - 1; i > 0; --i) {
    int j = (int)(g.next() % (uint64_t)(i + 1));
    std::swap(perm[i], perm[j]);
  }
  return TGX_OK;
}

extern "C" int tgx_generate_normal(int n, double stddev, uint64_t seed, double* out) {
  TGX_REQUIRE(n >= 0 && out, "tgx_generate_normal: bad arguments");
  rng g(seed);
  for (int i = 0; i < n; ++i) out[i] = stddev * g.normal();
  return TGX_OK;
}

// ---------------------------------------------------------------------------
// Gauss-Legendre nodes and weights on [-1, 1] (Newton iteration on P_n).
void tgx_gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; ++j) {
        double pm = p1;
        p1 = p0;
        p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * pm) / j;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = 0.0;
    for (int j = 1; j <= n; ++j) {
      double pm = p1;
      p1 = p0;
      p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * pm) / j;
    }
    dp = n * (z * p0 - p1) / (z * z - 1.0);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// Radial tables of every species of a model:
//   xq[s][q]       quadrature nodes in x = r / rcut_hard
//   B[s][n][q]     w_q psi_n(x_q) F(x_q) x_q^re
//   cen[s][n]      central-atom coefficient of l = 0 (central_weight,
//                  global_scaling, 1/sigma and Y00 folded in)
// psi_n is the ordered Gram-Schmidt basis of (1-x)^3, (1-x)^4, ... ("poly3"),
// preceded by exp(-x^2/(2 s0^2)) for "poly3gauss". The polynomial part comes
// from a Stieltjes recurrence in t = 1 - x with weight t^6 on the quadrature
// nodes, which is exact for polynomials and stays well conditioned.
int tgx_build_radial_tables(tgx_model_t* m, std::vector<double>& B, std::vector<double>& xq_all,
                            std::vector<double>& cen_all) {
  const int S = m->n_species;
  B.assign((size_t)TGX_MAX_SPECIES * TGX_MAX_NB * TGX_MAX_Q, 0.0);
  xq_all.assign((size_t)TGX_MAX_SPECIES * TGX_MAX_Q, 0.0);
  cen_all.assign((size_t)TGX_MAX_SPECIES * TGX_MAX_NB, 0.0);
  const double Y00 = 0.5 / std::sqrt(M_PI);
  for (int s = 0; s < S; ++s) {
    const double rc = m->rcut_hard[s];
    const double xs = m->rcut_soft[s] / rc;
    const double s0 = m->sig_r[s] / rc;
    const bool buffer = xs < 1.0;
    const int n1 = buffer ? m->nq_soft : m->nq_soft + m->nq_buffer;
    const int n2 = buffer ? m->nq_buffer : 0;
    const double x1 = buffer ? xs : 1.0;
    std::vector<double> xa, wa, x, w;
    tgx_gauss_legendre(n1, xa, wa);
    for (int i = 0; i < n1; ++i) { x.push_back(0.5 * (xa[i] + 1.0) * x1); w.push_back(0.5 * wa[i] * x1); }
    if (n2 > 0) {
      tgx_gauss_legendre(n2, xa, wa);
      for (int i = 0; i < n2; ++i) {
        x.push_back(xs + 0.5 * (xa[i] + 1.0) * (1.0 - xs));
        w.push_back(0.5 * wa[i] * (1.0 - xs));
      }
    }
    const int nq = (int)x.size();
    m->P.sp[s].nq = nq;
    auto dotw = [&](const std::vector<double>& f, const std::vector<double>& g) {
      double acc = 0.0;
      for (int q = 0; q < nq; ++q) acc += w[q] * f[q] * g[q];
      return acc;
    };
    const int nb = m->alpha_max[s];
    const int n_poly = m->basis == 1 ? nb - 1 : nb;
    std::vector<std::vector<double>> poly;
    {
      std::vector<double> cur(nq);
      for (int q = 0; q < nq; ++q) { double t = 1.0 - x[q]; cur[q] = t * t * t; }
      double nrm = std::sqrt(dotw(cur, cur));
      for (double& v : cur) v /= nrm;
      for (int k = 0; k < n_poly; ++k) {
        poly.push_back(cur);
        if (k + 1 == n_poly) break;
        std::vector<double> v(nq);
        for (int q = 0; q < nq; ++q) v[q] = (1.0 - x[q]) * cur[q];
        for (int pass = 0; pass < 2; ++pass)
          for (const auto& pj : poly) {
            double c = dotw(v, pj);
            for (int q = 0; q < nq; ++q) v[q] -= c * pj[q];
          }
        double nv = std::sqrt(dotw(v, v));
        for (int q = 0; q < nq; ++q) cur[q] = v[q] / nv;
      }
    }
    std::vector<std::vector<double>> basis;
    if (m->basis == 1) {
      std::vector<double> g(nq);
      for (int q = 0; q < nq; ++q) g[q] = std::exp(-x[q] * x[q] / (2.0 * s0 * s0));
      double ng = std::sqrt(dotw(g, g));
      for (double& v : g) v /= ng;
      basis.push_back(g);
      for (int k = 0; k < n_poly; ++k) {
        std::vector<double> v = poly[k];
        for (int pass = 0; pass < 2; ++pass)
          for (const auto& bj : basis) {
            double c = dotw(v, bj);
            for (int q = 0; q < nq; ++q) v[q] -= c * bj[q];
          }
        double nv = std::sqrt(dotw(v, v));
        for (double& t : v) t /= nv;
        basis.push_back(v);
      }
    } else {
      basis = poly;
    }
    for (int q = 0; q < nq; ++q) xq_all[(size_t)s * TGX_MAX_Q + q] = x[q];
    for (int n = 0; n < nb; ++n) {
      double c = 0.0;
      for (int q = 0; q < nq; ++q) {
        const double xx = x[q];
        const double F = xx <= xs ? 1.0 : std::exp(-m->nf[s] * (xx - xs) * (xx - xs) / ((1.0 - xs) * (1.0 - xs)));
        double en = 1.0;
        for (int e = 0; e < m->radial_enhancement; ++e) en *= xx;
        const double b = w[q] * basis[n][q] * F * en;
        B[((size_t)s * TGX_MAX_NB + n) * TGX_MAX_Q + q] = b;
        c += b * std::exp(-xx * xx / (2.0 * s0 * s0));
      }
      cen_all[(size_t)s * TGX_MAX_NB + n] = m->cw[s] * m->global_scaling[s] * Y00 * c / s0;
    }
  }
  return TGX_OK;
}
