// Host-side (FP64) real solid harmonics and their lowering table.
//
// S_lm(x, y, z) = ynorm[lm] q_l^|m|(z, r^2) {c_m, s_m}(x, y) is the homogeneous
// polynomial of degree l that equals the real spherical harmonic Y_lm on the
// unit sphere (lm = l^2 + l + m), built by the recurrences the kernels use:
//   c_m + i s_m = (x + i y)^m,
//   q_m^m = 1, q_{m+1}^m = (2m+1) z, q_l^m = ((2l-1) z q_{l-1}^m - (l+m-1) r^2 q_{l-2}^m) / (l-m).
//
// Every Cartesian derivative of S_lm is a homogeneous harmonic polynomial of
// degree l-1, hence a combination of the S_{l-1,m'}:
//   dS_lm/dx_c = sum_t coef[lm][c][t] S_{l-1}[idx[lm][c][t]].
// tgx_lowering_table finds these coefficients by least squares on points of the
// unit sphere and checks that the fit is exact; the force kernels use it to get
// the gradients of the harmonics from the harmonics one degree lower.
#pragma once

#include <cmath>
#include <vector>

inline void tgx_solid_harmonics(int L, const double* ynorm, double x, double y, double z, double* S, double* G) {
  const double r2 = x * x + y * y + z * z;
  double c = 1, cx = 0, cy = 0, s = 0, sx = 0, sy = 0;
  for (int m = 0; m <= L; ++m) {
    if (m > 0) {
      const double nc = c * x - s * y, ns = c * y + s * x;
      const double ncx = cx * x + c - sx * y, ncy = cy * x - sy * y - s;
      const double nsx = cx * y + sx * x + s, nsy = cy * y + c + sy * x;
      c = nc; s = ns; cx = ncx; cy = ncy; sx = nsx; sy = nsy;
    }
    double q = 1, qx = 0, qy = 0, qz = 0, p = 0, px = 0, py = 0, pz = 0;
    for (int l = m; l <= L; ++l) {
      if (l == m + 1) {
        p = q; px = qx; py = qy; pz = qz;
        q = (2.0 * m + 1.0) * z;
        qx = 0; qy = 0; qz = 2.0 * m + 1.0;
      } else if (l > m + 1) {
        const double aa = 2.0 * l - 1.0, bb = (double)(l + m - 1), inv = 1.0 / (double)(l - m);
        const double nq = (aa * z * q - bb * r2 * p) * inv;
        const double nqx = (aa * z * qx - bb * (2.0 * x * p + r2 * px)) * inv;
        const double nqy = (aa * z * qy - bb * (2.0 * y * p + r2 * py)) * inv;
        const double nqz = (aa * (q + z * qz) - bb * (2.0 * z * p + r2 * pz)) * inv;
        p = q; px = qx; py = qy; pz = qz;
        q = nq; qx = nqx; qy = nqy; qz = nqz;
      }
      const int base = l * l + l;
      const double N = ynorm[base + m];
      if (m == 0) {
        S[base] = N * q;
        G[3 * base] = N * qx; G[3 * base + 1] = N * qy; G[3 * base + 2] = N * qz;
      } else {
        S[base + m] = N * q * c;
        G[3 * (base + m)] = N * (qx * c + q * cx);
        G[3 * (base + m) + 1] = N * (qy * c + q * cy);
        G[3 * (base + m) + 2] = N * qz * c;
        S[base - m] = N * q * s;
        G[3 * (base - m)] = N * (qx * s + q * sx);
        G[3 * (base - m) + 1] = N * (qy * s + q * sy);
        G[3 * (base - m) + 2] = N * qz * s;
      }
    }
  }
}

// idx / coef: [NLM][3][KT]; unused slots have idx -1. Returns false when a
// gradient needs more than KT terms or the fit is not exact (max_resid).
inline bool tgx_lowering_table(int L, const double* ynorm, int KT, std::vector<int>& idx, std::vector<float>& coef,
                               double* max_resid) {
  const int NLM = (L + 1) * (L + 1);
  idx.assign((size_t)NLM * 3 * KT, -1);
  coef.assign((size_t)NLM * 3 * KT, 0.0f);
  *max_resid = 0.0;
  // fixed points on the unit sphere (golden-angle spiral)
  const int M = 6 * L + 24;
  std::vector<double> S((size_t)M * NLM), G((size_t)M * NLM * 3);
  for (int k = 0; k < M; ++k) {
    const double zz = 1.0 - (2.0 * k + 1.0) / M;
    const double rr = std::sqrt(1.0 - zz * zz), ph = 2.399963229728653 * k;
    tgx_solid_harmonics(L, ynorm, rr * std::cos(ph), rr * std::sin(ph), zz, &S[(size_t)k * NLM], &G[(size_t)k * NLM * 3]);
  }
  for (int l = 1; l <= L; ++l) {
    const int nb = 2 * l - 1, b0 = (l - 1) * (l - 1);
    // normal equations A^T A w = A^T g over the basis S_{l-1, .}
    std::vector<double> AtA((size_t)nb * nb, 0.0);
    for (int k = 0; k < M; ++k)
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) AtA[(size_t)i * nb + j] += S[(size_t)k * NLM + b0 + i] * S[(size_t)k * NLM + b0 + j];
    for (int lm = l * l; lm < (l + 1) * (l + 1); ++lm)
      for (int cc = 0; cc < 3; ++cc) {
        std::vector<double> a(AtA), w((size_t)nb, 0.0);
        for (int k = 0; k < M; ++k)
          for (int i = 0; i < nb; ++i) w[i] += S[(size_t)k * NLM + b0 + i] * G[((size_t)k * NLM + lm) * 3 + cc];
        // Gaussian elimination with partial pivoting
        for (int col = 0; col < nb; ++col) {
          int piv = col;
          for (int r = col + 1; r < nb; ++r)
            if (std::fabs(a[(size_t)r * nb + col]) > std::fabs(a[(size_t)piv * nb + col])) piv = r;
          if (piv != col) {
            for (int j = 0; j < nb; ++j) std::swap(a[(size_t)col * nb + j], a[(size_t)piv * nb + j]);
            std::swap(w[col], w[piv]);
          }
          for (int r = col + 1; r < nb; ++r) {
            const double f = a[(size_t)r * nb + col] / a[(size_t)col * nb + col];
            for (int j = col; j < nb; ++j) a[(size_t)r * nb + j] -= f * a[(size_t)col * nb + j];
            w[r] -= f * w[col];
          }
        }
        for (int r = nb - 1; r >= 0; --r) {
          double t = w[r];
          for (int j = r + 1; j < nb; ++j) t -= a[(size_t)r * nb + j] * w[j];
          w[r] = t / a[(size_t)r * nb + r];
        }
        double wmax = 0.0;
        for (int i = 0; i < nb; ++i) wmax = std::fmax(wmax, std::fabs(w[i]));
        int t = 0;
        for (int i = 0; i < nb; ++i) {
          if (std::fabs(w[i]) <= 1e-10 * wmax) { w[i] = 0.0; continue; }
          if (t == KT) return false;
          idx[((size_t)lm * 3 + cc) * KT + t] = b0 + i;
          coef[((size_t)lm * 3 + cc) * KT + t] = (float)w[i];
          ++t;
        }
        for (int k = 0; k < M; ++k) {
          double v = 0.0;
          for (int i = 0; i < nb; ++i) v += w[i] * S[(size_t)k * NLM + b0 + i];
          *max_resid = std::fmax(*max_resid, std::fabs(v - G[((size_t)k * NLM + lm) * 3 + cc]));
        }
      }
  }
  return *max_resid < 1e-9;
}
