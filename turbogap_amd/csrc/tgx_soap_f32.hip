// FP32 SOAP kernels for the mixed-precision and surrogate paths.
//
// Radial functions: each species' R_n(r) and dR_n/dr (amplitude, filter and
// global scaling included) are tabulated once per model on a uniform grid of
// K nodes on [0, rcut_hard] from the exact FP64 quadrature, and evaluated by
// cubic Hermite interpolation (error ~ h^4 R''''/384, below 1e-9 relative at
// K = 2048); the force uses the derivative of the same interpolant.
//
// Work split: one wavefront per site; each neighbour pair is shared by LP = 3
// lanes that split its real-harmonic recurrence by m (m = part, part+3, ...),
// so a typical site (about 21 pairs) fills the wavefront. Expansion and
// adjoint contractions run with lanes over (l, m) as in the FP64 kernels.
#include "tgx_internal.h"

namespace {

constexpr int WAVES = 4;
constexpr int LP = 3;           // lanes per pair
constexpr int CH = 64 / LP;     // pairs per pass (21)
constexpr int NMAXF = 16;       // stacked radial functions held per lane

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float wsum(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

struct F32Args {
  const int64_t* site_beg;
  const int* pr_spec;
  const double* pr_r;
  const double* pr_xyz;
  const int64_t* pr_gidx;
  const float* rtab;   // [S][K][2][MAX_NB]: R then dR per node
  int K;
};

struct SpecLite {
  float rc, inv_h, h, sig_t, sig_t_sc;
  int nb, off;
};

// Hermite interpolation of R_n and dR_n/dr at r for species s.
template <bool GRAD>
__device__ __forceinline__ void radial_tab(const F32Args& a, const SpecLite& sp, int s, float r, float (&R)[TGX_MAX_NB],
                                           float (&dR)[TGX_MAX_NB]) {
  float t = r * sp.inv_h;
  int k = (int)t;
  if (k > a.K - 2) k = a.K - 2;
  if (k < 0) k = 0;
  const float u = t - (float)k;
  const float u2 = u * u, u3 = u2 * u;
  const float h00 = 2.f * u3 - 3.f * u2 + 1.f, h10 = u3 - 2.f * u2 + u, h01 = -2.f * u3 + 3.f * u2, h11 = u3 - u2;
  const float g00 = 6.f * u2 - 6.f * u, g10 = 3.f * u2 - 4.f * u + 1.f, g01 = -6.f * u2 + 6.f * u, g11 = 3.f * u2 - 2.f * u;
  const float* n0 = a.rtab + ((size_t)s * a.K + k) * 2 * TGX_MAX_NB;
  const float* n1 = n0 + 2 * TGX_MAX_NB;
  const float hh = sp.h;
#pragma unroll
  for (int n = 0; n < TGX_MAX_NB; ++n) {
    if (n < sp.nb) {
      const float r0 = n0[n], d0 = n0[TGX_MAX_NB + n], r1 = n1[n], d1 = n1[TGX_MAX_NB + n];
      R[n] = h00 * r0 + h10 * hh * d0 + h01 * r1 + h11 * hh * d1;
      if (GRAD) dR[n] = (g00 * r0 + g01 * r1) * sp.inv_h + g10 * d0 + g11 * d1;
    } else {
      R[n] = 0.f;
      if (GRAD) dR[n] = 0.f;
    }
  }
}

// Harmonics with m = part, part + LP, ... streamed to f(l, lm, Y, gx, gy, gz).
template <bool GRAD, class F>
__device__ __forceinline__ void ylm_part(int L, int part, const float* ynorm, float ux, float uy, float uz, F&& f) {
  float c = 1, cx = 0, cy = 0, s = 0, sx = 0, sy = 0;
  for (int m = 0; m <= L; ++m) {
    if (m > 0) {
      const float nc = c * ux - s * uy, ns = c * uy + s * ux;
      if (GRAD) {
        const float ncx = cx * ux + c - sx * uy, ncy = cy * ux - sy * uy - s;
        const float nsx = cx * uy + sx * ux + s, nsy = cy * uy + c + sy * ux;
        cx = ncx; cy = ncy; sx = nsx; sy = nsy;
      }
      c = nc;
      s = ns;
    }
    if ((m % LP) != part) continue;
    float q = 1, qx = 0, qy = 0, qz = 0, p = 0, px = 0, py = 0, pz = 0;
    for (int l = m; l <= L; ++l) {
      if (l == m + 1) {
        p = q; px = qx; py = qy; pz = qz;
        q = (float)(2 * m + 1) * uz;
        qx = 0; qy = 0; qz = (float)(2 * m + 1);
      } else if (l > m + 1) {
        const float aa = (float)(2 * l - 1), bb = (float)(l + m - 1), inv = 1.f / (float)(l - m);
        const float nq = (aa * uz * q - bb * p) * inv;
        if (GRAD) {
          const float nqx = (aa * uz * qx - bb * (2.f * ux * p + px)) * inv;
          const float nqy = (aa * uz * qy - bb * (2.f * uy * p + py)) * inv;
          const float nqz = (aa * (q + uz * qz) - bb * (2.f * uz * p + pz)) * inv;
          px = qx; py = qy; pz = qz;
          qx = nqx; qy = nqy; qz = nqz;
        }
        p = q;
        q = nq;
      }
      const int base = l * l + l;
      const float N = ynorm[base + m];
      if (m == 0) {
        f(l, base, N * q, N * qx, N * qy, N * qz);
      } else {
        f(l, base + m, N * q * c, N * (qx * c + q * cx), N * (qy * c + q * cy), N * qz * c);
        f(l, base - m, N * q * s, N * (qx * s + q * sx), N * (qy * s + q * sy), N * qz * s);
      }
    }
  }
}

struct Shared {
  float* ynorm;   // [NLM]
  SpecLite* sp;   // [S]
};

__device__ void load_shared(const SoapParams& P, const F32Args& a, float* base, Shared& sh) {
  sh.ynorm = base;
  sh.sp = reinterpret_cast<SpecLite*>(base + ((P.NLM + 3) & ~3));
  for (int i = threadIdx.x; i < P.NLM; i += blockDim.x) sh.ynorm[i] = (float)P.ynorm[i];
  for (int s = threadIdx.x; s < P.n_species; s += blockDim.x) {
    SpecLite v;
    v.rc = (float)P.sp[s].rc;
    v.h = (float)(P.sp[s].rc / (a.K - 1));
    v.inv_h = (float)((a.K - 1) / P.sp[s].rc);
    v.sig_t = (float)P.sp[s].sig_t;
    v.sig_t_sc = (float)P.sp[s].sig_t_sc;
    v.nb = P.sp[s].nb;
    v.off = P.sp[s].off;
    sh.sp[s] = v;
  }
}

__host__ __device__ inline int shared_floats(int NLM, int S) { return ((NLM + 3) & ~3) + S * (int)(sizeof(SpecLite) / 4); }

// Per-wave LDS (floats): C [N NLM] | H [N NLM] (adjoint) | U
//   U: W [CH NLM] + Rf [CH NMAXF]  /  X [D]  /  Xa, Xb [CH NLM] + R, dR [CH NMAXF]
__host__ __device__ inline int u_floats(const SoapParams& P, bool adj) {
  int u = CH * P.NLM + CH * NMAXF;
  if (P.D > u) u = P.D;
  if (adj) {
    // the pair chunk overlays C and H (both dead by then) and U
    const int pc = 2 * CH * P.NLM + 2 * CH * NMAXF - 2 * P.N * P.NLM;
    if (pc > u) u = pc;
  }
  return u;
}
__host__ __device__ inline int wave_floats(const SoapParams& P, bool adj) {
  return (adj ? 2 : 1) * P.N * P.NLM + u_floats(P, adj);
}

// Damping A_l = exp(-l(l+1) k) and -dA_l/dr... as a table over l.
__device__ __forceinline__ void damping(const SpecLite& sp, float r, float (&A)[TGX_MAX_L + 1],
                                        float (&dA)[TGX_MAX_L + 1]) {
  const float st = sp.sig_t * (1.f + sp.sig_t_sc * r);
  const float dst = sp.sig_t * sp.sig_t_sc;
  const float kk = st * st / (2.f * r * r);
  const float dk = st * st / (r * r * r) - st * dst / (r * r);
  const float e1 = __expf(-kk), e2 = e1 * e1;
  float Al = 1.f, pw = 1.f;
#pragma unroll
  for (int l = 0; l <= TGX_MAX_L; ++l) {
    if (l > 0) { pw *= e2; Al *= pw; }
    A[l] = Al;
    dA[l] = Al * (float)(l * (l + 1)) * dk;
  }
}

// c(n, lm) of the site into C.
__device__ void coefficients(const SoapParams& P, const F32Args& a, const Shared& sh, int64_t b0, int nn, float* C,
                             float* W, float* Rf) {
  const int lane = threadIdx.x & 63;
  const int NLM = P.NLM, N = P.N, L = P.L;
  const int j = lane / LP, part = lane - j * LP;
  float acc[2][NMAXF];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int n = 0; n < NMAXF; ++n) acc[h][n] = 0.f;
  {
    const int s = a.pr_spec[b0];
    if (s >= 0 && lane == 0) {
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) {
        const int nl = n - sh.sp[s].off;
        if (nl >= 0 && nl < sh.sp[s].nb) acc[0][n] = (float)P.cen[s * TGX_MAX_NB + nl];
      }
    }
  }
  for (int k0 = 1; k0 < nn; k0 += CH) {
    const int cnt = min(CH, nn - k0);
    wave_sync();
    if (j < cnt) {
      const int64_t p = b0 + k0 + j;
      int s = a.pr_spec[p];
      const float r = (float)a.pr_r[p];
      if (s >= 0 && !(a.pr_r[p] < (double)P.sp[s].rc)) s = -1;
      float* w = W + j * NLM;
      if (s >= 0) {
        const SpecLite sp = sh.sp[s];
        const float inv = 1.f / r;
        const float ux = (float)a.pr_xyz[3 * p] * inv, uy = (float)a.pr_xyz[3 * p + 1] * inv,
                    uz = (float)a.pr_xyz[3 * p + 2] * inv;
        if (part == 0) {
          float R[TGX_MAX_NB], dR[TGX_MAX_NB];
          radial_tab<false>(a, sp, s, r, R, dR);
          float* rf = Rf + j * NMAXF;
#pragma unroll
          for (int n = 0; n < NMAXF; ++n) rf[n] = 0.f;
#pragma unroll
          for (int n = 0; n < TGX_MAX_NB; ++n)
            if (n < sp.nb) rf[sp.off + n] = R[n];
        }
        float A[TGX_MAX_L + 1], dA[TGX_MAX_L + 1];
        damping(sp, r, A, dA);
        ylm_part<false>(L, part, sh.ynorm, ux, uy, uz, [&](int l, int lm, float Y, float, float, float) { w[lm] = A[l] * Y; });
      } else {
        if (part == 0) {
          float* rf = Rf + j * NMAXF;
#pragma unroll
          for (int n = 0; n < NMAXF; ++n) rf[n] = 0.f;
        }
        for (int lm = part; lm < NLM; lm += LP) w[lm] = 0.f;
      }
    }
    wave_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lm = lane + 64 * h;
      if (lm < NLM)
        for (int jj = 0; jj < cnt; ++jj) {
          const float wv = W[jj * NLM + lm];
          const float* rf = Rf + jj * NMAXF;
#pragma unroll
          for (int n = 0; n < NMAXF; ++n) acc[h][n] += rf[n] * wv;
        }
    }
  }
  wave_sync();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lm = lane + 64 * h;
    if (lm < NLM)
#pragma unroll
      for (int n = 0; n < NMAXF; ++n)
        if (n < N) C[n * NLM + lm] = acc[h][n];
  }
  wave_sync();
}

__device__ float power(const SoapParams& P, const float* C, float* X) {
  const int lane = threadIdx.x & 63;
  const int NLM = P.NLM;
  float acc = 0.f;
  for (int d = lane; d < P.D; d += 64) {
    const int code = P.dtab[d];
    const int n = code & 0xff, np = (code >> 8) & 0xff, l = code >> 16;
    const float* c1 = C + n * NLM + l * l;
    const float* c2 = C + np * NLM + l * l;
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int m = 0; m < 2 * TGX_MAX_L + 1; m += 2) {
      if (m < 2 * l + 1) s0 += c1[m] * c2[m];
      if (m + 1 < 2 * l + 1) s1 += c1[m + 1] * c2[m + 1];
    }
    float s = s0 + s1;
    if (n != np) s *= (float)M_SQRT2;
    X[d] = s;
    acc += s * s;
  }
  const float nrm = sqrtf(wsum(acc));
  const float inv = nrm > 0.f ? 1.f / nrm : 0.f;
  for (int d = lane; d < P.D; d += 64) X[d] *= inv;
  wave_sync();
  return nrm;
}

__global__ __launch_bounds__(64 * WAVES) void k_desc_f32(SoapParams P, F32Args a, int n_sites, float* __restrict__ q_out,
                                                         float* __restrict__ pnorm_out) {
  extern __shared__ __align__(16) float smf[];
  Shared sh;
  load_shared(P, a, smf, sh);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int site = blockIdx.x * WAVES + w;
  if (site >= n_sites) return;
  float* C = smf + shared_floats(P.NLM, P.n_species) + w * wave_floats(P, false);
  float* U = C + P.N * P.NLM;
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  coefficients(P, a, sh, b0, nn, C, U, U + CH * P.NLM);
  const float nrm = power(P, C, U);
  float* qrow = q_out + (size_t)site * P.Dpad;
  for (int d = lane; d < P.Dpad; d += 64) qrow[d] = d < P.D ? U[d] : 0.f;
  if (lane == 0) pnorm_out[site] = nrm;
}

__global__ __launch_bounds__(64 * WAVES) void k_adjoint_f32(SoapParams P, F32Args a, int n_sites,
                                                            const float* __restrict__ q_in,
                                                            const float* __restrict__ pnorm_in,
                                                            const float* __restrict__ qss, double* __restrict__ fglob) {
  extern __shared__ __align__(16) float smf[];
  Shared sh;
  load_shared(P, a, smf, sh);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int site = blockIdx.x * WAVES + w;
  if (site >= n_sites) return;
  const int NLM = P.NLM, N = P.N, L = P.L, D = P.D, CN = N * NLM;
  float* C = smf + shared_floats(NLM, P.n_species) + w * wave_floats(P, true);
  float* H = C + CN;
  float* U = H + CN;
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  coefficients(P, a, sh, b0, nn, C, U, U + CH * NLM);
  float* X = U;
  const float* qrow = q_in + (size_t)site * P.Dpad;
  const float* srow = qss + (size_t)site * P.Dpad;
  float acc = 0.f;
  for (int d = lane; d < D; d += 64) acc += srow[d] * qrow[d];
  const float sq = wsum(acc);
  const float nrm = pnorm_in[site];
  const float inv_n = nrm > 0.f ? 1.f / nrm : 0.f;
  for (int d = lane; d < D; d += 64) X[d] = (srow[d] - sq * qrow[d]) * inv_n;
  wave_sync();
  const int L1 = L + 1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lm = lane + 64 * h;
    if (lm < NLM) {
      int l = (int)sqrtf((float)lm + 0.5f);
      if (l * l > lm) --l;
      float cc[NMAXF];
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) cc[n] = n < N ? C[n * NLM + lm] : 0.f;
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) {
        if (n >= N) break;
        float hs = 0.f;
#pragma unroll
        for (int np = 0; np < NMAXF; ++np) {
          if (np >= N) break;
          const int lo = n < np ? n : np, hi = n < np ? np : n;
          const int pidx = lo * N - lo * (lo - 1) / 2 + (hi - lo);
          hs += (n == np ? 2.f : (float)M_SQRT2) * X[pidx * L1 + l] * cc[np];
        }
        H[n * NLM + lm] = hs;
      }
    }
  }
  wave_sync();
  float hreg[2][NMAXF];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lm = lane + 64 * h;
#pragma unroll
    for (int n = 0; n < NMAXF; ++n) hreg[h][n] = (lm < NLM && n < N) ? H[n * NLM + lm] : 0.f;
  }
  float* Xa = C;
  float* Xb = Xa + CH * NLM;
  float* Rr = Xb + CH * NLM;
  float* dRr = Rr + CH * NMAXF;
  const int j = lane / LP, part = lane - j * LP;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int k0 = 1; k0 < nn; k0 += CH) {
    const int cnt = min(CH, nn - k0);
    int s = -1;
    float r = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
    int64_t p = 0;
    wave_sync();
    if (j < cnt) {
      p = b0 + k0 + j;
      s = a.pr_spec[p];
      if (s >= 0 && !(a.pr_r[p] < (double)P.sp[s].rc)) s = -1;
      if (part == 0) {
        float* rr_ = Rr + j * NMAXF;
        float* dr_ = dRr + j * NMAXF;
#pragma unroll
        for (int n = 0; n < NMAXF; ++n) { rr_[n] = 0.f; dr_[n] = 0.f; }
      }
      if (s >= 0) {
        r = (float)a.pr_r[p];
        const float inv = 1.f / r;
        ux = (float)a.pr_xyz[3 * p] * inv; uy = (float)a.pr_xyz[3 * p + 1] * inv; uz = (float)a.pr_xyz[3 * p + 2] * inv;
        if (part == 0) {
          const SpecLite sp = sh.sp[s];
          float R[TGX_MAX_NB], dR[TGX_MAX_NB];
          radial_tab<true>(a, sp, s, r, R, dR);
          float* rr_ = Rr + j * NMAXF;
          float* dr_ = dRr + j * NMAXF;
#pragma unroll
          for (int n = 0; n < TGX_MAX_NB; ++n)
            if (n < sp.nb) { rr_[sp.off + n] = R[n]; dr_[sp.off + n] = dR[n]; }
        }
      }
    }
    wave_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lm = lane + 64 * h;
      if (lm < NLM)
        for (int jj = 0; jj < cnt; ++jj) {
          const float* rj = Rr + jj * NMAXF;
          const float* dj = dRr + jj * NMAXF;
          float xa = 0.f, xb = 0.f;
#pragma unroll
          for (int n = 0; n < NMAXF; ++n) {
            xa += rj[n] * hreg[h][n];
            xb += dj[n] * hreg[h][n];
          }
          Xa[jj * NLM + lm] = xa;
          Xb[jj * NLM + lm] = xb;
        }
    }
    wave_sync();
    float f0 = 0.f, f1 = 0.f, f2 = 0.f;
    if (j < cnt && s >= 0) {
      const SpecLite sp = sh.sp[s];
      const float inv_r = 1.f / r;
      float A[TGX_MAX_L + 1], dA[TGX_MAX_L + 1];
      damping(sp, r, A, dA);
      const float* xa = Xa + j * NLM;
      const float* xb = Xb + j * NLM;
      ylm_part<true>(L, part, sh.ynorm, ux, uy, uz, [&](int l, int lm, float Y, float gx, float gy, float gz) {
        const float x = xa[lm], y = xb[lm];
        const float rad = dA[l] * x + A[l] * y;
        const float ang = A[l] * x * inv_r;
        const float lf = (float)l;
        f0 += rad * Y * ux + ang * (gx - lf * Y * ux);
        f1 += rad * Y * uy + ang * (gy - lf * Y * uy);
        f2 += rad * Y * uz + ang * (gz - lf * Y * uz);
      });
    }
    // combine the LP lanes of each pair (lanes j*LP .. j*LP+2)
    const float g0 = f0 + __shfl(f0, lane + 1) + __shfl(f0, lane + 2);
    const float g1 = f1 + __shfl(f1, lane + 1) + __shfl(f1, lane + 2);
    const float g2 = f2 + __shfl(f2, lane + 1) + __shfl(f2, lane + 2);
    if (j < cnt && part == 0) {
      const int64_t gi = a.pr_gidx[p];
      fglob[3 * gi] += (double)g0;
      fglob[3 * gi + 1] += (double)g1;
      fglob[3 * gi + 2] += (double)g2;
      s0 += g0;
      s1 += g1;
      s2 += g2;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 += __shfl_xor(s0, o);
    s1 += __shfl_xor(s1, o);
    s2 += __shfl_xor(s2, o);
  }
  if (lane == 0) {
    const int64_t gi = a.pr_gidx[b0];
    fglob[3 * gi] -= s0;
    fglob[3 * gi + 1] -= s1;
    fglob[3 * gi + 2] -= s2;
  }
}

// Radial table: R_n(r_k), dR_n/dr(r_k) in FP64 from the quadrature, stored FP32.
// One thread per (species, node).
__global__ void k_radial_table(SoapParams P, int K, float* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n_species * K) return;
  const int s = i / K, k = i - s * K;
  const SpeciesParams& sp = P.sp[s];
  const double r = sp.rc * (double)k / (double)(K - 1);
  const double inv_rc = sp.inv_rc;
  const double x = r * inv_rc;
  const double sig = sp.sig_r * (1.0 + sp.sig_r_sc * r) * inv_rc;
  const double dsig = sp.sig_r * sp.sig_r_sc * inv_rc;
  const double inv_s2 = 1.0 / (sig * sig);
  double I[TGX_MAX_NB], dI[TGX_MAX_NB];
  for (int n = 0; n < TGX_MAX_NB; ++n) { I[n] = 0; dI[n] = 0; }
  for (int q = 0; q < sp.nq; ++q) {
    const double dx = P.xq[s * TGX_MAX_Q + q] - x;
    const double g = exp(-0.5 * dx * dx * inv_s2);
    const double dg = g * (dx * inv_s2 * inv_rc + dx * dx * inv_s2 / sig * dsig);
    for (int n = 0; n < sp.nb; ++n) {
      const double b = P.B[((size_t)s * TGX_MAX_NB + n) * TGX_MAX_Q + q];
      I[n] += b * g;
      dI[n] += b * dg;
    }
  }
  const double a = sp.amp_a;
  const double Pp = 1.0 + 2.0 * x * x * x - 3.0 * x * x;
  const double dPp = (6.0 * x * x - 6.0 * x) * inv_rc;
  double amp, damp;
  if (a == 0.0) {
    amp = 1.0 / sig;
    damp = -dsig / (sig * sig);
  } else if (Pp <= 1.0e-10) {
    amp = 0.0;
    damp = 0.0;
  } else {
    const double Pa = pow(Pp, a);
    amp = Pa / sig;
    damp = a * pow(Pp, a - 1.0) * dPp / sig - Pa * dsig / (sig * sig);
  }
  amp *= sp.gs;
  damp *= sp.gs;
  float* o = tab + (size_t)i * 2 * TGX_MAX_NB;
  for (int n = 0; n < TGX_MAX_NB; ++n) {
    o[n] = n < sp.nb ? (float)(amp * I[n]) : 0.f;
    o[TGX_MAX_NB + n] = n < sp.nb ? (float)(damp * I[n] + amp * dI[n]) : 0.f;
  }
}

}  // namespace

static constexpr int KTAB = 2048;

int tgx_radial_table_build(tgx_ctx* ctx, tgx_model_t* m) {
  TGX_TRY(m->rtab.ensure(sizeof(float) * (size_t)m->n_species * KTAB * 2 * TGX_MAX_NB));
  const int n = m->n_species * KTAB;
  TGX_LAUNCH(ctx, "radial_table", k_radial_table, dim3((n + 127) / 128), dim3(128), 0, m->P, KTAB, m->rtab.as<float>());
  return TGX_OK;
}

static F32Args f32_args(tgx_model_t* m) {
  F32Args a;
  a.site_beg = m->site_beg.as<int64_t>();
  a.pr_spec = m->pr_spec.as<int>();
  a.pr_r = m->pr_r.as<double>();
  a.pr_xyz = m->pr_xyz.as<double>();
  a.pr_gidx = m->pr_gidx.as<int64_t>();
  a.rtab = m->rtab.as<float>();
  a.K = KTAB;
  return a;
}

static size_t f32_lds(const SoapParams& P, bool adj) {
  return sizeof(float) * ((size_t)shared_floats(P.NLM, P.n_species) + (size_t)WAVES * wave_floats(P, adj)) + 64;
}

int tgx_desc_f32(tgx_ctx* ctx, tgx_model_t* m) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_REQUIRE(m->P.N <= NMAXF, "FP32 descriptor kernel holds at most %d radial functions", NMAXF);
  const size_t lds = f32_lds(m->P, false);
  TGX_REQUIRE(lds <= 160 * 1024, "FP32 descriptor kernel needs %zu bytes of LDS", lds);
  TGX_LAUNCH(ctx, "soap_desc_f32", k_desc_f32, dim3((ns + WAVES - 1) / WAVES), dim3(64 * WAVES), lds, m->P, f32_args(m),
             ns, m->q32.as<float>(), m->pnorm.as<float>());
  return TGX_OK;
}

int tgx_adjoint_f32(tgx_ctx* ctx, tgx_model_t* m) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_REQUIRE(m->P.N <= NMAXF, "FP32 adjoint kernel holds at most %d radial functions", NMAXF);
  const size_t lds = f32_lds(m->P, true);
  TGX_REQUIRE(lds <= 160 * 1024, "FP32 adjoint kernel needs %zu bytes of LDS", lds);
  TGX_LAUNCH(ctx, "soap_adjoint_f32", k_adjoint_f32, dim3((ns + WAVES - 1) / WAVES), dim3(64 * WAVES), lds, m->P,
             f32_args(m), ns, m->q32.as<float>(), m->pnorm.as<float>(), m->qss32.as<float>(), ctx->fglob.as<double>());
  return TGX_OK;
}
