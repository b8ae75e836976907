// FP32 SOAP kernels for the mixed-precision and surrogate paths.
//
// Radial functions: each species' R_n(r) and dR_n/dr (amplitude, filter and
// global scaling included) are tabulated once per model on a uniform grid of
// K nodes on [0, rcut_hard] from the exact FP64 quadrature, and evaluated by
// cubic Hermite interpolation (error ~ h^4 R''''/384, below 1e-9 relative at
// K = 2048); the force uses the derivative of the same interpolant.
//
// Three kernels per block and step:
//   k_pair_basis   one thread per pair: the damped harmonics W_lm = A_l(r) Y_lm
//                  and R_n, dR_n/dr (stacked by species) as one row per pair,
//                  staged in LDS and written as contiguous 64-row panels;
//   k_site_desc    one wavefront per site, lanes over lm: c(n,lm) = sum_j R_n W_lm
//                  from the site's rows (coalesced row loads, R by broadcast),
//                  then the power spectrum and q = p/|p|;
//   k_site_adjoint one wavefront per site: c again, H(n,lm) = sum_n' S c with
//                  S from g = (Qss - (Qss.q) q)/|p|, and for every pair
//                  f = u T + V / r with
//                    T = sum_lm (l(l+1) k' x + y - l x / r) W_lm,
//                    V = sum_lm x (A_l/A_(l-1)) sum_t C_t W_(l-1, m_t),
//                  x = sum_n R_n H, y = sum_n dR_n H (k' = -dk/dr, k = st^2/(2r^2));
//                  the gradients of the solid harmonics come from the degree
//                  below through the lowering table C (tgx_sph_host.h), so no
//                  harmonic is recomputed.
#include "tgx_internal.h"
#include "tgx_sph_host.h"

namespace {

constexpr int WAVES = 4;
constexpr int NMAXF = 16;  // stacked radial functions held per lane
constexpr int KLOW = 2;    // terms per Cartesian component in the lowering table

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
  float* basis;        // [pairs][ldb]: W (NLM, padded to roff) | R (16 slots) | dR (16 slots)
  int ldb, roff;
  int split8;          // every species has <= 8 radial functions, stacked at 0 or 8
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

typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16), aligned(4)));
typedef float f32x8_t __attribute__((ext_vector_type(8), aligned(4)));  // rows are only dword-aligned in general

// All damped harmonics of one direction: w[lm] = A[l] Y_lm(u), the recurrences
// of tgx_sph_host.h at r = 1.
__device__ __forceinline__ void harmonics_all(int L, const float* ynorm, float ux, float uy, float uz,
                                              const float (&A)[TGX_MAX_L + 1], float* w) {
  float c = 1, s = 0;
  for (int m = 0; m <= L; ++m) {
    if (m > 0) {
      const float nc = c * ux - s * uy, ns = c * uy + s * ux;
      c = nc;
      s = ns;
    }
    float q = 1, p = 0;
    for (int l = m; l <= L; ++l) {
      if (l == m + 1) {
        p = q;
        q = (float)(2 * m + 1) * uz;
      } else if (l > m + 1) {
        const float nq = ((float)(2 * l - 1) * uz * q - (float)(l + m - 1) * p) / (float)(l - m);
        p = q;
        q = nq;
      }
      const int base = l * l + l;
      const float v = A[l] * ynorm[base + m] * q;
      if (m == 0) {
        w[base] = v;
      } else {
        w[base + m] = v * c;
        w[base - m] = v * s;
      }
    }
  }
}

// The same with l_max known at compile time: every index is a constant.
template <int LT>
__device__ __forceinline__ void harmonics_fixed(const float* ynorm, float ux, float uy, float uz,
                                                const float (&A)[TGX_MAX_L + 1], float* w) {
  float c = 1, s = 0;
#pragma unroll
  for (int m = 0; m <= LT; ++m) {
    if (m > 0) {
      const float nc = c * ux - s * uy, ns = c * uy + s * ux;
      c = nc;
      s = ns;
    }
    float q = 1, p = 0;
#pragma unroll
    for (int l = m; l <= LT; ++l) {
      if (l == m + 1) {
        p = q;
        q = (float)(2 * m + 1) * uz;
      } else if (l > m + 1) {
        const float nq = ((float)(2 * l - 1) * uz * q - (float)(l + m - 1) * p) * (1.0f / (float)(l - m));
        p = q;
        q = nq;
      }
      const int base = l * l + l;
      const float v = A[l] * ynorm[base + m] * q;
      if (m == 0) {
        w[base] = v;
      } else {
        w[base + m] = v * c;
        w[base - m] = v * s;
      }
    }
  }
}

constexpr int PB = 64;  // pairs (threads) per k_pair_basis workgroup

// One pair per thread. Rows leave through LDS in two passes so that the
// staging buffer holds only the wider part (harmonics and padding, roff
// columns) at a time: first the damped harmonics, then R and dR. The smaller
// buffer lets more workgroups share a CU.
__global__ __launch_bounds__(PB) void k_pair_basis(SoapParams P, F32Args a, int64_t n_pairs) {
  extern __shared__ __align__(16) float smf[];
  Shared sh;
  load_shared(P, a, smf, sh);
  const int LS = a.roff + 1;  // odd row pitch: the per-thread row writes hit distinct banks
  float* st = smf + shared_floats(P.NLM, P.n_species);
  const int t = threadIdx.x;
  float* row = st + t * LS;
  __syncthreads();  // the shared tables are loaded
  const int64_t p0 = (int64_t)blockIdx.x * PB;
  const int64_t p = p0 + t;
  int s = -1;
  double rd = 0.0;
  if (p < n_pairs) {
    s = a.pr_spec[p];
    rd = a.pr_r[p];
    // the self pair (r = 0) and neighbours at or beyond their species' cutoff carry nothing
    if (s >= 0 && !(rd > 0.0 && rd < P.sp[s].rc)) s = -1;
  }
  const int64_t left = n_pairs - p0;
  const int nrows = left < PB ? (int)left : PB;
  float* dst = a.basis + p0 * a.ldb;
  float R[TGX_MAX_NB], dR[TGX_MAX_NB];
  SpecLite sp{};
  // pass 1: harmonics and the padding up to roff
  if (s >= 0) {
    sp = sh.sp[s];
    const float r = (float)rd, inv = 1.f / r;
    const float ux = (float)a.pr_xyz[3 * p] * inv, uy = (float)a.pr_xyz[3 * p + 1] * inv,
                uz = (float)a.pr_xyz[3 * p + 2] * inv;
    radial_tab<true>(a, sp, s, r, R, dR);
    float A[TGX_MAX_L + 1], dA[TGX_MAX_L + 1];
    damping(sp, r, A, dA);
    if (P.L == 8) harmonics_fixed<8>(sh.ynorm, ux, uy, uz, A, row);
    else harmonics_all(P.L, sh.ynorm, ux, uy, uz, A, row);
    for (int i = P.NLM; i < a.roff; ++i) row[i] = 0.f;
  } else {
    for (int i = 0; i < a.roff; ++i) row[i] = 0.f;
  }
  __syncthreads();
  for (int rr = 0; rr < nrows; ++rr)
    for (int cc = t; cc < a.roff; cc += PB) dst[(size_t)rr * a.ldb + cc] = st[rr * LS + cc];
  __syncthreads();  // pass 1 read out: the rows are reused
  // pass 2: R and dR (2 NMAXF columns from roff), the block's species at its offset
  for (int i = 0; i < 2 * NMAXF; ++i) row[i] = 0.f;
  if (s >= 0) {
#pragma unroll
    for (int n = 0; n < TGX_MAX_NB; ++n)
      if (n < sp.nb) {
        row[sp.off + n] = R[n];
        row[NMAXF + sp.off + n] = dR[n];
      }
  }
  __syncthreads();
  for (int rr = 0; rr < nrows; ++rr)
    for (int cc = t; cc < 2 * NMAXF; cc += PB) dst[(size_t)rr * a.ldb + a.roff + cc] = st[rr * LS + cc];
}

// k_pair_basis for l_max = 8 (roff = 84): the pair's 81 damped harmonics stay
// in registers and leave in three passes of 28 columns, then R and dR in one
// pass of 32, through a 64 x 33 staging buffer (8.4 KB instead of 64 x 85
// floats), so that LDS no longer bounds the number of resident workgroups.
// Each pass stores its 64 rows x W columns as one flat run over the threads.
constexpr int PB8_W = 28;  // harmonic columns per pass
constexpr int PB8_LS = 33; // staging pitch (odd)

__global__ __launch_bounds__(PB) void k_pair_basis8(SoapParams P, F32Args a, int64_t n_pairs) {
  extern __shared__ __align__(16) float smf[];
  Shared sh;
  load_shared(P, a, smf, sh);
  float* st = smf + shared_floats(P.NLM, P.n_species);
  const int t = threadIdx.x;
  float* row = st + t * PB8_LS;
  __syncthreads();  // the shared tables are loaded
  const int64_t p0 = (int64_t)blockIdx.x * PB;
  const int64_t p = p0 + t;
  int s = -1;
  double rd = 0.0;
  if (p < n_pairs) {
    s = a.pr_spec[p];
    rd = a.pr_r[p];
    // the self pair (r = 0) and neighbours at or beyond their species' cutoff carry nothing
    if (s >= 0 && !(rd > 0.0 && rd < P.sp[s].rc)) s = -1;
  }
  const int64_t left = n_pairs - p0;
  const int nrows = left < PB ? (int)left : PB;
  float* dst = a.basis + p0 * a.ldb;
  float W[3 * PB8_W];
  float R[TGX_MAX_NB], dR[TGX_MAX_NB];
  SpecLite sp{};
#pragma unroll
  for (int i = 0; i < 3 * PB8_W; ++i) W[i] = 0.f;
  if (s >= 0) {
    sp = sh.sp[s];
    const float r = (float)rd, inv = 1.f / r;
    const float ux = (float)a.pr_xyz[3 * p] * inv, uy = (float)a.pr_xyz[3 * p + 1] * inv,
                uz = (float)a.pr_xyz[3 * p + 2] * inv;
    radial_tab<true>(a, sp, s, r, R, dR);
    float A[TGX_MAX_L + 1], dA[TGX_MAX_L + 1];
    damping(sp, r, A, dA);
    harmonics_fixed<8>(sh.ynorm, ux, uy, uz, A, W);
  }
  // rows go out in 16-byte stores: ldb and the column offsets 28 pass and roff
  // are multiples of 4, and the basis array is 16-byte aligned
  constexpr int W4 = PB8_W / 4;
  static_assert(PB8_W % 4 == 0 && (2 * NMAXF) % 4 == 0, "whole 16-byte pieces per pass");
#pragma unroll
  for (int pass = 0; pass < 3; ++pass) {
#pragma unroll
    for (int k = 0; k < PB8_W; ++k) row[k] = W[pass * PB8_W + k];
    __syncthreads();
    for (int e = t; e < nrows * W4; e += PB) {
      const int rr = e / W4, c4 = e - rr * W4;
      const float* sr = st + rr * PB8_LS + 4 * c4;
      *reinterpret_cast<float4*>(dst + (size_t)rr * a.ldb + pass * PB8_W + 4 * c4) = float4{sr[0], sr[1], sr[2], sr[3]};
    }
    __syncthreads();  // the staging rows are rewritten by the next pass
  }
  // R and dR (2 NMAXF columns from roff), the block's species at its offset
#pragma unroll
  for (int i = 0; i < 2 * NMAXF; ++i) row[i] = 0.f;
  if (s >= 0) {
#pragma unroll
    for (int n = 0; n < TGX_MAX_NB; ++n)
      if (n < sp.nb) {
        row[sp.off + n] = R[n];
        row[NMAXF + sp.off + n] = dR[n];
      }
  }
  __syncthreads();
  for (int e = t; e < nrows * (NMAXF / 2); e += PB) {
    const int rr = e / (NMAXF / 2), c4 = e - rr * (NMAXF / 2);
    const float* sr = st + rr * PB8_LS + 4 * c4;
    *reinterpret_cast<float4*>(dst + (size_t)rr * a.ldb + a.roff + 4 * c4) = float4{sr[0], sr[1], sr[2], sr[3]};
  }
}

// c(n, lm) of a site in registers, lanes over lm (lm = lane and lane + 64).
// SPLIT allows the species-split accumulation (fewer FMAs, more registers).
template <bool SPLIT>
__device__ __forceinline__ void site_coefficients(const SoapParams& P, const F32Args& a, int64_t b0, int nn, int lane,
                                                  float (&acc)[2][NMAXF]) {
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int n = 0; n < NMAXF; ++n) acc[h][n] = 0.f;
  {
    const int s = a.pr_spec[b0];
    if (s >= 0 && lane == 0) {
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) {
        const int nl = n - P.sp[s].off;
        if (nl >= 0 && nl < P.sp[s].nb) acc[0][n] = (float)P.cen[s * TGX_MAX_NB + nl];
      }
    }
  }
  const bool v0 = lane < P.NLM, v1 = lane + 64 < P.NLM;
  // chunks of CCH pairs: all of a chunk's row loads are issued before its FMAs
  constexpr int CCH = 8;
  for (int j0 = 1; j0 < nn; j0 += CCH) {
    float w0[CCH], w1[CCH];
#pragma unroll
    for (int u = 0; u < CCH; ++u) {
      const float* row = a.basis + (b0 + j0 + u) * a.ldb;
      const bool ok = j0 + u < nn;
      w0[u] = ok && v0 ? row[lane] : 0.f;
      w1[u] = ok && v1 ? row[lane + 64] : 0.f;
    }
    if (SPLIT && a.split8) {
      // at most 8 radial functions per species, stacked at 0 or 8: each pair
      // feeds only its species' half, selected by an unrolled species loop so
      // that every accumulator index stays static
      int off[CCH];
#pragma unroll
      for (int u = 0; u < CCH; ++u) off[u] = (j0 + u < nn && a.pr_spec[b0 + j0 + u] > 0) ? 8 : 0;
#pragma unroll
      for (int sh = 0; sh < 2; ++sh) {
#pragma unroll
        for (int u = 0; u < CCH; ++u) {
          if (j0 + u < nn && off[u] == 8 * sh) {
            const f32x8_t rv = *reinterpret_cast<const f32x8_t*>(a.basis + (b0 + j0 + u) * a.ldb + a.roff + 8 * sh);
#pragma unroll
            for (int n = 0; n < 8; ++n) {
              acc[0][8 * sh + n] += rv[n] * w0[u];
              acc[1][8 * sh + n] += rv[n] * w1[u];
            }
          }
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < CCH; ++u) {
        if (j0 + u < nn) {
          const f32x16_t rv = *reinterpret_cast<const f32x16_t*>(a.basis + (b0 + j0 + u) * a.ldb + a.roff);
#pragma unroll
          for (int n = 0; n < NMAXF; ++n) {
            acc[0][n] += rv[n] * w0[u];
            acc[1][n] += rv[n] * w1[u];
          }
        }
      }
    }
  }
}

// per wavefront: c [N][NLM], then the spectrum X [D] over it (each a multiple of 4 floats)
__host__ __device__ inline int desc_wave_floats(const SoapParams& P) {
  const int c = (P.N * P.NLM + 3) & ~3, x = (P.D + 3) & ~3;
  return c > x ? c : x;
}
constexpr int DESC_NT = 8;  // most lm tiles of 16 of k_site_desc (NLM <= 128)
constexpr int APASS = 16;  // pair rows staged per pass of the adjoint
// per wavefront: max(X [D], APASS rows) then the pass's pair forces [APASS][3]
__host__ __device__ inline int adj_rows_floats(const SoapParams& P, int ldb) {
  const int r = APASS * ldb;
  return P.D > r ? ((P.D + 3) & ~3) : r;
}
__host__ __device__ inline int adj_wave_floats(const SoapParams& P, int ldb) { return adj_rows_floats(P, ldb) + 16 * APASS; }

// One wavefront per site; l_max is a template argument (NLM = (LL + 1)^2 <=
// 128), so the coefficient and power-spectrum stages are chains of matrix-core
// steps without run-time guards: with a guarded MFMA the compiler copies every
// accumulator out of and back into the accumulation registers on each step, and
// each guarded load becomes its own branch with a full wait.
template <int LL>
__global__ __launch_bounds__(64 * WAVES, 5) void k_site_desc(SoapParams P, F32Args a, int n_sites,
                                                          float* __restrict__ q_out, __bf16* __restrict__ qb_out,
                                                          float* __restrict__ pnorm_out) {
  constexpr int L1 = LL + 1, NLM = L1 * L1, NT = (NLM + 15) / 16;
  static_assert(NT <= DESC_NT, "at most 128 harmonics");
  extern __shared__ __align__(16) float smf[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int site = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + w);
  if (site >= n_sites) return;
  const int N = P.N;
  float* C = smf + w * desc_wave_floats(P);
  float* X = C;  // the spectrum overwrites c once every read of c is in registers
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  const int li = lane & 15, lg = lane >> 4;
  // the central atom's term, lm = 0 (read here, so that it overlaps the rows)
  float cadd[4];
  bool cuse[4];
  {
    const int s = a.pr_spec[b0];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int nl = s >= 0 ? 4 * lg + r - P.sp[s].off : -1;
      cuse[r] = s >= 0 && li == 0 && nl >= 0 && nl < P.sp[s].nb;
      cadd[r] = cuse[r] ? (float)P.cen[s * TGX_MAX_NB + nl] : 0.f;
    }
  }
  // c(n, lm) = cen + sum_j R_n(j) W_lm(j) on the matrix cores, K over the
  // neighbour rows (4 per step): lane (n = lane % 16, j = k0 + lane / 16) holds
  // R_n(j) and, per lm tile t, W_(16 t + lane % 16)(j), both read straight from
  // the rows; accumulator t holds c(4 (lane / 16) + r, 16 t + lane % 16).
  // Every load is in bounds without a guard: rows past the site's last one
  // read that last row and columns past NLM read column NLM - 1. A row past the
  // end contributes R = 0 (its products are zero), and the columns past NLM
  // are never stored, so c is what the guarded form gives, bit for bit.
  f32x4_t cacc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) cacc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const int np1 = nn - 1;
  const float* rows = a.basis + (b0 + 1) * a.ldb;
  constexpr int KS = 4;  // k-steps per pass (16 rows): all of a pass's loads are issued before its MFMAs
  for (int k0 = 0; k0 < np1; k0 += 4 * KS) {
    float av[KS], bv[KS][NT];
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      const float* row = rows + (size_t)min(k0 + 4 * u + lg, np1 - 1) * a.ldb;
      av[u] = row[a.roff + li];
#pragma unroll
      for (int t = 0; t < NT; ++t) bv[u][t] = row[min(16 * t + li, NLM - 1)];
    }
    // every k-step of the pass, also those past the site's rows (R = 0 there):
    // a guard, even a uniform one, brings back the accumulator copies
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      const float ar = k0 + 4 * u + lg < np1 ? av[u] : 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) cacc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(ar, bv[u][t], cacc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (cuse[r]) cacc[0][r] += cadd[r];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int lm = 16 * t + li;
    if (lm < NLM)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (4 * lg + r < N) C[(4 * lg + r) * NLM + lm] = cacc[t][r];
  }
  wave_sync();
  // power spectrum p(n <= n', l) = w sum_m c(n, lm) c(n', lm), w = 1 or sqrt 2:
  // lane (n' = lane % 16, k = 4 q + lane / 16) holds c(n', l^2 + k); the MFMA of
  // that column with itself leaves g(4 (lane / 16) + r, n') in accumulator r.
  // The degrees are independent chains (stored unnormalised; the norm is
  // applied on the way out).
  int xo[4];
  float xw[4];
  bool xv[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int n = 4 * lg + r, np = li;
    xv[r] = n <= np && np < N;
    xo[r] = (n * N - n * (n - 1) / 2 + (np - n)) * L1;
    xw[r] = n == np ? 1.f : (float)M_SQRT2;
  }
  const float* Cn = C + min(li, N - 1) * NLM;
  // every k-step's column of c, read before X overwrites c
  float cv[LL + 1][(2 * LL + 4) / 4];
#pragma unroll
  for (int l = 0; l <= LL; ++l)
#pragma unroll
    for (int q = 0; 4 * q < 2 * l + 1; ++q) {
      const int k = 4 * q + lg;
      const float v = Cn[l * l + min(k, 2 * l)];
      cv[l][q] = li < N && k < 2 * l + 1 ? v : 0.f;
    }
  wave_sync();
  float acc = 0.f;
#pragma unroll
  for (int l = 0; l <= LL; ++l) {
    f32x4_t g = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; 4 * q < 2 * l + 1; ++q) g = __builtin_amdgcn_mfma_f32_16x16x4f32(cv[l][q], cv[l][q], g, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (xv[r]) {
        const float sv = xw[r] * g[r];
        X[xo[r] + l] = sv;
        acc += sv * sv;
      }
  }
  const float nrm = sqrtf(wsum(acc));
  const float inv = nrm > 0.f ? 1.f / nrm : 0.f;
  wave_sync();
  // the BF16 row in 16-byte stores of eight components (Dpad is a multiple of 8)
  const int D = P.D;
  uint4* qbrow = reinterpret_cast<uint4*>(qb_out + (size_t)site * P.Dpad);
  float* qrow = q_out ? q_out + (size_t)site * P.Dpad : nullptr;
  for (int d0 = 8 * lane; d0 < P.Dpad; d0 += 512) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = d0 + i < D ? X[d0 + i] * inv : 0.f;
    unsigned pk[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      pk[i] = (unsigned)__builtin_bit_cast(unsigned short, (__bf16)v[2 * i]) |
              ((unsigned)__builtin_bit_cast(unsigned short, (__bf16)v[2 * i + 1]) << 16);
    qbrow[d0 / 8] = uint4{pk[0], pk[1], pk[2], pk[3]};
    if (qrow) {  // the FP32 copy, for the adjoint kernel that reads it
      reinterpret_cast<float4*>(qrow + d0)[0] = float4{v[0], v[1], v[2], v[3]};
      reinterpret_cast<float4*>(qrow + d0)[1] = float4{v[4], v[5], v[6], v[7]};
    }
  }
  if (q_out && lane == 0) pnorm_out[site] = nrm;
}

// Sum over each 16-lane row, left in the row's last lane (lanes 15, 31, 47, 63):
// prefix sums by DPP row shifts.
__device__ __forceinline__ float row_sum16(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xf, 0xf, false));
  return v;
}

// Sum over the wavefront of a value, available in every lane: DPP row shifts
// build each 16-lane row's sum in the row's last lane, and the four row sums
// are read out through scalar registers.
__device__ __forceinline__ float wave_total(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xf, 0xf, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xf, 0xf, false));
  const int b = __builtin_bit_cast(int, v);
  return (__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 15)) +
          __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 31))) +
         (__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 47)) +
          __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 63)));
}


__global__ __launch_bounds__(64 * WAVES) void k_site_adjoint(SoapParams P, F32Args a, int n_sites,
                                                             const int* __restrict__ low_idx,
                                                             const float* __restrict__ low_coef,
                                                             const float* __restrict__ q_in,
                                                             const float* __restrict__ pnorm_in,
                                                             const float* __restrict__ qss, double* __restrict__ fglob) {
  extern __shared__ __align__(16) float smf[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int site = __builtin_amdgcn_readfirstlane(blockIdx.x * WAVES + w);
  if (site >= n_sites) return;
  const int NLM = P.NLM, N = P.N, D = P.D, L1 = P.L + 1;
  float* X = smf + w * adj_wave_floats(P, a.ldb);
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  float hreg[2][NMAXF];
  site_coefficients<true>(P, a, b0, nn, lane, hreg);
  // g = (Qss - (Qss . q) q) / |p|
  const float* qrow = q_in + (size_t)site * P.Dpad;
  const float* srow = qss + (size_t)site * P.Dpad;
  float sacc = 0.f;
  for (int d = lane; d < D; d += 64) sacc += srow[d] * qrow[d];
  const float sq = wave_total(sacc);
  const float nrm = pnorm_in[site];
  const float inv_n = nrm > 0.f ? 1.f / nrm : 0.f;
  for (int d = lane; d < D; d += 64) X[d] = (srow[d] - sq * qrow[d]) * inv_n;
  wave_sync();
  // H(n, lm) = sum_n' w(n, n') g(n, n', l) c(n', lm), w = 2 on the diagonal and
  // sqrt(2) off it; each stored g(n <= n') serves both orders. With 16 radial
  // functions the triangle is walked by fully static code.
  int lh[2];
  bool vh[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lm = lane + 64 * h;
    vh[h] = lm < NLM;
    int l = (int)sqrtf((float)lm + 0.5f);
    if (l * l > lm) --l;
    lh[h] = l;
  }
  if (N == NMAXF) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float hs[NMAXF];
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) hs[n] = 0.f;
      const float* Xl = X + (vh[h] ? lh[h] : 0);
      int pidx = 0;
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) {
        hs[n] += 2.f * Xl[pidx * L1] * hreg[h][n];
        ++pidx;
#pragma unroll
        for (int np = n + 1; np < NMAXF; ++np) {
          const float g = (float)M_SQRT2 * Xl[pidx * L1];
          hs[n] += g * hreg[h][np];
          hs[np] += g * hreg[h][n];
          ++pidx;
        }
      }
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) hreg[h][n] = vh[h] ? hs[n] : 0.f;
    }
  } else {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int l = lh[h];
      float cc[NMAXF], hs[NMAXF];
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) {
        cc[n] = hreg[h][n];
        hs[n] = 0.f;
      }
      if (vh[h]) {
        int pidx = 0;
#pragma unroll
        for (int n = 0; n < NMAXF; ++n) {
          if (n < N) {
            hs[n] += 2.f * X[pidx * L1 + l] * cc[n];
            ++pidx;
#pragma unroll
            for (int np = n + 1; np < NMAXF; ++np) {
              if (np < N) {
                const float g = (float)M_SQRT2 * X[pidx * L1 + l];
                hs[n] += g * cc[np];
                hs[np] += g * cc[n];
                ++pidx;
              }
            }
          }
        }
      }
#pragma unroll
      for (int n = 0; n < NMAXF; ++n) hreg[h][n] = hs[n];
    }
  }
  // lowering table entries of this lane's lm
  int li[2][3][KLOW];
  float lc[2][3][KLOW];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int t = 0; t < KLOW; ++t) {
        // unused terms read W_00 with a zero coefficient, so the sums need no branch
        const int k = ((lane + 64 * h) * 3 + c) * KLOW + t;
        const int ix = vh[h] ? low_idx[k] : -1;
        li[h][c][t] = ix >= 0 ? ix : 0;
        lc[h][c][t] = ix >= 0 ? low_coef[k] : 0.f;
      }
  // Pair passes: the rows of up to APASS consecutive pairs are copied into LDS
  // (over X, which H no longer needs) by one coalesced sweep, so each pass costs
  // one memory round trip. Pairs without a contribution (other species, beyond
  // their cutoff) have all-zero rows and come out with a zero force. Each pair's
  // four lane sums (T, V) are reduced within 16-lane rows by DPP and the four
  // row partials go to LDS; at the end of the pass lane u finishes pair u.
  wave_sync();
  float* rows = X;
  float* part = X + adj_rows_floats(P, a.ldb);  // [APASS][4 values][4 rows]
  const int ldb = a.ldb, roff = a.roff;
  float fs0 = 0.f, fs1 = 0.f, fs2 = 0.f;  // this lane's share of the self-pair force
  for (int j0 = 1; j0 < nn; j0 += APASS) {
    const int cnt = min(APASS, nn - j0);
    {
      const float4* src = reinterpret_cast<const float4*>(a.basis + (b0 + j0) * ldb);
      float4* dst = reinterpret_cast<float4*>(rows);
      const int n4 = cnt * ldb / 4;
      for (int t = lane; t < n4; t += 64) dst[t] = src[t];
    }
    // per-pair scalars of the pass, one lane per pair
    float pr = 1.f, px = 0.f, py = 0.f, pz = 0.f;
    int ps = 0;
    if (lane < cnt) {
      const int64_t p = b0 + j0 + lane;
      pr = (float)a.pr_r[p];
      px = (float)a.pr_xyz[3 * p];
      py = (float)a.pr_xyz[3 * p + 1];
      pz = (float)a.pr_xyz[3 * p + 2];
      ps = a.pr_spec[p];
      if (ps < 0) ps = 0;  // its row is zero: any species' damping will do
    }
    wave_sync();
    for (int u = 0; u < cnt; ++u) {
      const float* row = rows + u * ldb;
      const float r = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, pr), u));
      const int s = __builtin_amdgcn_readlane(ps, u);
      const float inv_r = __builtin_amdgcn_rcpf(r);
      float xa[2] = {0.f, 0.f}, xb[2] = {0.f, 0.f};
      if (a.split8) {
        // at most 8 radial functions per species, stacked at 0 or 8: half the FMAs
        const int off = P.sp[s].off;
        const float* rr = row + roff + off;
        const float* dd = row + roff + NMAXF + off;
        if (off == 0) {
#pragma unroll
          for (int n = 0; n < 8; ++n) {
            xa[0] += rr[n] * hreg[0][n];
            xa[1] += rr[n] * hreg[1][n];
            xb[0] += dd[n] * hreg[0][n];
            xb[1] += dd[n] * hreg[1][n];
          }
        } else {
#pragma unroll
          for (int n = 0; n < 8; ++n) {
            xa[0] += rr[n] * hreg[0][8 + n];
            xa[1] += rr[n] * hreg[1][8 + n];
            xb[0] += dd[n] * hreg[0][8 + n];
            xb[1] += dd[n] * hreg[1][8 + n];
          }
        }
      } else {
#pragma unroll
        for (int n = 0; n < NMAXF; ++n) {
          const float rn = row[roff + n], dn = row[roff + NMAXF + n];
          xa[0] += rn * hreg[0][n];
          xa[1] += rn * hreg[1][n];
          xb[0] += dn * hreg[0][n];
          xb[1] += dn * hreg[1][n];
        }
      }
      const float sgt = (float)P.sp[s].sig_t, sgs = (float)P.sp[s].sig_t_sc;
      const float stv = sgt * (1.f + sgs * r), dst = sgt * sgs;
      const float kk = stv * stv * 0.5f * inv_r * inv_r;
      const float dk = stv * stv * inv_r * inv_r * inv_r - stv * dst * inv_r * inv_r;
      float T = 0.f, V0 = 0.f, V1 = 0.f, V2 = 0.f;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float lf = (float)lh[h];
        const float Wv = vh[h] ? row[lane + 64 * h] : 0.f;
        T += (lf * (lf + 1.f) * dk * xa[h] + xb[h] - lf * xa[h] * inv_r) * Wv;
        const float xr = xa[h] * __expf(-2.f * lf * kk);
        float g[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int t = 0; t < KLOW; ++t) g[c] += lc[h][c][t] * row[li[h][c][t]];
        V0 += xr * g[0];
        V1 += xr * g[1];
        V2 += xr * g[2];
      }
      T = row_sum16(T);
      V0 = row_sum16(V0);
      V1 = row_sum16(V1);
      V2 = row_sum16(V2);
      if ((lane & 15) == 15) {
        float* pu = part + 16 * u + (lane >> 4);
        pu[0] = T;
        pu[4] = V0;
        pu[8] = V1;
        pu[12] = V2;
      }
    }
    wave_sync();
    if (lane < cnt) {
      const float4* pu = reinterpret_cast<const float4*>(part + 16 * lane);
      const float4 t4 = pu[0], a4 = pu[1], b4 = pu[2], c4 = pu[3];
      const float T = (t4.x + t4.y) + (t4.z + t4.w);
      const float V0 = (a4.x + a4.y) + (a4.z + a4.w);
      const float V1 = (b4.x + b4.y) + (b4.z + b4.w);
      const float V2 = (c4.x + c4.y) + (c4.z + c4.w);
      const float inv_r = __builtin_amdgcn_rcpf(pr);
      const float f0 = px * inv_r * T + V0 * inv_r, f1 = py * inv_r * T + V1 * inv_r, f2 = pz * inv_r * T + V2 * inv_r;
      const int64_t gi = a.pr_gidx[b0 + j0 + lane];
      fglob[3 * gi] += (double)f0;
      fglob[3 * gi + 1] += (double)f1;
      fglob[3 * gi + 2] += (double)f2;
      fs0 += f0;
      fs1 += f1;
      fs2 += f2;
    }
    wave_sync();  // rows and part are rewritten by the next pass
  }
  fs0 = wave_total(fs0);
  fs1 = wave_total(fs1);
  fs2 = wave_total(fs2);
  if (lane == 0) {
    const int64_t gi = a.pr_gidx[b0];
    fglob[3 * gi] -= (double)fs0;
    fglob[3 * gi + 1] -= (double)fs1;
    fglob[3 * gi + 2] -= (double)fs2;
  }
}

// ---- MFMA adjoint for l_max = 8 -----------------------------------------------
// One workgroup (four wavefronts) per site; every per-site product runs on the
// matrix cores (v_mfma_f32_16x16x4f32, lanes t supply A[t % 16][t / 16] and
// B[t / 16][t % 16] and hold D[4 (t / 16) + r][t % 16]):
//   c(n, lm)  = cen + sum_j R_n(j) W_lm(j)             (K over the site's pairs)
//   p(n<=n',l) = w sum_m c(n, lm) c(n', lm)            (Gram per l, as k_site_desc)
//   g          = (Qss - (Qss.p) p / |p|^2) / |p|       (q is not read: p is recomputed)
//   H_l        = S_l C_l, S_l(n, n') = w(n, n') g(n, n', l)
//   x^T = H^T R^T, y^T = H^T dR^T per tile of 16 pairs, and for the gradients
//   of the harmonics x~_c^T = H~_c^T R^T with the raised H~_c(n, lm') =
//   sum over (lm, t) lowering onto lm' of coef_t H(n, lm): the lowering of
//   W_(l-1) in V_c = sum_lm x_lm rho_l sum_t coef_t W_(l-1, m_t) moves onto H,
//   V_c = sum_lm' rho_(l'+1) W_lm' x~_c(lm'), rho_l = A_l / A_(l-1) = exp(-2 l k).
// Then per pair T = sum_lm ((l(l+1) k' - l / r) x + y) W_lm and
// f = u T + V / r, as k_site_adjoint. The site's pair rows stay in LDS from the
// c stage to the forces when there are at most AM_NRES of them (one read of the
// rows); longer sites stream their rows twice. Pairs of a 16-pair tile share a
// species when every species has at most 8 radial functions (stacked at 0 and
// 8), so each tile needs half the k-steps. Every global read of the site is
// issued at the start, and pair forces are added with f64 atomics (one per
// entry and launch, so the sum equals a plain +=) so that no read of the
// force array sits on the site's critical path.
constexpr int AM_NLM = 81;   // (l_max + 1)^2
constexpr int AM_NT = 6;     // lm tiles of 16 (the last holds lm 80 only)
constexpr int AM_HP = 20;    // LDS pitch of the [lm][n] rows of C and H
constexpr int AM_NRES = 32;  // pair rows kept in LDS through the whole site
constexpr int AM_WIN = 64;   // neighbour entries per list window
constexpr int AM_KR = 4;     // raising terms per (lm', component)
constexpr int AM_MF = 8;     // LDS floats per window entry: 1/r, k, k', u (3), pair index (int64)
constexpr int AM_QV = 2;     // Qss quads per thread (D <= 2048): floats 4 q .. 4 q + 3, q = thread + 256 j
constexpr float AM_LOG2E = 1.4426950408889634f;


__host__ __device__ inline int am_lds_floats(const SoapParams& P, int ldb) {
  return AM_NRES * ldb + AM_NLM * AM_HP + ((P.D + 3) & ~3) + 8 + AM_WIN * AM_MF + 2 * 4 * 16 * 4 + 2 * AM_WIN + 4;
}

// Degree l of lm = l^2 + l + m. sqrt(lm + 1/2) lies at least 1/(4 l + 4) above
// l and below l + 1, far more than the error of v_sqrt_f32 (one ulp), so the
// truncation of the hardware square root is exact for every lm below 2^20
// (checked exhaustively with two ulp of error either way); the correctly
// rounded sqrtf sequence and its integer correction cost five times the
// instructions, and the adjoint evaluates twelve degrees per lane and site.
__device__ __forceinline__ int am_degree(int lm) { return (int)__builtin_amdgcn_sqrtf((float)lm + 0.5f); }

// One window of neighbour entries [win0, win0 + wcnt) (wave 0): the per-pair
// data of the contributing pairs (the block's species, 0 < r < rcut of that
// species) into meta, their window indices into one list per group (the
// species when tiles share one, else a single group), the counts into gcnt.
// The loads are independent of each other, so a window costs one round trip.
struct AmEntry {
  int grp;
  float ir, k, kp, ux, uy, uz;
  int64_t gi;
};

// The raw loads of one window entry (independent of each other).
struct AmRaw {
  int s;
  double r, x, y, z;
  int64_t gi;
};

__device__ __forceinline__ AmRaw am_load_raw(const F32Args& a, int64_t p, bool valid) {
  AmRaw e{-1, 0.0, 0.0, 0.0, 0.0, 0};
  if (!valid) return e;
  e.s = a.pr_spec[p];
  e.r = a.pr_r[p];
  e.x = a.pr_xyz[3 * p];
  e.y = a.pr_xyz[3 * p + 1];
  e.z = a.pr_xyz[3 * p + 2];
  e.gi = a.pr_gidx[p];
  return e;
}

__device__ __forceinline__ AmEntry am_entry(const SoapParams& P, const F32Args& a, const AmRaw& q) {
  AmEntry e{};
  e.grp = -1;
  e.gi = q.gi;
  if (!(q.s >= 0 && q.r > 0.0 && q.r < P.sp[q.s].rc)) return e;
  e.grp = a.split8 ? q.s : 0;
  const float r = (float)q.r, ir = __builtin_amdgcn_rcpf(r);
  const float sgt = (float)P.sp[q.s].sig_t, sgs = (float)P.sp[q.s].sig_t_sc;
  const float stv = sgt * (1.f + sgs * r), dst = sgt * sgs;
  e.ir = ir;
  e.k = stv * stv * 0.5f * ir * ir;
  e.kp = stv * stv * ir * ir * ir - stv * dst * ir * ir;  // -dk/dr
  e.ux = (float)q.x * ir;
  e.uy = (float)q.y * ir;
  e.uz = (float)q.z * ir;
  return e;
}

// Everything a site needs from global memory before its first stage: the
// first AM_NRES pair rows (this thread's share), this thread's Qss values and,
// in wave 0, the first window's entries, the self pair's index and species.
struct AmPrefetch {
  int64_t b0;
  int np1;
  float4 rv[4];
  float4 qv[AM_QV];
  AmRaw raw;
  int64_t gself;
  int sself;
};

__device__ __forceinline__ void am_prefetch(const SoapParams& P, const F32Args& a, const float* __restrict__ qss,
                                            int site, int w, int lane, AmPrefetch& f) {
  f.b0 = a.site_beg[site];
  f.np1 = (int)(a.site_beg[site + 1] - f.b0) - 1;
  const int n4 = min(AM_NRES, max(f.np1, 0)) * a.ldb / 4;
  const float4* src = reinterpret_cast<const float4*>(a.basis + (f.b0 + 1) * a.ldb);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int t = threadIdx.x + 256 * u;
    f.rv[u] = t < n4 ? src[t] : float4{0.f, 0.f, 0.f, 0.f};
  }
  const float4* srow = reinterpret_cast<const float4*>(qss + (size_t)site * P.Dpad);
#pragma unroll
  for (int j = 0; j < AM_QV; ++j) {
    const int q = threadIdx.x + 256 * j;
    float4 v = 4 * q < P.D ? srow[q] : float4{0.f, 0.f, 0.f, 0.f};
    if (4 * q + 3 >= P.D) {  // the row's padding columns
      if (4 * q + 1 >= P.D) v.y = 0.f;
      if (4 * q + 2 >= P.D) v.z = 0.f;
      if (4 * q + 3 >= P.D) v.w = 0.f;
    }
    f.qv[j] = v;
  }
  f.raw = AmRaw{-1, 0.0, 0.0, 0.0, 0.0, 0};
  f.gself = 0;
  f.sself = -1;
  if (w == 0) {
    f.raw = am_load_raw(a, f.b0 + 1 + lane, lane < min(AM_WIN, max(f.np1, 0)));
    f.gself = a.pr_gidx[f.b0];
    f.sself = a.pr_spec[f.b0];
  }
}

// Sum over the four 16-lane rows of a wavefront, (row 0 + row 1) + (row 2 + row 3)
// in every lane (the order of xor-16 then xor-32 shuffles): v_permlane16_swap
// pairs rows 0/1 and 2/3, v_permlane32_swap the two halves, both on the VALU.
__device__ __forceinline__ float am_row_sum(float v) {
  const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
  const float h = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
  const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(h), __float_as_uint(h), false, false);
  return __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
}

__device__ __forceinline__ void am_store_entry(const AmEntry& e, int lane, float* meta, int* lists, int* gcnt) {
  if (e.grp >= 0) {
    float* mt = meta + lane * AM_MF;
    mt[0] = e.ir;
    mt[1] = e.k;
    mt[2] = e.kp;
    mt[3] = e.ux;
    mt[4] = e.uy;
    mt[5] = e.uz;
    *reinterpret_cast<int64_t*>(mt + 6) = e.gi;
  }
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    const uint64_t bal = __ballot(e.grp == g);
    if (e.grp == g) lists[g * AM_WIN + __popcll(bal & ((1ull << lane) - 1ull))] = lane;
    if (lane == 0) gcnt[g] = __popcll(bal);
  }
}

// Five workgroups per CU (96 VGPRs, 30.9 KB of LDS): the raising table is read
// through the caches instead of being copied into each workgroup's LDS (6 KB),
// which took the kernel from four to five resident workgroups, 0.455 -> 0.408 ms
// at 65,536 atoms, a 20-byte register spill included.
__global__ __launch_bounds__(256, 5) void k_site_adjoint_mfma(SoapParams P, F32Args a, int n_sites,
                                                              const int* __restrict__ raise_idx,
                                                              const float* __restrict__ raise_coef,
                                                              const float* __restrict__ qss,
                                                              double* __restrict__ fglob) {
  extern __shared__ __align__(16) float smf[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lg = lane >> 4;
  const int ldb = a.ldb, roff = a.roff, N = P.N, D = P.D, L1 = P.L + 1;
  float* rows = smf;                        // [AM_NRES][ldb]
  float* CH = rows + AM_NRES * ldb;         // [AM_NLM][AM_HP]: C, then H in place
  float* X = CH + AM_NLM * AM_HP;           // [D]: p, then g
  float* red = X + ((D + 3) & ~3);          // [8]
  float* meta = red + 8;                    // [AM_WIN][AM_MF]
  float* part = meta + AM_WIN * AM_MF;      // [2 buffers][4 waves][16][4]
  int* lists = reinterpret_cast<int*>(part + 2 * 4 * 16 * 4);  // [2][AM_WIN]
  int* gcnt = lists + 2 * AM_WIN;           // [4]
  const int tx = w, te = 4 + (w >> 1), tpart = w & 1, tu = w;
  const int site = blockIdx.x;
  AmPrefetch pf;
  am_prefetch(P, a, qss, site, w, lane, pf);
  const int64_t b0 = pf.b0;
  const int np1 = pf.np1;  // neighbour entries (entry 0 is the self pair)
  const int64_t p1 = b0 + 1;
  const bool resident = np1 <= AM_NRES;
  const int cnt0 = min(AM_NRES, max(np1, 0));
  float4 qv[AM_QV];
#pragma unroll
  for (int j = 0; j < AM_QV; ++j) qv[j] = pf.qv[j];
  const int64_t gself = pf.gself;
  const int sself = pf.sself;
  {
    float4* dst = reinterpret_cast<float4*>(rows);
    const int n4 = cnt0 * ldb / 4;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int t = threadIdx.x + 256 * u;
      if (t < n4) dst[t] = pf.rv[u];
    }
  }
  if (w == 0) am_store_entry(am_entry(P, a, pf.raw), lane, meta, lists, gcnt);
  __syncthreads();
  // ---- c(n, lm): this wave's lm tiles w and w + 4
  f32x4_t cacc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
  for (int c0 = 0; c0 < np1; c0 += AM_NRES) {
    const int cnt = min(AM_NRES, np1 - c0);
    if (c0 > 0) {  // a site longer than AM_NRES: the next chunk of rows
      __syncthreads();
      const float4* src = reinterpret_cast<const float4*>(a.basis + (p1 + c0) * ldb);
      float4* dst = reinterpret_cast<float4*>(rows);
      const int n4 = cnt * ldb / 4;
      for (int t = threadIdx.x; t < n4; t += 256) dst[t] = src[t];
      __syncthreads();
    }
#pragma unroll 2
    for (int k0 = 0; k0 < cnt; k0 += 4) {
      const int j = k0 + lg;
      const bool ok = j < cnt;
      const float av = ok ? rows[j * ldb + roff + li] : 0.f;  // R_n(j), n = li
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int lm = 16 * (w + 4 * t) + li;
        if (w + 4 * t < AM_NT) {
          const float bv = (ok && lm < AM_NLM) ? rows[j * ldb + lm] : 0.f;
          cacc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, cacc[t], 0, 0, 0);
        }
      }
    }
  }
  if (w == 0 && li == 0) {  // the central atom's term, lm = 0
    const int s = sself;
    if (s >= 0) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int nl = 4 * lg + r - P.sp[s].off;
        if (nl >= 0 && nl < P.sp[s].nb) cacc[0][r] += (float)P.cen[s * TGX_MAX_NB + nl];
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int lm = 16 * (w + 4 * t) + li;
    if (w + 4 * t < AM_NT && lm < AM_NLM) *reinterpret_cast<f32x4_t*>(CH + lm * AM_HP + 4 * lg) = cacc[t];
  }
  __syncthreads();

  // ---- power spectrum into X, degrees l = w, w + 4, w + 8
  for (int l = w; l < L1; l += 4) {
    const int K = 2 * l + 1;
    f32x4_t g = {0.f, 0.f, 0.f, 0.f};
    float v[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const int k = 4 * q + lg;
      // (24-bit products: full-rate on the VALU, where a 32-bit product is quarter-rate)
      v[q] = (li < N && k < K) ? CH[__umul24(l * l + min(k, 2 * l), AM_HP) + li] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (4 * q < K) g = __builtin_amdgcn_mfma_f32_16x16x4f32(v[q], v[q], g, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 4 * lg + r, np = li;
      if (n <= np && np < N) {
        const int pidx = __umul24(n, N) - n * (n - 1) / 2 + (np - n);
        X[__umul24(pidx, L1) + l] = n == np ? g[r] : (float)M_SQRT2 * g[r];
      }
    }
  }
  __syncthreads();

  // ---- g = (Qss - (Qss.p) p / |p|^2) / |p|
  {
    float s1 = 0.f, s2 = 0.f;
#pragma unroll
    for (int j = 0; j < AM_QV; ++j) {
      const int q = threadIdx.x + 256 * j;
      if (4 * q < D) {  // X holds (D + 3) / 4 whole quads; entries past D are not the spectrum's
        float4 x = *reinterpret_cast<const float4*>(X + 4 * q);
        if (4 * q + 3 >= D) {
          if (4 * q + 1 >= D) x.y = 0.f;
          if (4 * q + 2 >= D) x.z = 0.f;
          if (4 * q + 3 >= D) x.w = 0.f;
        }
        s1 += (qv[j].x * x.x + qv[j].y * x.y) + (qv[j].z * x.z + qv[j].w * x.w);
        s2 += (x.x * x.x + x.y * x.y) + (x.z * x.z + x.w * x.w);
      }
    }
    s1 = wave_total(s1);
    s2 = wave_total(s2);
    if (lane == 0) {
      red[2 * w] = s1;
      red[2 * w + 1] = s2;
    }
  }
  __syncthreads();
  {
    const float qp = (red[0] + red[2]) + (red[4] + red[6]);
    const float pp = (red[1] + red[3]) + (red[5] + red[7]);
    const float inv_n = pp > 0.f ? 1.f / sqrtf(pp) : 0.f;
    const float cq = pp > 0.f ? qp / pp : 0.f;
#pragma unroll
    for (int j = 0; j < AM_QV; ++j) {
      const int q = threadIdx.x + 256 * j;
      if (4 * q < D) {
        float4* xp = reinterpret_cast<float4*>(X + 4 * q);
        const float4 x = *xp;
        *xp = float4{(qv[j].x - cq * x.x) * inv_n, (qv[j].y - cq * x.y) * inv_n, (qv[j].z - cq * x.z) * inv_n,
                     (qv[j].w - cq * x.w) * inv_n};
      }
    }
  }
  __syncthreads();

  // ---- H_l = S_l C_l in place: tiles l = 0..7 (2l + 1 columns), then lm 64..79 and lm 80 of l = 8
  {
    int sidx[4];
    float swt[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int n = li, np = 4 * s + lg;
      if (n < N && np < N) {
        const int u = min(n, np), v = max(n, np);
        sidx[s] = __umul24(__umul24(u, N) - u * (u - 1) / 2 + (v - u), L1);
        swt[s] = n == np ? 2.f : (float)M_SQRT2;
      } else {
        sidx[s] = 0;
        swt[s] = 0.f;
      }
    }
    for (int ht = w; ht < 10; ht += 4) {
      const int l = ht < 8 ? ht : 8;
      const int base = ht < 8 ? l * l : (ht == 8 ? 64 : 80);
      const int width = ht < 8 ? 2 * l + 1 : (ht == 8 ? 16 : 1);
      f32x4_t h = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float av = swt[s] * X[sidx[s] + l];
        const float bv = li < width ? CH[(base + li) * AM_HP + 4 * s + lg] : 0.f;
        h = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, h, 0, 0, 0);
      }
      if (li < width) *reinterpret_cast<f32x4_t*>(CH + (base + li) * AM_HP + 4 * lg) = h;
    }
  }
  __syncthreads();

  // ---- this wave's operands for the pair tiles: H for lm tile w and half of
  // tile 4 + w / 2 (its x part for even w, y part for odd w), raised H~_c for
  // lm' tile w; k-step kk covers n = 4 kk + lane / 16
  float hx[4], he[4], hr[3][4];
  {
    const int lmx = 16 * tx + li, lme = 16 * te + li, lmr = 16 * tu + li;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int n = 4 * kk + lg;
      hx[kk] = CH[lmx * AM_HP + n];
      he[kk] = lme < AM_NLM ? CH[lme * AM_HP + n] : 0.f;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        float v = 0.f;
#pragma unroll
        for (int t = 0; t < AM_KR; ++t) {
          const int e = (lmr * 3 + c) * AM_KR + t;
          v += raise_coef[e] * CH[raise_idx[e] + n];
        }
        hr[c][kk] = v;
      }
    }
  }
  // degrees of this lane's output rows: lm = 16 tile + 4 (lane / 16) + r
  float lx1[4], lx2[4], le1[4], le2[4], lr1[4];
  bool ve[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int l = am_degree(16 * tx + 4 * lg + r);
    lx2[r] = (float)l;
    lx1[r] = lx2[r] * (lx2[r] + 1.f);  // exact in FP32, without a quarter-rate integer product
    const int lme = 16 * te + 4 * lg + r;
    ve[r] = lme < AM_NLM;
    const int le = am_degree(min(lme, AM_NLM - 1));
    le2[r] = (float)le;
    le1[r] = le2[r] * (le2[r] + 1.f);
    lr1[r] = -2.f * AM_LOG2E * (float)(am_degree(16 * tu + 4 * lg + r) + 1);
  }

  // ---- pair forces, windows of AM_WIN neighbour entries, tiles of 16 pairs
  float fs0 = 0.f, fs1 = 0.f, fs2 = 0.f;  // wave 0, lanes 0..15: the self pair's sum
  const int ng = a.split8 ? 2 : 1;
  int tcount = 0;
  for (int win0 = 0; win0 < np1; win0 += AM_WIN) {
    if (win0 > 0) {  // a site with more than AM_WIN entries: the next window's lists
      if (w == 0) {
        const AmRaw q = am_load_raw(a, p1 + win0 + lane, lane < min(AM_WIN, np1 - win0));
        am_store_entry(am_entry(P, a, q), lane, meta, lists, gcnt);
      }
      __syncthreads();
    }
    for (int g = 0; g < ng; ++g) {
      const int gc = gcnt[g];
      for (int mt0 = 0; mt0 < gc; mt0 += 16, ++tcount) {
        const int mrow = mt0 + li;
        const bool pv = mrow < gc;
        const int wl = pv ? lists[g * AM_WIN + mrow] : 0;
        const float* rw;
        if (resident) {
          rw = rows + (win0 + wl) * ldb;
        } else {
          // stage the tile's rows in slots 0..15 (the previous tile's reads ended at its barrier)
          const int q = threadIdx.x >> 4, c4 = threadIdx.x & 15;
          __syncthreads();
          if (mt0 + q < gc) {
            const float4* src =
                reinterpret_cast<const float4*>(a.basis + (p1 + win0 + lists[g * AM_WIN + mt0 + q]) * ldb);
            float4* dst = reinterpret_cast<float4*>(rows + q * ldb);
            for (int t = c4; t < ldb / 4; t += 16) dst[t] = src[t];
          }
          __syncthreads();
          rw = rows + li * ldb;
        }
        f32x4_t xa = {0.f, 0.f, 0.f, 0.f}, ya = xa, xe = xa, x0 = xa, x1 = xa, x2 = xa;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          if (!a.split8 || (kk >> 1) == g) {
            const int n = 4 * kk + lg;
            const float bR = pv ? rw[roff + n] : 0.f;
            const float bD = pv ? rw[roff + NMAXF + n] : 0.f;
            xa = __builtin_amdgcn_mfma_f32_16x16x4f32(hx[kk], bR, xa, 0, 0, 0);
            ya = __builtin_amdgcn_mfma_f32_16x16x4f32(hx[kk], bD, ya, 0, 0, 0);
            xe = __builtin_amdgcn_mfma_f32_16x16x4f32(he[kk], tpart ? bD : bR, xe, 0, 0, 0);
            x0 = __builtin_amdgcn_mfma_f32_16x16x4f32(hr[0][kk], bR, x0, 0, 0, 0);
            x1 = __builtin_amdgcn_mfma_f32_16x16x4f32(hr[1][kk], bR, x1, 0, 0, 0);
            x2 = __builtin_amdgcn_mfma_f32_16x16x4f32(hr[2][kk], bR, x2, 0, 0, 0);
          }
        }
        float T = 0.f, V0 = 0.f, V1 = 0.f, V2 = 0.f;
        if (pv) {
          const float* mt = meta + wl * AM_MF;
          const float ir = mt[0], k = mt[1], kp = mt[2];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float W = rw[16 * tx + 4 * lg + r];
            T += ((lx1[r] * kp - lx2[r] * ir) * xa[r] + ya[r]) * W;
            if (ve[r]) {
              const float We = rw[16 * te + 4 * lg + r];
              T += (tpart ? xe[r] : (le1[r] * kp - le2[r] * ir) * xe[r]) * We;
            }
            const float Wr = rw[16 * tu + 4 * lg + r] * __builtin_amdgcn_exp2f(lr1[r] * k);
            V0 += Wr * x0[r];
            V1 += Wr * x1[r];
            V2 += Wr * x2[r];
          }
        }
        // the four lane rows hold the same pair: (row 0 + row 1) + (row 2 + row 3)
        T = am_row_sum(T);
        V0 = am_row_sum(V0);
        V1 = am_row_sum(V1);
        V2 = am_row_sum(V2);
        float* pb = part + (tcount & 1) * (4 * 16 * 4);
        if (lg == 0) *reinterpret_cast<f32x4_t*>(pb + (w * 16 + li) * 4) = f32x4_t{T, V0, V1, V2};
        __syncthreads();
        if (w == 0 && lane < 16 && pv) {
          const f32x4_t q0 = *reinterpret_cast<const f32x4_t*>(pb + lane * 4);
          const f32x4_t q1 = *reinterpret_cast<const f32x4_t*>(pb + (16 + lane) * 4);
          const f32x4_t q2 = *reinterpret_cast<const f32x4_t*>(pb + (32 + lane) * 4);
          const f32x4_t q3 = *reinterpret_cast<const f32x4_t*>(pb + (48 + lane) * 4);
          const f32x4_t t4 = (q0 + q1) + (q2 + q3);
          const float* mt = meta + wl * AM_MF;
          const float ir = mt[0];
          const float f0 = mt[3] * t4[0] + t4[1] * ir;
          const float f1 = mt[4] * t4[0] + t4[2] * ir;
          const float f2 = mt[5] * t4[0] + t4[3] * ir;
          // each entry gets one addition per launch: the sum is the same as a plain += (no read back here)
          const int64_t gi = *reinterpret_cast<const int64_t*>(mt + 6);
          unsafeAtomicAdd(fglob + 3 * gi, (double)f0);
          unsafeAtomicAdd(fglob + 3 * gi + 1, (double)f1);
          unsafeAtomicAdd(fglob + 3 * gi + 2, (double)f2);
          fs0 += f0;
          fs1 += f1;
          fs2 += f2;
        }
      }
    }
  }
  if (w == 0) {
    fs0 = wave_total(fs0);
    fs1 = wave_total(fs1);
    fs2 = wave_total(fs2);
    if (lane == 0) {
      unsafeAtomicAdd(fglob + 3 * gself, -(double)fs0);
      unsafeAtomicAdd(fglob + 3 * gself + 1, -(double)fs1);
      unsafeAtomicAdd(fglob + 3 * gself + 2, -(double)fs2);
    }
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
  // lowering table of the solid harmonics (host, FP64), read by the FP32 adjoint
  std::vector<double> yn((size_t)m->P.NLM);
  TGX_CHECK_HIP(hipMemcpy(yn.data(), m->P.ynorm, sizeof(double) * yn.size(), hipMemcpyDeviceToHost));
  std::vector<int> li;
  std::vector<float> lc;
  double resid = 0.0;
  TGX_REQUIRE(tgx_lowering_table(m->l_max, yn.data(), KLOW, li, lc, &resid),
              "lowering table of the solid harmonics failed (l_max %d, residual %g)", m->l_max, resid);
  TGX_TRY(m->low_idx.ensure(sizeof(int) * li.size()));
  TGX_TRY(m->low_coef.ensure(sizeof(float) * lc.size()));
  TGX_CHECK_HIP(hipMemcpyAsync(m->low_idx.p, li.data(), sizeof(int) * li.size(), hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(m->low_coef.p, lc.data(), sizeof(float) * lc.size(), hipMemcpyHostToDevice, ctx->stream));
  // the raising table of the MFMA adjoint (l_max = 8): for every lm' below the
  // top degree and component c, the (lm, coef) pairs whose lowering lands on lm'
  m->have_raise = false;
  if (m->l_max == 8) {
    const int NLM = (m->l_max + 1) * (m->l_max + 1), NR = m->l_max * m->l_max;
    std::vector<int> ri((size_t)NR * 3 * AM_KR, 0);
    std::vector<float> rc((size_t)NR * 3 * AM_KR, 0.f);
    std::vector<int> cnt((size_t)NR * 3, 0);
    bool fits = true;
    for (int lm = 0; lm < NLM && fits; ++lm)
      for (int c = 0; c < 3; ++c)
        for (int t = 0; t < KLOW; ++t) {
          const int ix = li[((size_t)lm * 3 + c) * KLOW + t];
          if (ix < 0) continue;
          if (ix >= NR || cnt[(size_t)ix * 3 + c] == AM_KR) {
            fits = false;
            break;
          }
          const size_t k = ((size_t)ix * 3 + c) * AM_KR + cnt[(size_t)ix * 3 + c]++;
          ri[k] = lm * AM_HP;  // the row offset in the kernel's [lm][n] LDS array
          rc[k] = lc[((size_t)lm * 3 + c) * KLOW + t];
        }
    if (fits) {
      TGX_TRY(m->raise_idx.ensure(sizeof(int) * ri.size()));
      TGX_TRY(m->raise_coef.ensure(sizeof(float) * rc.size()));
      TGX_CHECK_HIP(hipMemcpyAsync(m->raise_idx.p, ri.data(), sizeof(int) * ri.size(), hipMemcpyHostToDevice, ctx->stream));
      TGX_CHECK_HIP(hipMemcpyAsync(m->raise_coef.p, rc.data(), sizeof(float) * rc.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
      m->have_raise = true;
    }
  }
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

// The block's FP32 adjoint runs on the matrix-core kernel (l_max = 8 with the
// raising table built).
static bool adjoint_on_mfma(const tgx_model_t* m, const F32Args& a) {
  return m->have_raise && m->P.L == 8 && m->P.NLM == AM_NLM && a.roff >= AM_NLM;
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
  a.roff = (m->P.NLM + 3) & ~3;
  a.ldb = a.roff + 2 * NMAXF;  // 16-byte aligned rows: R and dR load as whole vectors
  a.basis = m->pbasis.as<float>();
  const SoapParams& P = m->P;
  a.split8 = P.N <= NMAXF && P.n_species <= 2;
  for (int k = 0; k < P.n_species; ++k)
    if (P.sp[k].nb > 8 || P.sp[k].off != 8 * k) a.split8 = 0;
  return a;
}

// Pair basis of the current block, then its descriptors.
int tgx_desc_f32(tgx_ctx* ctx, tgx_model_t* m) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_REQUIRE(m->P.N <= NMAXF, "FP32 descriptor kernel holds at most %d radial functions", NMAXF);
  TGX_REQUIRE(m->P.NLM <= 16 * DESC_NT, "FP32 descriptor kernel holds at most %d harmonics (l_max %d)", 16 * DESC_NT,
              m->l_max);
  const int64_t np = m->n_pairs_b;
  F32Args a = f32_args(m);
  TGX_TRY(m->pbasis.ensure(sizeof(float) * (size_t)(np > 0 ? np : 1) * a.ldb));
  a.basis = m->pbasis.as<float>();
  if (np > 0 && m->P.L == 8 && a.roff == 3 * PB8_W && 2 * NMAXF < PB8_LS) {
    const size_t lds = sizeof(float) * ((size_t)shared_floats(m->P.NLM, m->P.n_species) + (size_t)PB * PB8_LS);
    TGX_LAUNCH(ctx, "soap_pair_basis", k_pair_basis8, dim3((unsigned)((np + PB - 1) / PB)), dim3(PB), lds, m->P, a,
               np);
  } else if (np > 0) {
    const size_t lds = sizeof(float) * ((size_t)shared_floats(m->P.NLM, m->P.n_species) + (size_t)PB * (a.roff + 1));
    TGX_REQUIRE(lds <= 64 * 1024, "pair basis needs %zu bytes of LDS", lds);
    TGX_LAUNCH(ctx, "soap_pair_basis", k_pair_basis, dim3((unsigned)((np + PB - 1) / PB)), dim3(PB), lds, m->P, a, np);
  }
  const size_t lds = sizeof(float) * (size_t)WAVES * desc_wave_floats(m->P);
  TGX_REQUIRE(lds <= 160 * 1024, "FP32 descriptor kernel needs %zu bytes of LDS", lds);
  TGX_TRY(m->qbf.ensure(sizeof(__bf16) * (size_t)tgx_round_up(ns, 128) * m->Dpad));
  // the matrix-core adjoint forms the power spectrum from the coefficients
  // again: the FP32 descriptors and norms are written only for the other one
  const bool fp32_copy = !adjoint_on_mfma(m, a);
  TGX_REQUIRE(m->Dpad % 8 == 0, "FP32 descriptor rows need a multiple of 8 columns (%d)", m->Dpad);
  float* q32 = fp32_copy ? m->q32.as<float>() : nullptr;
  float* pn = fp32_copy ? m->pnorm.as<float>() : nullptr;
  const dim3 grid((ns + WAVES - 1) / WAVES), block(64 * WAVES);
  switch (m->P.L) {
#define TGX_DESC_L(ll) \
  case ll:             \
    TGX_LAUNCH(ctx, "soap_desc_f32", k_site_desc<ll>, grid, block, lds, m->P, a, ns, q32, m->qbf.as<__bf16>(), pn); \
    break;
    TGX_DESC_L(0)
    TGX_DESC_L(1)
    TGX_DESC_L(2)
    TGX_DESC_L(3)
    TGX_DESC_L(4)
    TGX_DESC_L(5)
    TGX_DESC_L(6)
    TGX_DESC_L(7)
    TGX_DESC_L(8)
    TGX_DESC_L(9)
    TGX_DESC_L(10)
#undef TGX_DESC_L
    default:
      TGX_REQUIRE(false, "FP32 descriptor kernel: l_max %d", m->P.L);
  }
  return TGX_OK;
}

// Pair forces of the current block from its FP32 Qss, reading the pair basis
// that tgx_desc_f32 left for the block.
int tgx_adjoint_f32(tgx_ctx* ctx, tgx_model_t* m) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_REQUIRE(m->P.N <= NMAXF, "FP32 adjoint kernel holds at most %d radial functions", NMAXF);
  F32Args a = f32_args(m);
  TGX_REQUIRE(m->pbasis.p && m->pbasis.bytes >= sizeof(float) * (size_t)m->n_pairs_b * a.ldb,
              "FP32 adjoint: the pair basis of the block was not built");
  if (adjoint_on_mfma(m, a)) {
    // l_max = 8: the matrix-core kernel, one workgroup per site
    const size_t lds = sizeof(float) * (size_t)am_lds_floats(m->P, a.ldb);
    TGX_REQUIRE(lds <= 64 * 1024, "FP32 adjoint kernel needs %zu bytes of LDS", lds);
    TGX_LAUNCH(ctx, "soap_adjoint_f32", k_site_adjoint_mfma, dim3(ns), dim3(256), lds, m->P, a, ns,
               m->raise_idx.as<int>(), m->raise_coef.as<float>(), m->qss32.as<float>(), ctx->fglob.as<double>());
    return TGX_OK;
  }
  const size_t lds = sizeof(float) * (size_t)WAVES * adj_wave_floats(m->P, a.ldb);
  TGX_REQUIRE(lds <= 160 * 1024, "FP32 adjoint kernel needs %zu bytes of LDS", lds);
  TGX_LAUNCH(ctx, "soap_adjoint_f32", k_site_adjoint, dim3((ns + WAVES - 1) / WAVES), dim3(64 * WAVES), lds, m->P, a,
             ns, m->low_idx.as<int>(), m->low_coef.as<float>(), m->q32.as<float>(), m->pnorm.as<float>(),
             m->qss32.as<float>(), ctx->fglob.as<double>());
  return TGX_OK;
}
