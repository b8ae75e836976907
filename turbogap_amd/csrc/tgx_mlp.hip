// MLP surrogate of the per-site GAP energy: training and inference kernels.
//
// Network per soap_turbo block: q (D) -> W1 (H x D) -> SiLU -> W2 (H x H) ->
// SiLU -> w3 (H) -> E. Weights are kept as FP32 masters (Adam) and used as
// BF16 in every matrix product, with FP32 accumulation (v_mfma_f32_16x16x32_bf16
// through tgx_gemm_bf16_nt). Targets are normalised, E_n = (E - shift)/scale.
//
// Loss of a batch of B environments:
//   L = w_E mean (E_hat - E_n)^2 + w_G mean |P (g_hat - g_n)|^2
// where g = dE/dq is the descriptor-space gradient (the GAP's -Qss), and
// P = 1 - q q^T removes the component along q that no force sees (dq/dr is
// orthogonal to q for a normalised descriptor). The gradient term needs
// d/dtheta of u . g_hat(q) with u = P(g_hat - g_n) held fixed: a forward-mode
// pass along u through the network followed by reverse mode (double
// backpropagation), written out below with SiLU's first and second derivatives.
//
// Forces of the surrogate come from the SOAP adjoint kernel with
// Qss = -dE_hat/dq, the same route as the GAP's.
#include <cmath>
#include <algorithm>
#include <memory>
#include <numeric>

#include "tgx_internal.h"
#include "tgx_util.h"

int tgx_soap_descriptors_sites(tgx_ctx* ctx, tgx_model_t* m, const int* d_sites, int n_list, double* q_out,
                               double* pnorm_out);

namespace {

__device__ __forceinline__ float silu(float z) { return tgx_silu(z); }
__device__ __forceinline__ float silu1(float z) { return tgx_silu1(z); }
__device__ __forceinline__ float silu2(float z) { return tgx_silu2(z); }

// layer 1: z1 += b1; h1 = silu(z1) (bf16 and fp32); s1 = silu'(z1); s1pp = silu''(z1)
__global__ void k_l1(int64_t n, int H, float* __restrict__ z1, const float* __restrict__ b1, __bf16* __restrict__ h1b,
                     float* __restrict__ h1, float* __restrict__ s1, float* __restrict__ s1pp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float z = z1[i] + b1[i % H];
  z1[i] = z;
  const float h = silu(z);
  h1b[i] = (__bf16)h;
  if (h1) h1[i] = h;
  s1[i] = silu1(z);
  if (s1pp) s1pp[i] = silu2(z);
}

// layer 2 and output, one wavefront per row: z2 += b2, E = w3 . silu(z2) + b3,
// a2 = w3 * silu'(z2); optional training copies.
__global__ __launch_bounds__(256) void k_l2(int rows, int H, float* __restrict__ z2, const float* __restrict__ b2,
                                            const float* __restrict__ w3, const float* __restrict__ b3,
                                            float* __restrict__ E, __bf16* __restrict__ a2b, float* __restrict__ a2,
                                            float* __restrict__ h2, float* __restrict__ s2p,
                                            float* __restrict__ s2pp) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  float acc = 0.f;
  for (int j = lane; j < H; j += 64) {
    const size_t k = (size_t)row * H + j;
    const float z = z2[k] + b2[j];
    z2[k] = z;
    const float h = silu(z), d1 = silu1(z);
    acc += w3[j] * h;
    const float a = w3[j] * d1;
    a2b[k] = (__bf16)a;
    if (a2) a2[k] = a;
    if (h2) h2[k] = h;
    if (s2p) s2p[k] = d1;
    if (s2pp) s2pp[k] = silu2(z);
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) E[row] = acc + b3[0];
}

// a1 = v1 * silu'(z1)
__global__ void k_a1(int64_t n, const float* __restrict__ v1, const float* __restrict__ s1, float* __restrict__ a1,
                     __bf16* __restrict__ a1b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float a = v1[i] * s1[i];
  if (a1) a1[i] = a;
  a1b[i] = (__bf16)a;
}

// Per row: residuals of the normalised targets, u = P(g - g_n), loss partials,
// dE = 2 w_E (E - E_n)/B. One wavefront per row.
__global__ __launch_bounds__(256) void k_loss(int rows, int Dp, const float* __restrict__ E, const double* __restrict__ e_ref,
                                              const float* __restrict__ g, const float* __restrict__ g_ref,
                                              const float* __restrict__ q, double shift, double inv_scale, float w_e,
                                              float* __restrict__ u, __bf16* __restrict__ ub, float* __restrict__ dE,
                                              double* __restrict__ le, double* __restrict__ lg) {
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const float* gr = g + (size_t)row * Dp;
  const float* gf = g_ref + (size_t)row * Dp;
  const float* qr = q + (size_t)row * Dp;
  float qd = 0.f;
  for (int j = lane; j < Dp; j += 64) qd += qr[j] * (gr[j] - (float)(gf[j] * inv_scale));
  for (int o = 32; o > 0; o >>= 1) qd += __shfl_xor(qd, o);
  float uu = 0.f;
  for (int j = lane; j < Dp; j += 64) {
    const size_t k = (size_t)row * Dp + j;
    const float v = (gr[j] - (float)(gf[j] * inv_scale)) - qr[j] * qd;
    u[k] = v;
    ub[k] = (__bf16)v;
    uu += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) uu += __shfl_xor(uu, o);
  if (lane == 0) {
    const double r = (double)E[row] - (e_ref[row] - shift) * inv_scale;
    dE[row] = (float)(2.0 * w_e * r / rows);
    le[row] = r * r;
    lg[row] = uu;
  }
}

// tau1 = silu'(z1) * tau0
__global__ void k_t1(int64_t n, const float* __restrict__ s1, const float* __restrict__ t0, float* __restrict__ t1,
                     __bf16* __restrict__ t1b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = s1[i] * t0[i];
  t1[i] = t;
  t1b[i] = (__bf16)t;
}

// gamma2 = silu'(z2) tau2, zeta2 = w3 tau2 silu''(z2)
__global__ void k_z2(int64_t n, int H, const float* __restrict__ t2, const float* __restrict__ s2p,
                     const float* __restrict__ s2pp, const float* __restrict__ w3, float* __restrict__ gam2,
                     float* __restrict__ zet2, __bf16* __restrict__ zet2b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  gam2[i] = s2p[i] * t2[i];
  const float z = w3[i % H] * t2[i] * s2pp[i];
  zet2[i] = z;
  zet2b[i] = (__bf16)z;
}

// zeta1 = v1 tau0 silu''(z1) + (W2^T zeta2) silu'(z1)
__global__ void k_z1(int64_t n, const float* __restrict__ v1, const float* __restrict__ t0,
                     const float* __restrict__ s1pp, const float* __restrict__ wz, const float* __restrict__ s1,
                     float* __restrict__ zet1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  zet1[i] = v1[i] * t0[i] * s1pp[i] + wz[i] * s1[i];
}

// Transposed BF16 operands of the weight-gradient GEMMs, with the batch of the
// gradient's two terms concatenated along K (2B columns):
//   out[c][b]     = bf16(dE[b] P[b][c] + beta U[b][c])   (dE null: P[b][c]; U null: no second term)
//   out[c][B + b] = bf16(gamma V[b][c])
// for b < B, c < cols. 64 x 64 tiles go through LDS so that the reads run along
// c and the writes along b.
__global__ __launch_bounds__(256) void k_pack_t(int B, int cols, const float* __restrict__ dE,
                                                const float* __restrict__ P, float beta, const float* __restrict__ U,
                                                float gamma, const float* __restrict__ V, __bf16* __restrict__ out) {
  __shared__ float tile[64][65];
  const int c0 = blockIdx.x * 64, b0 = blockIdx.y * 64, half = blockIdx.z;
  const int tc = threadIdx.x & 63, tr = threadIdx.x >> 6;
#pragma unroll 4
  for (int i = 0; i < 16; ++i) {
    const int rr = tr + 4 * i, b = b0 + rr, c = c0 + tc;
    float v = 0.f;
    if (b < B && c < cols) {
      const size_t k = (size_t)b * cols + c;
      if (half == 0) {
        v = dE ? dE[b] * P[k] : P[k];
        if (U) v += beta * U[k];
      } else {
        v = gamma * V[k];
      }
    }
    tile[rr][tc] = v;
  }
  __syncthreads();
#pragma unroll 4
  for (int i = 0; i < 16; ++i) {
    const int cc = tr + 4 * i, c = c0 + cc, b = b0 + tc;
    if (c < cols && b < B) out[(size_t)c * (2 * (size_t)B) + (size_t)half * B + b] = (__bf16)tile[tc][cc];
  }
}

// out[i] = sum_z part[z][i] over the split-K partials, in z order.
__global__ void k_splitk_sum(int64_t n, int nz, const float* __restrict__ part, float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < nz; ++z) s += part[(size_t)z * n + i];
  out[i] = s;
}

// Column sums out[j] = sum_b (dE[b] X[b][j] + c Y[b][j]) (X null: X = 1; Y null:
// no second term) of up to four arrays at once (blockIdx.z), in two fixed-order
// passes: every 64-row chunk of 64 columns summed by 4 row groups in FP64, then
// the chunks added in order.
struct ColTask {
  const float* X;
  const float* Y;
  float c;
  int cols;
  float* out;
};
struct ColTasks {
  ColTask t[4];
  const float* dE;
  int rows;
};

__global__ __launch_bounds__(256) void k_colsum_part(ColTasks ts, double* __restrict__ part) {
  __shared__ double sh[4][64];
  const ColTask& t = ts.t[blockIdx.z];
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + tx;
  const int r0 = blockIdx.y * 64, r1 = min(ts.rows, r0 + 64);
  double acc = 0.0;
  if (j < t.cols)
    for (int b = r0 + ty; b < r1; b += 4) {
      double v = (double)ts.dE[b] * (t.X ? (double)t.X[(size_t)b * t.cols + j] : 1.0);
      if (t.Y) v += (double)t.c * (double)t.Y[(size_t)b * t.cols + j];
      acc += v;
    }
  sh[ty][tx] = acc;
  __syncthreads();
  if (ty == 0 && j < t.cols)
    part[((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x * 64 + j] = (sh[0][tx] + sh[1][tx]) + (sh[2][tx] + sh[3][tx]);
}

__global__ void k_colsum_fin(ColTasks ts, int nch, int ldp, const double* __restrict__ part) {
  const ColTask& t = ts.t[blockIdx.z];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= t.cols) return;
  double s = 0.0;
  for (int k = 0; k < nch; ++k) s += part[((size_t)blockIdx.z * nch + k) * ldp + j];
  t.out[j] = (float)s;
}

__global__ void k_adam(int64_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 copies of W1 [H][Dp], W1^T [Dp][H], W2 [H][H], W2^T [H][H].
__global__ void k_pack(int H, int Dp, const float* __restrict__ W1, const float* __restrict__ W2, __bf16* __restrict__ W1b,
                       __bf16* __restrict__ W1Tb, __bf16* __restrict__ W2b, __bf16* __restrict__ W2Tb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n1 = (int64_t)H * Dp;
  if (i < n1) {
    const int h = (int)(i / Dp), d = (int)(i % Dp);
    const __bf16 v = (__bf16)W1[i];
    W1b[i] = v;
    W1Tb[(size_t)d * H + h] = v;
  } else if (i < n1 + (int64_t)H * H) {
    const int64_t k = i - n1;
    const int a = (int)(k / H), b = (int)(k % H);
    const __bf16 v = (__bf16)W2[k];
    W2b[k] = v;
    W2Tb[(size_t)b * H + a] = v;
  }
}

// Gather dataset rows into the batch buffers.
__global__ void k_gather_rows(int rows, int Dp, const int* __restrict__ idx, const float* __restrict__ q,
                              const float* __restrict__ g, const double* __restrict__ e, float* __restrict__ bq,
                              __bf16* __restrict__ bqb, float* __restrict__ bg, double* __restrict__ be) {
  const int r = blockIdx.x;
  if (r >= rows) return;
  const int src = idx[r];
  for (int j = threadIdx.x; j < Dp; j += blockDim.x) {
    const float v = q[(size_t)src * Dp + j];
    bq[(size_t)r * Dp + j] = v;
    bqb[(size_t)r * Dp + j] = (__bf16)v;
    bg[(size_t)r * Dp + j] = g[(size_t)src * Dp + j];
  }
  if (threadIdx.x == 0) be[r] = e[src];
}

// Dataset rows from a model's FP64 results: q, e and g = dE/dq = -Qss.
__global__ void k_append(int n_take, int stride, int offset, int Dp, const double* __restrict__ q64,
                         const double* __restrict__ qss, const double* __restrict__ e_site, float* __restrict__ dq,
                         float* __restrict__ dg, double* __restrict__ de) {
  const int r = blockIdx.x;
  if (r >= n_take) return;
  const int s = offset + r * stride;
  for (int j = threadIdx.x; j < Dp; j += blockDim.x) {
    dq[(size_t)r * Dp + j] = (float)q64[(size_t)s * Dp + j];
    dg[(size_t)r * Dp + j] = (float)(-qss[(size_t)s * Dp + j]);
  }
  if (threadIdx.x == 0) de[r] = e_site[s];
}

// Surrogate output: member energies e_k = shift_k + scale_k E_k (FP64),
// e_site = mean_k e_k, spread = max_k e_k - min_k e_k, E_k = b3_k plus the nt
// partials of member k's row in epart ([K][ldm][nt]). With more
// than one member, sites whose spread exceeds the threshold are appended to
// gated[] (stat[0] counts them; the order of the list does not matter, every
// listed site is re-evaluated on its own) and stat[1] keeps the largest spread
// (as the bits of a non-negative float, which order like the values).
__global__ void k_surrogate_out(int n, int K, const double* __restrict__ epart, int ldm, int nt,
                                const float* __restrict__ b3, const double* __restrict__ shift,
                                const double* __restrict__ scale, double* __restrict__ e_site, float threshold,
                                int* __restrict__ stat, int* __restrict__ gated) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < n;
  double s = 0.0, lo = 1e300, hi = -1e300;
  if (in) {
    for (int k = 0; k < K; ++k) {
      // member k's network output: its second layer's tile partials plus b3
      const double* pk = epart + ((size_t)k * ldm + i) * nt;
      double ek = (double)b3[k];
      for (int t = 0; t < nt; ++t) ek += pk[t];
      const double v = shift[k] + scale[k] * ek;
      s += v;
      lo = v < lo ? v : lo;
      hi = v > hi ? v : hi;
    }
    e_site[i] = s / K;
  }
  if (K > 1) {
    // the largest spread as the largest bit pattern (spreads are >= 0, so the
    // integer order is the float order, NaN above all), one atomic per
    // wavefront instead of one per site on the same word
    const float sp = in ? (float)(hi - lo) : 0.f;
    int m = __float_as_int(sp);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(&stat[1], m);
    if (in && sp > threshold) gated[atomicAdd(&stat[0], 1)] = i;
  }
}

// Gated sites take the FP64 GAP energy and gradient.
__global__ void k_gate_replace(int ng, const int* __restrict__ gated, int Dp, const double* __restrict__ e_gap,
                               const double* __restrict__ qss_gap, double* __restrict__ e_site,
                               float* __restrict__ qss32) {
  const int r = blockIdx.x;
  if (r >= ng) return;
  const int s = gated[r];
  for (int j = threadIdx.x; j < Dp; j += blockDim.x) qss32[(size_t)s * Dp + j] = (float)qss_gap[(size_t)r * Dp + j];
  if (threadIdx.x == 0) e_site[s] = e_gap[r];
}

inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// ---- force-loss training (whole structures) -------------------------------------
// Batch rows of a block's centres: q (FP32 and BF16) and the FP64 site
// energies; rows past ns are zero.
__global__ void k_force_batch(int ns, int Dp, const double* __restrict__ q64, const double* __restrict__ e_site,
                              float* __restrict__ bq, __bf16* __restrict__ bqb, double* __restrict__ be) {
  const int r = blockIdx.x;
  const bool in = r < ns;
  for (int j = threadIdx.x; j < Dp; j += blockDim.x) {
    const float v = in ? (float)q64[(size_t)r * Dp + j] : 0.f;
    bq[(size_t)r * Dp + j] = v;
    bqb[(size_t)r * Dp + j] = (__bf16)v;
  }
  if (threadIdx.x == 0) be[r] = in ? e_site[r] : 0.0;
}

// -(g_hat - g) per centre, the -dE/dq rows the pair-force kernel takes: g_hat =
// scale G is the network's gradient, g = -Qss the FP64 GAP's.
__global__ void k_force_dg(int64_t n, double scale, const float* __restrict__ G, const double* __restrict__ qss,
                           double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = -(scale * (double)G[i] + qss[i]);
}

// Energy residual weights dE_b = 2 w_e r_b / ns, r_b = E_b - (e_b - shift) / scale,
// and r_b^2, for the block's centres; zero on the padding rows.
__global__ void k_force_eres(int ns, int B, const float* __restrict__ E, const double* __restrict__ e_ref,
                             double shift, double inv_scale, float w_e, float* __restrict__ dE,
                             double* __restrict__ le) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= B) return;
  const double x = r < ns ? (double)E[r] - (e_ref[r] - shift) * inv_scale : 0.0;
  dE[r] = (float)(2.0 * (double)w_e * x / ns);
  le[r] = x * x;
}

// u_i = coef sum_p der_p^T F_nbr(p) over the list entries p of centre i (the
// self entry included): the force residual carried back to the centre's
// descriptor. One workgroup per row, threads over components; padding rows and
// columns are zero.
__global__ __launch_bounds__(256) void k_force_jvp(int ns, int D, int Dp, const int64_t* __restrict__ site_beg,
                                                   const int* __restrict__ pr_nbr, const double* __restrict__ der,
                                                   const double* __restrict__ F, double coef, float* __restrict__ u,
                                                   __bf16* __restrict__ ub) {
  const int i = blockIdx.x;
  float* ur = u + (size_t)i * Dp;
  __bf16* ubr = ub + (size_t)i * Dp;
  const int64_t p0 = i < ns ? site_beg[i] : 0, p1 = i < ns ? site_beg[i + 1] : 0;
  for (int d = threadIdx.x; d < Dp; d += blockDim.x) {
    double acc = 0.0;
    if (d < D)
      for (int64_t p = p0; p < p1; ++p) {
        const double* f = F + 3 * (size_t)pr_nbr[p];
        const double* j3 = der + ((size_t)p * D + d) * 3;
        acc += j3[0] * f[0] + j3[1] * f[1] + j3[2] * f[2];
      }
    const float v = (float)(coef * acc);
    ur[d] = v;
    ubr[d] = (__bf16)v;
  }
}

}  // namespace

struct tgx_mlp {
  tgx_ctx* ctx = nullptr;
  int D = 0, Dp = 0, H = 0, maxB = 0;
  int64_t nparam = 0;
  int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, ow3 = 0, ob3 = 0;
  dbuf theta, grad, mom, vel, W1b, W1Tb, W2b, W2Tb;
  int64_t step = 0;
  double shift = 0.0, scale = 1.0;
  bool norm_set = false;
  // dataset
  int n = 0, cap = 0;
  dbuf dq, dg, de;
  // batch work space
  dbuf idx, bq, bqb, bg, be, z1, h1, h1b, s1, s1pp, z2, h2, a2, a2b, s2p, s2pp, E, v1, a1, a1b, G, u, ub, t0, t1, t1b,
      t2, gam2, zet2, zet2b, wz, zet1, dE, packA, packB, cpart, kpart, lossb;
  dbuf fdg;  // force-loss step: -(g_hat - g) rows in FP64
  uint64_t seed = 0;
};

static int mlp_workspace(tgx_mlp* m, int B) {
  const size_t H = m->H, Dp = m->Dp, b = (size_t)B;
  dbuf* f_bh[] = {&m->z1, &m->h1, &m->s1, &m->s1pp, &m->z2, &m->h2, &m->a2, &m->s2p, &m->s2pp, &m->v1, &m->a1,
                  &m->t0, &m->t1, &m->t2, &m->gam2, &m->zet2, &m->wz, &m->zet1};
  for (dbuf* x : f_bh) TGX_TRY(x->ensure(sizeof(float) * b * H));
  dbuf* h_bh[] = {&m->h1b, &m->a2b, &m->a1b, &m->t1b, &m->zet2b};
  for (dbuf* x : h_bh) TGX_TRY(x->ensure(sizeof(__bf16) * b * H));
  dbuf* f_bd[] = {&m->bq, &m->bg, &m->G, &m->u};
  for (dbuf* x : f_bd) TGX_TRY(x->ensure(sizeof(float) * b * Dp));
  TGX_TRY(m->bqb.ensure(sizeof(__bf16) * b * Dp));
  TGX_TRY(m->ub.ensure(sizeof(__bf16) * b * Dp));
  TGX_TRY(m->be.ensure(sizeof(double) * b));
  TGX_TRY(m->E.ensure(sizeof(float) * b));
  TGX_TRY(m->dE.ensure(sizeof(float) * b));
  TGX_TRY(m->lossb.ensure(sizeof(double) * 2 * b));
  TGX_TRY(m->idx.ensure(sizeof(int) * b));
  // weight-gradient operands over the concatenated batch, column-sum partials
  TGX_TRY(m->packA.ensure(sizeof(__bf16) * H * 2 * b));
  TGX_TRY(m->packB.ensure(sizeof(__bf16) * (Dp > H ? Dp : H) * 2 * b));
  TGX_TRY(m->cpart.ensure(sizeof(double) * 4 * ((b + 63) / 64) * (size_t)tgx_round_up((int64_t)H, 64)));
  return TGX_OK;
}

static int mlp_pack(tgx_mlp* m) {
  tgx_ctx* ctx = m->ctx;
  const float* th = m->theta.as<float>();
  const int64_t n = (int64_t)m->H * m->Dp + (int64_t)m->H * m->H;
  TGX_LAUNCH(ctx, "mlp_pack", k_pack, g1(n), dim3(256), 0, m->H, m->Dp, th + m->oW1, th + m->oW2, m->W1b.as<__bf16>(),
             m->W1Tb.as<__bf16>(), m->W2b.as<__bf16>(), m->W2Tb.as<__bf16>());
  return TGX_OK;
}

extern "C" int tgx_mlp_create(tgx_ctx* ctx, int dim, int hidden, int max_batch, uint64_t seed, tgx_mlp** out) {
  TGX_REQUIRE(ctx && out && dim > 0 && hidden > 0 && hidden % 128 == 0 && max_batch > 0 && max_batch % 128 == 0,
              "tgx_mlp_create: need dim > 0, hidden and max_batch multiples of 128");
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  tgx_mlp* m = new tgx_mlp();
  m->ctx = ctx;
  m->D = dim;
  m->Dp = (int)tgx_round_up(dim, 128);
  m->H = hidden;
  m->maxB = max_batch;
  m->seed = seed;
  const int64_t H = hidden, Dp = m->Dp;
  m->oW1 = 0;
  m->ob1 = m->oW1 + H * Dp;
  m->oW2 = m->ob1 + H;
  m->ob2 = m->oW2 + H * H;
  m->ow3 = m->ob2 + H;
  m->ob3 = m->ow3 + H;
  m->nparam = tgx_round_up(m->ob3 + 1, 64);
  auto fail = [&](int st) { delete m; return st; };
  int st;
  if ((st = m->theta.ensure(sizeof(float) * m->nparam)) || (st = m->grad.ensure(sizeof(float) * m->nparam)) ||
      (st = m->mom.ensure(sizeof(float) * m->nparam)) || (st = m->vel.ensure(sizeof(float) * m->nparam)) ||
      (st = m->W1b.ensure(sizeof(__bf16) * H * Dp)) || (st = m->W1Tb.ensure(sizeof(__bf16) * H * Dp)) ||
      (st = m->W2b.ensure(sizeof(__bf16) * H * H)) || (st = m->W2Tb.ensure(sizeof(__bf16) * H * H)) ||
      (st = mlp_workspace(m, max_batch)))
    return fail(st);
  // initial weights: W1 ~ N(0, 1) over the D real columns (|q| = 1), W2 ~ N(0, 1/H),
  // w3 ~ N(0, 1/H), biases zero
  std::vector<double> w((size_t)m->nparam, 0.0), tmp;
  std::vector<double> r1((size_t)H * dim), r2((size_t)H * H), r3((size_t)H);
  tgx_generate_normal((int)r1.size(), 1.0, seed, r1.data());
  tgx_generate_normal((int)r2.size(), 1.0 / std::sqrt((double)H), seed + 1, r2.data());
  tgx_generate_normal((int)r3.size(), 1.0 / std::sqrt((double)H), seed + 2, r3.data());
  for (int64_t h = 0; h < H; ++h)
    for (int d = 0; d < dim; ++d) w[m->oW1 + h * Dp + d] = r1[(size_t)h * dim + d];
  for (int64_t k = 0; k < H * H; ++k) w[m->oW2 + k] = r2[(size_t)k];
  for (int64_t k = 0; k < H; ++k) w[m->ow3 + k] = r3[(size_t)k];
  std::vector<float> wf(w.begin(), w.end());
  if (hipMemcpyAsync(m->theta.p, wf.data(), sizeof(float) * wf.size(), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    return fail(TGX_ERR_HIP);
  if ((st = tgx_zero(ctx, m->mom.p, sizeof(float) * m->nparam)) || (st = tgx_zero(ctx, m->vel.p, sizeof(float) * m->nparam)) ||
      (st = mlp_pack(m)))
    return fail(st);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(TGX_ERR_HIP);
  *out = m;
  return TGX_OK;
}

extern "C" int tgx_mlp_destroy(tgx_mlp* m) {
  if (!m) return TGX_OK;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  dbuf* all[] = {&m->theta, &m->grad, &m->mom, &m->vel, &m->W1b, &m->W1Tb, &m->W2b, &m->W2Tb, &m->dq, &m->dg, &m->de,
                 &m->idx, &m->bq, &m->bqb, &m->bg, &m->be, &m->z1, &m->h1, &m->h1b, &m->s1, &m->s1pp, &m->z2, &m->h2,
                 &m->a2, &m->a2b, &m->s2p, &m->s2pp, &m->E, &m->v1, &m->a1, &m->a1b, &m->G, &m->u, &m->ub, &m->t0,
                 &m->t1, &m->t1b, &m->t2, &m->gam2, &m->zet2, &m->zet2b, &m->wz, &m->zet1, &m->dE, &m->packA,
                 &m->packB, &m->cpart, &m->kpart, &m->lossb, &m->fdg};
  for (dbuf* b : all) b->release();
  delete m;
  return TGX_OK;
}

// Forward (energy and input gradient) on B rows of bqb (BF16) held in the batch
// buffers. With training copies when `train`.
static int mlp_forward(tgx_mlp* m, int B, const __bf16* xb, bool train, float* E_out, float* G_out, double gscale) {
  tgx_ctx* ctx = m->ctx;
  const int H = m->H, Dp = m->Dp;
  const float* th = m->theta.as<float>();
  const int64_t nbh = (int64_t)B * H;
  TGX_TRY(tgx_gemm_bf16_nt(ctx, "mlp_l1", B, H, Dp, xb, Dp, m->W1b.as<__bf16>(), Dp, m->z1.as<float>(), H, 1.0));
  TGX_LAUNCH(ctx, "mlp_act1", k_l1, g1(nbh), dim3(256), 0, nbh, H, m->z1.as<float>(), th + m->ob1, m->h1b.as<__bf16>(),
             train ? m->h1.as<float>() : nullptr, m->s1.as<float>(), train ? m->s1pp.as<float>() : nullptr);
  TGX_TRY(tgx_gemm_bf16_nt(ctx, "mlp_l2", B, H, H, m->h1b.as<__bf16>(), H, m->W2b.as<__bf16>(), H, m->z2.as<float>(), H,
                           1.0));
  TGX_LAUNCH(ctx, "mlp_act2", k_l2, dim3((B + 3) / 4), dim3(256), 0, B, H, m->z2.as<float>(), th + m->ob2, th + m->ow3,
             th + m->ob3, E_out, m->a2b.as<__bf16>(), train ? m->a2.as<float>() : nullptr,
             train ? m->h2.as<float>() : nullptr, train ? m->s2p.as<float>() : nullptr,
             train ? m->s2pp.as<float>() : nullptr);
  TGX_TRY(tgx_gemm_bf16_nt(ctx, "mlp_back2", B, H, H, m->a2b.as<__bf16>(), H, m->W2Tb.as<__bf16>(), H, m->v1.as<float>(),
                           H, 1.0));
  TGX_LAUNCH(ctx, "mlp_a1", k_a1, g1(nbh), dim3(256), 0, nbh, m->v1.as<float>(), m->s1.as<float>(),
             train ? m->a1.as<float>() : nullptr, m->a1b.as<__bf16>());
  TGX_TRY(tgx_gemm_bf16_nt(ctx, "mlp_back1", B, Dp, H, m->a1b.as<__bf16>(), H, m->W1Tb.as<__bf16>(), H, G_out, Dp,
                           gscale));
  return TGX_OK;
}

// C[M][N] = A[M][K] B[N][K]^T (BF16 operands, FP32) for the weight gradients, whose
// M x N grid has few tiles against a long K: K is split into nz chunks run as a
// batch of GEMMs into FP32 partials, which are then added in chunk order.
static int wgrad_gemm(tgx_mlp* m, const char* name, int M, int N, int K, const __bf16* A, const __bf16* B, float* C) {
  tgx_ctx* ctx = m->ctx;
  const int tiles = (M / 128) * (N / 128);
  int nz = 1;
  while (nz < 32 && tiles * nz * 2 <= 1024 && (K / (nz * 2)) % 64 == 0 && K / (nz * 2) >= 256) nz *= 2;
  if (nz == 1) return tgx_gemm_bf16_nt(ctx, name, M, N, K, A, K, B, K, C, N, 1.0);
  const int64_t mn = (int64_t)M * N;
  TGX_TRY(m->kpart.ensure(sizeof(float) * (size_t)(nz * mn)));
  tgx_gemm_desc d;
  d.epi = TGX_EPI_STORE_F32;
  d.M = M;
  d.N = N;
  d.K = K / nz;
  d.batch = nz;
  d.A = A;
  d.lda = K;
  d.sA = K / nz;
  d.B = B;
  d.ldb = K;
  d.sB = K / nz;
  d.C = m->kpart.p;
  d.ldc = N;
  d.sC = mn;
  d.scale = 1.0;
  TGX_TRY(tgx_gemm_bf16(ctx, name, d));
  TGX_LAUNCH(ctx, "mlp_splitk_sum", k_splitk_sum, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, mn, nz,
             (const float*)m->kpart.as<float>(), C);
  return TGX_OK;
}

static int mlp_backward_dirs(tgx_mlp* m, int B, float cG);

// Loss and weight gradients of the batch in the batch buffers (bq, bqb, bg, be).
static int mlp_backward(tgx_mlp* m, int B, float w_e, float w_g, double* loss_e, double* loss_g) {
  tgx_ctx* ctx = m->ctx;
  const int Dp = m->Dp;
  TGX_TRY(mlp_forward(m, B, m->bqb.as<__bf16>(), true, m->E.as<float>(), m->G.as<float>(), 1.0));
  double* le = m->lossb.as<double>();
  double* lg = le + B;
  TGX_LAUNCH(ctx, "mlp_loss", k_loss, dim3((B + 3) / 4), dim3(256), 0, B, Dp, m->E.as<float>(), m->be.as<double>(),
             m->G.as<float>(), m->bg.as<float>(), m->bq.as<float>(), m->shift, 1.0 / m->scale, w_e, m->u.as<float>(),
             m->ub.as<__bf16>(), m->dE.as<float>(), le, lg);
  TGX_TRY(mlp_backward_dirs(m, B, 2.0f * w_g / (float)B));
  if (loss_e || loss_g) {
    double se = 0.0, sg = 0.0;
    TGX_TRY(tgx_device_sum(ctx, le, B, &se));
    TGX_TRY(tgx_device_sum(ctx, lg, B, &sg));
    if (loss_e) *loss_e = se / B;
    if (loss_g) *loss_g = sg / B;
  }
  return TGX_OK;
}

// Weight gradients of a batch whose forward pass (train = true) has run, from
// the energy residual weights dE (per row) and the gradient-loss direction u
// (FP32 and BF16 rows) with its weight cG: the reverse-mode derivative of
// sum_b dE_b E_b + cG u_b . g_b with dE and u fixed.
static int mlp_backward_dirs(tgx_mlp* m, int B, float cG) {
  tgx_ctx* ctx = m->ctx;
  const int H = m->H, Dp = m->Dp;
  const int64_t nbh = (int64_t)B * H;
  float* th = m->theta.as<float>();
  float* gr = m->grad.as<float>();
  // forward mode along u, then reverse
  TGX_TRY(tgx_gemm_bf16_nt(ctx, "mlp_jvp1", B, H, Dp, m->ub.as<__bf16>(), Dp, m->W1b.as<__bf16>(), Dp, m->t0.as<float>(),
                           H, 1.0));
  TGX_LAUNCH(ctx, "mlp_t1", k_t1, g1(nbh), dim3(256), 0, nbh, m->s1.as<float>(), m->t0.as<float>(), m->t1.as<float>(),
             m->t1b.as<__bf16>());
  TGX_TRY(tgx_gemm_bf16_nt(ctx, "mlp_jvp2", B, H, H, m->t1b.as<__bf16>(), H, m->W2b.as<__bf16>(), H, m->t2.as<float>(),
                           H, 1.0));
  TGX_LAUNCH(ctx, "mlp_z2", k_z2, g1(nbh), dim3(256), 0, nbh, H, m->t2.as<float>(), m->s2p.as<float>(),
             m->s2pp.as<float>(), th + m->ow3, m->gam2.as<float>(), m->zet2.as<float>(), m->zet2b.as<__bf16>());
  TGX_TRY(tgx_gemm_bf16_nt(ctx, "mlp_vjp2", B, H, H, m->zet2b.as<__bf16>(), H, m->W2Tb.as<__bf16>(), H,
                           m->wz.as<float>(), H, 1.0));
  TGX_LAUNCH(ctx, "mlp_z1", k_z1, g1(nbh), dim3(256), 0, nbh, m->v1.as<float>(), m->t0.as<float>(), m->s1pp.as<float>(),
             m->wz.as<float>(), m->s1.as<float>(), m->zet1.as<float>());
  // dW1 = a1^T (dE q + cG u) + cG zeta1^T q = [a1 | cG zeta1]^T [dE q + cG u ; q], one
  // BF16 GEMM with FP32 accumulation over the concatenated batch (K = 2B)
  const dim3 gH(H / 64, B / 64, 2), gD(Dp / 64, B / 64, 2);
  __bf16* pA = m->packA.as<__bf16>();
  __bf16* pB = m->packB.as<__bf16>();
  TGX_LAUNCH(ctx, "mlp_packA1", k_pack_t, gH, dim3(256), 0, B, H, (const float*)nullptr, m->a1.as<float>(), 0.0f,
             (const float*)nullptr, cG, m->zet1.as<float>(), pA);
  TGX_LAUNCH(ctx, "mlp_packB1", k_pack_t, gD, dim3(256), 0, B, Dp, m->dE.as<float>(), m->bq.as<float>(), cG,
             m->u.as<float>(), 1.0f, m->bq.as<float>(), pB);
  TGX_TRY(wgrad_gemm(m, "mlp_dW1", H, Dp, 2 * B, pA, pB, gr + m->oW1));
  // dW2 = a2^T (dE h1 + cG tau1) + cG zeta2^T h1 = [a2 | cG zeta2]^T [dE h1 + cG tau1 ; h1]
  TGX_LAUNCH(ctx, "mlp_packA2", k_pack_t, gH, dim3(256), 0, B, H, (const float*)nullptr, m->a2.as<float>(), 0.0f,
             (const float*)nullptr, cG, m->zet2.as<float>(), pA);
  TGX_LAUNCH(ctx, "mlp_packB2", k_pack_t, gH, dim3(256), 0, B, H, m->dE.as<float>(), m->h1.as<float>(), cG,
             m->t1.as<float>(), 1.0f, m->h1.as<float>(), pB);
  TGX_TRY(wgrad_gemm(m, "mlp_dW2", H, H, 2 * B, pA, pB, gr + m->oW2));
  // biases and output weights: fixed-order column sums over 64-row chunks
  ColTasks ct{};
  ct.dE = m->dE.as<float>();
  ct.rows = B;
  ct.t[0] = ColTask{m->a1.as<float>(), m->zet1.as<float>(), cG, H, gr + m->ob1};
  ct.t[1] = ColTask{m->a2.as<float>(), m->zet2.as<float>(), cG, H, gr + m->ob2};
  ct.t[2] = ColTask{m->h2.as<float>(), m->gam2.as<float>(), cG, H, gr + m->ow3};
  ct.t[3] = ColTask{nullptr, nullptr, 0.0f, 1, gr + m->ob3};
  const int nch = (B + 63) / 64, ldp = (int)tgx_round_up(H, 64);
  TGX_LAUNCH(ctx, "mlp_colsum", k_colsum_part, dim3(ldp / 64, nch, 4), dim3(256), 0, ct, m->cpart.as<double>());
  TGX_LAUNCH(ctx, "mlp_colsum_fin", k_colsum_fin, dim3((ldp + 63) / 64, 1, 4), dim3(64), 0, ct, nch, ldp,
             (const double*)m->cpart.as<double>());
  return TGX_OK;
}

extern "C" int tgx_mlp_add_samples(tgx_mlp* m, int model_id, int stride, int offset, int* n_total) {
  TGX_REQUIRE(m && stride >= 1 && offset >= 0, "tgx_mlp_add_samples: bad arguments");
  tgx_ctx* ctx = m->ctx;
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  TGX_REQUIRE(model_id >= 0 && model_id < (int)ctx->models.size(), "model %d not defined", model_id);
  tgx_model_t* md = ctx->models[model_id];
  TGX_REQUIRE(md->Dpad == m->Dp && md->D == m->D, "descriptor size of model %d does not match the network", model_id);
  TGX_REQUIRE(md->q64.p && md->qss64.p && md->e_site.p && md->n_sites_b > 0,
              "tgx_mlp_add_samples: run an FP64 tgx_compute first");
  const int ns = md->n_sites_b;
  const int take = offset < ns ? (ns - offset + stride - 1) / stride : 0;
  if (m->n + take > m->cap) {
    const int ncap = std::max(m->cap * 2, m->n + take);
    dbuf q2, g2, e2;
    TGX_TRY(q2.ensure(sizeof(float) * (size_t)ncap * m->Dp));
    TGX_TRY(g2.ensure(sizeof(float) * (size_t)ncap * m->Dp));
    TGX_TRY(e2.ensure(sizeof(double) * (size_t)ncap));
    if (m->n) {
      TGX_CHECK_HIP(hipMemcpyAsync(q2.p, m->dq.p, sizeof(float) * (size_t)m->n * m->Dp, hipMemcpyDeviceToDevice, ctx->stream));
      TGX_CHECK_HIP(hipMemcpyAsync(g2.p, m->dg.p, sizeof(float) * (size_t)m->n * m->Dp, hipMemcpyDeviceToDevice, ctx->stream));
      TGX_CHECK_HIP(hipMemcpyAsync(e2.p, m->de.p, sizeof(double) * (size_t)m->n, hipMemcpyDeviceToDevice, ctx->stream));
      TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    }
    m->dq.release(); m->dg.release(); m->de.release();
    m->dq = q2; m->dg = g2; m->de = e2;
    q2.p = g2.p = e2.p = nullptr;
    m->cap = ncap;
  }
  if (take > 0)
    TGX_LAUNCH(ctx, "mlp_append", k_append, dim3(take), dim3(256), 0, take, stride, offset, m->Dp, md->q64.as<double>(),
               md->qss64.as<double>(), md->e_site.as<double>(), m->dq.as<float>() + (size_t)m->n * m->Dp,
               m->dg.as<float>() + (size_t)m->n * m->Dp, m->de.as<double>() + m->n);
  m->n += take;
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (n_total) *n_total = m->n;
  return TGX_OK;
}

extern "C" int tgx_mlp_set_normalization(tgx_mlp* m, double e_shift, double e_scale) {
  TGX_REQUIRE(m && e_scale > 0.0, "tgx_mlp_set_normalization: scale must be positive");
  m->shift = e_shift;
  m->scale = e_scale;
  m->norm_set = true;
  return TGX_OK;
}

extern "C" int tgx_mlp_get_normalization(tgx_mlp* m, double* e_shift, double* e_scale) {
  TGX_REQUIRE(m, "null network");
  if (e_shift) *e_shift = m->shift;
  if (e_scale) *e_scale = m->scale;
  return TGX_OK;
}

extern "C" int tgx_mlp_dataset_stats(tgx_mlp* m, int n_use, double* e_mean, double* e_std) {
  TGX_REQUIRE(m && n_use > 0 && n_use <= m->n, "tgx_mlp_dataset_stats: bad range");
  TGX_CHECK_HIP(hipSetDevice(m->ctx->device));
  std::vector<double> e((size_t)n_use);
  TGX_CHECK_HIP(hipMemcpy(e.data(), m->de.p, sizeof(double) * n_use, hipMemcpyDeviceToHost));
  double mu = 0.0;
  for (double v : e) mu += v;
  mu /= n_use;
  double var = 0.0;
  for (double v : e) var += (v - mu) * (v - mu);
  if (e_mean) *e_mean = mu;
  if (e_std) *e_std = std::sqrt(var / n_use);
  return TGX_OK;
}

static int adam_step(tgx_mlp* m, float lr) {
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - (float)std::pow(b1, (double)m->step), c2 = 1.f - (float)std::pow(b2, (double)m->step);
  TGX_LAUNCH(m->ctx, "mlp_adam", k_adam, g1(m->nparam), dim3(256), 0, m->nparam, m->theta.as<float>(),
             m->grad.as<float>(), m->mom.as<float>(), m->vel.as<float>(), lr, b1, b2, c1, c2, 1e-8f);
  // padded columns of W1 stay zero: their gradient is zero (q is zero there)
  return mlp_pack(m);
}

extern "C" int tgx_mlp_train(tgx_mlp* m, int n_train, int epochs, int batch, float lr, float w_e, float w_g,
                             uint64_t seed, double* loss_e, double* loss_g) {
  TGX_REQUIRE(m && n_train > 0 && n_train <= m->n && epochs >= 1 && batch > 0 && batch % 128 == 0 && batch <= m->maxB,
              "tgx_mlp_train: need 0 < n_train <= dataset size and batch a multiple of 128 within max_batch");
  TGX_REQUIRE(m->norm_set, "tgx_mlp_train: set the energy normalisation first");
  tgx_ctx* ctx = m->ctx;
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  std::vector<int> perm((size_t)n_train);
  std::iota(perm.begin(), perm.end(), 0);
  double le_sum = 0.0, lg_sum = 0.0;
  int nb = 0;
  std::vector<int> pidx(1);
  for (int ep = 0; ep < epochs; ++ep) {
    std::vector<int> p((size_t)n_train);
    TGX_TRY(tgx_generate_permutation(n_train, seed + (uint64_t)ep * 7919u, p.data()));
    le_sum = lg_sum = 0.0;
    nb = 0;
    for (int b0 = 0; b0 + batch <= n_train; b0 += batch) {
      TGX_CHECK_HIP(hipMemcpyAsync(m->idx.p, p.data() + b0, sizeof(int) * batch, hipMemcpyHostToDevice, ctx->stream));
      TGX_LAUNCH(ctx, "mlp_gather", k_gather_rows, dim3(batch), dim3(256), 0, batch, m->Dp, m->idx.as<int>(),
                 m->dq.as<float>(), m->dg.as<float>(), m->de.as<double>(), m->bq.as<float>(), m->bqb.as<__bf16>(),
                 m->bg.as<float>(), m->be.as<double>());
      double le = 0.0, lg = 0.0;
      const bool last = b0 + 2 * batch > n_train;
      TGX_TRY(mlp_backward(m, batch, w_e, w_g, (ep == epochs - 1) ? &le : nullptr, (ep == epochs - 1) ? &lg : nullptr));
      TGX_TRY(adam_step(m, lr));
      le_sum += le;
      lg_sum += lg;
      ++nb;
      (void)last;
    }
  }
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (loss_e) *loss_e = nb ? le_sum / nb : 0.0;
  if (loss_g) *loss_g = nb ? lg_sum / nb : 0.0;
  return TGX_OK;
}

// Errors on dataset rows [begin, end): energy MAE / RMSE (eV) and the RMS of
// |P (g_hat - g)| (eV per unit descriptor length).
extern "C" int tgx_mlp_evaluate(tgx_mlp* m, int begin, int end, double* e_mae, double* e_rmse, double* g_rmse) {
  TGX_REQUIRE(m && begin >= 0 && end <= m->n && end > begin, "tgx_mlp_evaluate: bad range");
  tgx_ctx* ctx = m->ctx;
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  double ae = 0.0, se = 0.0, sg = 0.0;
  std::vector<float> E((size_t)m->maxB);
  std::vector<double> er((size_t)m->maxB), lg((size_t)m->maxB);
  for (int b0 = begin; b0 < end; b0 += m->maxB) {
    const int rows = std::min(m->maxB, end - b0);
    const int B = (int)tgx_round_up(rows, 128);
    std::vector<int> idx((size_t)B);
    for (int r = 0; r < B; ++r) idx[r] = b0 + std::min(r, rows - 1);
    TGX_CHECK_HIP(hipMemcpyAsync(m->idx.p, idx.data(), sizeof(int) * B, hipMemcpyHostToDevice, ctx->stream));
    TGX_LAUNCH(ctx, "mlp_gather", k_gather_rows, dim3(B), dim3(256), 0, B, m->Dp, m->idx.as<int>(), m->dq.as<float>(),
               m->dg.as<float>(), m->de.as<double>(), m->bq.as<float>(), m->bqb.as<__bf16>(), m->bg.as<float>(),
               m->be.as<double>());
    TGX_TRY(mlp_forward(m, B, m->bqb.as<__bf16>(), false, m->E.as<float>(), m->G.as<float>(), 1.0));
    double* le = m->lossb.as<double>();
    TGX_LAUNCH(ctx, "mlp_loss", k_loss, dim3((B + 3) / 4), dim3(256), 0, B, m->Dp, m->E.as<float>(), m->be.as<double>(),
               m->G.as<float>(), m->bg.as<float>(), m->bq.as<float>(), m->shift, 1.0 / m->scale, 1.0f,
               m->u.as<float>(), m->ub.as<__bf16>(), m->dE.as<float>(), le, le + B);
    TGX_CHECK_HIP(hipMemcpyAsync(er.data(), le, sizeof(double) * B, hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(lg.data(), le + B, sizeof(double) * B, hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int r = 0; r < rows; ++r) {
      const double d = std::sqrt(er[r]) * m->scale;
      ae += d;
      se += d * d;
      sg += lg[r] * m->scale * m->scale;
    }
  }
  const int n = end - begin;
  if (e_mae) *e_mae = ae / n;
  if (e_rmse) *e_rmse = std::sqrt(se / n);
  if (g_rmse) *g_rmse = std::sqrt(sg / n);
  return TGX_OK;
}

extern "C" int tgx_mlp_forward_host(tgx_mlp* m, int batch, const float* q, float* e, float* g) {
  TGX_REQUIRE(m && q && batch > 0 && batch <= m->maxB, "tgx_mlp_forward_host: bad arguments");
  tgx_ctx* ctx = m->ctx;
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  const int B = (int)tgx_round_up(batch, 128);
  std::vector<float> qp((size_t)B * m->Dp, 0.f);
  for (int r = 0; r < batch; ++r)
    for (int d = 0; d < m->D; ++d) qp[(size_t)r * m->Dp + d] = q[(size_t)r * m->D + d];
  TGX_CHECK_HIP(hipMemcpyAsync(m->bq.p, qp.data(), sizeof(float) * qp.size(), hipMemcpyHostToDevice, ctx->stream));
  TGX_TRY(tgx_f32_to_bf16(ctx, m->bq.as<float>(), m->bqb.as<__bf16>(), (int64_t)B * m->Dp));
  TGX_TRY(mlp_forward(m, B, m->bqb.as<__bf16>(), false, m->E.as<float>(), m->G.as<float>(), 1.0));
  std::vector<float> E((size_t)B), G((size_t)B * m->Dp);
  TGX_CHECK_HIP(hipMemcpyAsync(E.data(), m->E.p, sizeof(float) * B, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(G.data(), m->G.p, sizeof(float) * G.size(), hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < batch; ++r) {
    if (e) e[r] = E[r];
    if (g)
      for (int d = 0; d < m->D; ++d) g[(size_t)r * m->D + d] = G[(size_t)r * m->Dp + d];
  }
  return TGX_OK;
}

// Loss and weight gradients of one batch given on the host (normalised units:
// e_ref and g_ref are already (E - shift)/scale and g/scale when shift = 0 and
// scale = 1 are set). Gradient layout as tgx_mlp_get_weights.
extern "C" int tgx_mlp_gradients_host(tgx_mlp* m, int batch, const float* q, const float* e_ref, const float* g_ref,
                                      float w_e, float w_g, float* dW1, float* db1, float* dW2, float* db2,
                                      float* dw3, float* db3, double* loss_e, double* loss_g) {
  TGX_REQUIRE(m && q && e_ref && g_ref && batch > 0 && batch % 128 == 0 && batch <= m->maxB,
              "tgx_mlp_gradients_host: batch must be a multiple of 128 within max_batch");
  tgx_ctx* ctx = m->ctx;
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  const int B = batch, Dp = m->Dp, D = m->D, H = m->H;
  std::vector<float> qp((size_t)B * Dp, 0.f), gp((size_t)B * Dp, 0.f);
  std::vector<double> ep((size_t)B);
  for (int r = 0; r < B; ++r) {
    ep[r] = e_ref[r];
    for (int d = 0; d < D; ++d) {
      qp[(size_t)r * Dp + d] = q[(size_t)r * D + d];
      gp[(size_t)r * Dp + d] = g_ref[(size_t)r * D + d];
    }
  }
  TGX_CHECK_HIP(hipMemcpyAsync(m->bq.p, qp.data(), sizeof(float) * qp.size(), hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(m->bg.p, gp.data(), sizeof(float) * gp.size(), hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(m->be.p, ep.data(), sizeof(double) * ep.size(), hipMemcpyHostToDevice, ctx->stream));
  TGX_TRY(tgx_f32_to_bf16(ctx, m->bq.as<float>(), m->bqb.as<__bf16>(), (int64_t)B * Dp));
  TGX_TRY(mlp_backward(m, B, w_e, w_g, loss_e, loss_g));
  (void)H;
  (void)D;
  return tgx_mlp_get_gradients(m, dW1, db1, dW2, db2, dw3, db3);
}

// The gradient of the last backward pass, in the layout of tgx_mlp_get_weights.
extern "C" int tgx_mlp_get_gradients(tgx_mlp* m, float* dW1, float* db1, float* dW2, float* db2, float* dw3,
                                     float* db3) {
  TGX_REQUIRE(m, "null network");
  tgx_ctx* ctx = m->ctx;
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  const int H = m->H, D = m->D, Dp = m->Dp;
  std::vector<float> gr((size_t)m->nparam);
  TGX_CHECK_HIP(hipMemcpyAsync(gr.data(), m->grad.p, sizeof(float) * gr.size(), hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (int h = 0; h < H; ++h)
    for (int d = 0; d < D; ++d)
      if (dW1) dW1[(size_t)h * D + d] = gr[m->oW1 + (size_t)h * Dp + d];
  for (int k = 0; k < H; ++k) {
    if (db1) db1[k] = gr[m->ob1 + k];
    if (db2) db2[k] = gr[m->ob2 + k];
    if (dw3) dw3[k] = gr[m->ow3 + k];
  }
  if (dW2) std::memcpy(dW2, gr.data() + m->oW2, sizeof(float) * (size_t)H * H);
  if (db3) db3[0] = gr[m->ob3];
  return TGX_OK;
}

extern "C" int tgx_mlp_adam_step(tgx_mlp* m, float lr) {
  TGX_REQUIRE(m && lr >= 0.0f, "tgx_mlp_adam_step: bad arguments");
  TGX_CHECK_HIP(hipSetDevice(m->ctx->device));
  TGX_TRY(adam_step(m, lr));
  TGX_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  return TGX_OK;
}

// Energy and force loss of whole structures: the centres of every listed
// block (the last FP64 plain evaluation of the current system, which holds
// q, E, -Qss and dq/dr of every centre) form one batch per network. With the
// network gradients g_hat_i = scale G_i and the GAP's g_i = -Qss_i, the force
// residual dF_j = -sum_i (g_hat_i - g_i) . dq_i/dr_j is gathered over all
// blocks (the pair-force kernel and the fixed-order gather), and
//   L = w_e sum_blocks mean_i r_i^2 + w_f / n_atoms sum_j |dF_j|^2,
// r_i the normalised energy residual. dL/dG_i = cG u_i with cG = 2 w_f / n_atoms
// and u_i = -scale sum_p (dq_i/dr_nbr(p))^T dF_nbr(p), which the network's
// forward-then-reverse pass takes like the gradient-loss direction. The weight
// gradients are left in each network (tgx_mlp_get_gradients, tgx_mlp_adam_step);
// the context's pair-force scratch is overwritten.
extern "C" int tgx_mlp_force_gradients(tgx_ctx* ctx, int n_blocks, tgx_mlp* const* nets, const int* model_ids,
                                       double w_e, double w_f, double* loss_e, double* loss_f) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(n_blocks >= 1 && nets && model_ids, "tgx_mlp_force_gradients: bad arguments");
  TGX_REQUIRE(ctx->have_list && ctx->n_sites > 0, "tgx_mlp_force_gradients: no neighbour list");
  std::vector<tgx_model_t*> md((size_t)n_blocks);
  for (int b = 0; b < n_blocks; ++b) {
    tgx_mlp* m = nets[b];
    TGX_REQUIRE(m && m->ctx == ctx, "tgx_mlp_force_gradients: network %d belongs to another context", b);
    TGX_REQUIRE(model_ids[b] >= 0 && model_ids[b] < (int)ctx->models.size(), "model %d not defined", model_ids[b]);
    for (int c = 0; c < b; ++c)
      TGX_REQUIRE(model_ids[c] != model_ids[b], "tgx_mlp_force_gradients: model %d listed twice", model_ids[b]);
    md[b] = ctx->models[model_ids[b]];
    TGX_REQUIRE(md[b]->D == m->D && md[b]->Dpad == m->Dp, "descriptor size of model %d does not match network %d",
                model_ids[b], b);
    TGX_REQUIRE(m->norm_set, "tgx_mlp_force_gradients: set the energy normalisation of network %d first", b);
    TGX_REQUIRE(md[b]->der_whole && md[b]->soap_der.p && md[b]->qss64.p,
                "tgx_mlp_force_gradients: model %d needs a whole-block FP64 plain evaluation with forces first",
                model_ids[b]);
    TGX_REQUIRE(md[b]->n_sites_b <= m->maxB, "model %d has %d centres, more than network %d's max_batch %d",
                model_ids[b], md[b]->n_sites_b, b, m->maxB);
  }
  const int n_at = ctx->n_sites;
  std::vector<int> Bs((size_t)n_blocks);
  // forward pass of every network on its block's centres
  for (int b = 0; b < n_blocks; ++b) {
    tgx_mlp* m = nets[b];
    const int ns = md[b]->n_sites_b;
    const int B = (int)tgx_round_up(ns > 0 ? ns : 1, 128);
    Bs[(size_t)b] = B;
    TGX_LAUNCH(ctx, "mlp_force_batch", k_force_batch, dim3(B), dim3(256), 0, ns, m->Dp, md[b]->q64.as<double>(),
               md[b]->e_site.as<double>(), m->bq.as<float>(), m->bqb.as<__bf16>(), m->be.as<double>());
    TGX_TRY(mlp_forward(m, B, m->bqb.as<__bf16>(), true, m->E.as<float>(), m->G.as<float>(), 1.0));
  }
  // force residual of all blocks together
  const size_t P3 = 3 * (size_t)(ctx->n_pairs > 0 ? ctx->n_pairs : 1);
  TGX_TRY(ctx->fglob.ensure(sizeof(double) * P3));
  TGX_TRY(tgx_zero(ctx, ctx->fglob.p, sizeof(double) * P3));
  for (int b = 0; b < n_blocks; ++b) {
    tgx_mlp* m = nets[b];
    const int ns = md[b]->n_sites_b;
    if (ns == 0) continue;
    const int64_t n = (int64_t)ns * m->Dp;
    TGX_TRY(m->fdg.ensure(sizeof(double) * (size_t)n));
    TGX_LAUNCH(ctx, "mlp_force_dg", k_force_dg, g1(n), dim3(256), 0, n, m->scale, (const float*)m->G.as<float>(),
               (const double*)md[b]->qss64.as<double>(), m->fdg.as<double>());
    TGX_TRY(tgx_plain_pair_forces_q(ctx, md[b], m->fdg.as<double>()));
  }
  TGX_TRY(ctx->fres.ensure(sizeof(double) * 3 * (size_t)n_at));
  TGX_TRY(tgx_gather_pair_forces(ctx, ctx->fglob.as<double>(), ctx->fres.as<double>(), nullptr, nullptr, nullptr));
  // per network: energy residuals, the force direction, the weight gradients
  const float cF = (float)(2.0 * w_f / n_at);
  double se = 0.0;
  for (int b = 0; b < n_blocks; ++b) {
    tgx_mlp* m = nets[b];
    const int ns = md[b]->n_sites_b, B = Bs[(size_t)b];
    double* le = m->lossb.as<double>();
    TGX_LAUNCH(ctx, "mlp_force_eres", k_force_eres, dim3((B + 255) / 256), dim3(256), 0, ns > 0 ? ns : 1, B,
               (const float*)m->E.as<float>(), (const double*)m->be.as<double>(), m->shift, 1.0 / m->scale,
               (float)w_e, m->dE.as<float>(), le);
    TGX_LAUNCH(ctx, "mlp_force_jvp", k_force_jvp, dim3(B), dim3(256), 0, ns, m->D, m->Dp,
               md[b]->site_beg.as<int64_t>(), md[b]->pr_nbr.as<int>(), md[b]->soap_der.as<double>(),
               (const double*)ctx->fres.as<double>(), -m->scale, m->u.as<float>(), m->ub.as<__bf16>());
    TGX_TRY(mlp_backward_dirs(m, B, cF));
    if (loss_e && ns > 0) {
      double e = 0.0;
      TGX_TRY(tgx_device_sum(ctx, le, B, &e));
      se += e / ns;
    }
  }
  if (loss_e) *loss_e = se;
  if (loss_f) {
    std::vector<double> F(3 * (size_t)n_at);
    TGX_CHECK_HIP(hipMemcpyAsync(F.data(), ctx->fres.p, sizeof(double) * F.size(), hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    double sf = 0.0;
    for (double v : F) sf += v * v;
    *loss_f = sf / n_at;
  }
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_mlp_get_weights(tgx_mlp* m, float* W1, float* b1, float* W2, float* b2, float* w3, float* b3) {
  TGX_REQUIRE(m, "null network");
  TGX_CHECK_HIP(hipSetDevice(m->ctx->device));
  std::vector<float> th((size_t)m->nparam);
  TGX_CHECK_HIP(hipMemcpy(th.data(), m->theta.p, sizeof(float) * th.size(), hipMemcpyDeviceToHost));
  const int H = m->H, D = m->D, Dp = m->Dp;
  for (int h = 0; h < H; ++h) {
    if (W1) std::memcpy(W1 + (size_t)h * D, th.data() + m->oW1 + (size_t)h * Dp, sizeof(float) * D);
    if (b1) b1[h] = th[m->ob1 + h];
    if (b2) b2[h] = th[m->ob2 + h];
    if (w3) w3[h] = th[m->ow3 + h];
  }
  if (W2) std::memcpy(W2, th.data() + m->oW2, sizeof(float) * (size_t)H * H);
  if (b3) b3[0] = th[m->ob3];
  return TGX_OK;
}

extern "C" int tgx_mlp_set_weights(tgx_mlp* m, const float* W1, const float* b1, const float* W2, const float* b2,
                                   const float* w3, const float* b3) {
  TGX_REQUIRE(m && W1 && b1 && W2 && b2 && w3 && b3, "tgx_mlp_set_weights: null input");
  tgx_ctx* ctx = m->ctx;
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  std::vector<float> th((size_t)m->nparam, 0.f);
  const int H = m->H, D = m->D, Dp = m->Dp;
  for (int h = 0; h < H; ++h) {
    std::memcpy(th.data() + m->oW1 + (size_t)h * Dp, W1 + (size_t)h * D, sizeof(float) * D);
    th[m->ob1 + h] = b1[h];
    th[m->ob2 + h] = b2[h];
    th[m->ow3 + h] = w3[h];
  }
  std::memcpy(th.data() + m->oW2, W2, sizeof(float) * (size_t)H * H);
  th[m->ob3] = b3[0];
  TGX_CHECK_HIP(hipMemcpyAsync(m->theta.p, th.data(), sizeof(float) * th.size(), hipMemcpyHostToDevice, ctx->stream));
  TGX_TRY(tgx_zero(ctx, m->mom.p, sizeof(float) * m->nparam));
  TGX_TRY(tgx_zero(ctx, m->vel.p, sizeof(float) * m->nparam));
  m->step = 0;
  TGX_TRY(mlp_pack(m));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

// ---- surrogate inside tgx_compute ---------------------------------------------
// Members are networks installed per model; their BF16 weights and biases are
// copied so the network can be destroyed afterwards. Before a run the first K
// members are packed side by side so that one GEMM evaluates every member's
// first layer (W1 rows stacked, [K H][D]) and one GEMM sums every member's
// input gradient (W1^T columns stacked, [D][K H]).
struct SurrogateMember {
  int H = 0;
  double shift = 0, scale = 1;
  dbuf W1b, W1Tb, W2b, W2Tb, bias;  // bias: b1 | b2 | w3 | b3
};
struct EnsemblePack {
  int K = 0, H = 0;
  bool valid = false;
  dbuf W1cat, W1Tcat, W2s, W2Ts;  // BF16: [K H][Dp], [Dp][K H], [K][H][H], [K][H][H]
  dbuf b1, b2, w3, b3, zscale;    // FP32: [K H], [K H], [K H], [K], [K]
  dbuf shift, scale;              // FP64: [K], [K]
  void release() {
    dbuf* all[] = {&W1cat, &W1Tcat, &W2s, &W2Ts, &b1, &b2, &w3, &b3, &zscale, &shift, &scale};
    for (dbuf* b : all) b->release();
    valid = false;
  }
};
struct SurrogateSlot {
  tgx_ctx* ctx = nullptr;
  int model = -1;
  std::vector<SurrogateMember*> members;
  EnsemblePack pack;
};
static std::vector<std::unique_ptr<SurrogateSlot>> g_slots;

static SurrogateSlot& slot(tgx_ctx* ctx, int model_id) {
  for (auto& s : g_slots)
    if (s->ctx == ctx && s->model == model_id) return *s;
  g_slots.emplace_back(new SurrogateSlot());
  g_slots.back()->ctx = ctx;
  g_slots.back()->model = model_id;
  return *g_slots.back();
}

void tgx_surrogate_release(tgx_ctx* ctx) {
  for (auto& s : g_slots)
    if (s->ctx == ctx) {
      for (auto* mb : s->members) {
        if (!mb) continue;
        mb->W1b.release(); mb->W1Tb.release(); mb->W2b.release(); mb->W2Tb.release(); mb->bias.release();
        delete mb;
      }
      s->members.clear();
      s->pack.release();
      s->ctx = nullptr;
    }
}

extern "C" int tgx_surrogate_install(tgx_ctx* ctx, int model_id, int member, tgx_mlp* m) {
  TGX_REQUIRE(ctx && m && m->ctx == ctx, "tgx_surrogate_install: network belongs to another context");
  TGX_CHECK_HIP(hipSetDevice(ctx->device));
  TGX_REQUIRE(model_id >= 0 && model_id < (int)ctx->models.size(), "model %d not defined", model_id);
  tgx_model_t* md = ctx->models[model_id];
  TGX_REQUIRE(md->D == m->D && md->Dpad == m->Dp, "network dimension %d does not match model %d", m->D, model_id);
  TGX_REQUIRE(member >= 0 && member < 16, "member index %d outside 0..15", member);
  SurrogateSlot& sl = slot(ctx, model_id);
  sl.pack.valid = false;
  if ((int)sl.members.size() <= member) sl.members.resize(member + 1, nullptr);
  if (!sl.members[member]) sl.members[member] = new SurrogateMember();
  SurrogateMember* s = sl.members[member];
  const size_t H = m->H, Dp = m->Dp;
  s->H = m->H;
  s->shift = m->shift;
  s->scale = m->scale;
  TGX_TRY(s->W1b.ensure(sizeof(__bf16) * H * Dp));
  TGX_TRY(s->W1Tb.ensure(sizeof(__bf16) * H * Dp));
  TGX_TRY(s->W2b.ensure(sizeof(__bf16) * H * H));
  TGX_TRY(s->W2Tb.ensure(sizeof(__bf16) * H * H));
  TGX_TRY(s->bias.ensure(sizeof(float) * (3 * H + 1)));
  TGX_CHECK_HIP(hipMemcpyAsync(s->W1b.p, m->W1b.p, sizeof(__bf16) * H * Dp, hipMemcpyDeviceToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(s->W1Tb.p, m->W1Tb.p, sizeof(__bf16) * H * Dp, hipMemcpyDeviceToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(s->W2b.p, m->W2b.p, sizeof(__bf16) * H * H, hipMemcpyDeviceToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(s->W2Tb.p, m->W2Tb.p, sizeof(__bf16) * H * H, hipMemcpyDeviceToDevice, ctx->stream));
  const float* th = m->theta.as<float>();
  float* b = s->bias.as<float>();
  TGX_CHECK_HIP(hipMemcpyAsync(b, th + m->ob1, sizeof(float) * H, hipMemcpyDeviceToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(b + H, th + m->ob2, sizeof(float) * H, hipMemcpyDeviceToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(b + 2 * H, th + m->ow3, sizeof(float) * H, hipMemcpyDeviceToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(b + 3 * H, th + m->ob3, sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

// Pack the first K members of a slot (device-to-device copies, once per
// install or change of K).
static int ensemble_pack(tgx_ctx* ctx, SurrogateSlot& sl, int K, int Dp) {
  EnsemblePack& pk = sl.pack;
  if (pk.valid && pk.K == K) return TGX_OK;
  const size_t H = sl.members[0]->H, KH = (size_t)K * H;
  TGX_TRY(pk.W1cat.ensure(sizeof(__bf16) * KH * Dp));
  TGX_TRY(pk.W1Tcat.ensure(sizeof(__bf16) * KH * Dp));
  TGX_TRY(pk.W2s.ensure(sizeof(__bf16) * KH * H));
  TGX_TRY(pk.W2Ts.ensure(sizeof(__bf16) * KH * H));
  TGX_TRY(pk.b1.ensure(sizeof(float) * KH));
  TGX_TRY(pk.b2.ensure(sizeof(float) * KH));
  TGX_TRY(pk.w3.ensure(sizeof(float) * KH));
  TGX_TRY(pk.b3.ensure(sizeof(float) * K));
  TGX_TRY(pk.zscale.ensure(sizeof(float) * K));
  TGX_TRY(pk.shift.ensure(sizeof(double) * K));
  TGX_TRY(pk.scale.ensure(sizeof(double) * K));
  std::vector<float> zs(K);
  std::vector<double> sh(K), sc(K);
  const hipMemcpyKind d2d = hipMemcpyDeviceToDevice;
  for (int k = 0; k < K; ++k) {
    SurrogateMember* s = sl.members[k];
    const float* b = s->bias.as<float>();
    TGX_CHECK_HIP(hipMemcpyAsync(pk.W1cat.as<__bf16>() + k * H * Dp, s->W1b.p, sizeof(__bf16) * H * Dp, d2d, ctx->stream));
    TGX_CHECK_HIP(hipMemcpy2DAsync(pk.W1Tcat.as<__bf16>() + k * H, sizeof(__bf16) * KH, s->W1Tb.p, sizeof(__bf16) * H,
                                   sizeof(__bf16) * H, Dp, d2d, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(pk.W2s.as<__bf16>() + k * H * H, s->W2b.p, sizeof(__bf16) * H * H, d2d, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(pk.W2Ts.as<__bf16>() + k * H * H, s->W2Tb.p, sizeof(__bf16) * H * H, d2d, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(pk.b1.as<float>() + k * H, b, sizeof(float) * H, d2d, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(pk.b2.as<float>() + k * H, b + H, sizeof(float) * H, d2d, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(pk.w3.as<float>() + k * H, b + 2 * H, sizeof(float) * H, d2d, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(pk.b3.as<float>() + k, b + 3 * H, sizeof(float), d2d, ctx->stream));
    // the input gradient of member k enters the sum -(1/K) sum_k scale_k dE_k/dq
    zs[k] = (float)(-s->scale / K);
    sh[k] = s->shift;
    sc[k] = s->scale;
  }
  TGX_CHECK_HIP(hipMemcpyAsync(pk.zscale.p, zs.data(), sizeof(float) * K, hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(pk.shift.p, sh.data(), sizeof(double) * K, hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(pk.scale.p, sc.data(), sizeof(double) * K, hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  pk.K = K;
  pk.H = (int)H;
  pk.valid = true;
  return TGX_OK;
}

extern "C" int tgx_surrogate_gate(tgx_ctx* ctx, int n_members, double threshold) {
  TGX_REQUIRE(ctx && n_members >= 1 && n_members <= 16, "tgx_surrogate_gate: 1..16 members");
  ctx->gate_members = n_members;
  ctx->gate_threshold = threshold > 0.0 ? threshold : 1e30;
  return TGX_OK;
}

extern "C" int tgx_surrogate_stats(tgx_ctx* ctx, double* gated_fraction, double* max_spread) {
  TGX_REQUIRE(ctx, "null context");
  if (gated_fraction) *gated_fraction = ctx->gated_fraction;
  if (max_spread) *max_spread = ctx->max_spread;
  return TGX_OK;
}

// First phase of a block: FP32 descriptors, the ensemble's energies and
// gradients, and the device-side gate marks. tgx_surrogate_finish completes the
// block once the caller has read the number of gated sites.
//   L1    : [h1 | s1] = silu / silu' (q W1cat^T + b1)       one GEMM, N = K H
//   L2    : z2_k = h1_k W2_k^T, fused with its activation:   batched over k
//           a2_k = w3_k silu'(z2_k + b2_k) and the partial sums of
//           w3_k silu(z2_k + b2_k), from which the output kernel forms E_k
//   back2 : a1_k = -(scale_k/K) (a2_k W2_k) o s1_k           batched over k
//   back1 : Qss = a1 W1Tcat^T = -(1/K) sum_k scale_k dE_k/dq  one GEMM, depth K H
int tgx_surrogate_eval(tgx_ctx* ctx, tgx_model_t* m, bool selected) {
  int model_id = -1;
  for (int k = 0; k < (int)ctx->models.size(); ++k)
    if (ctx->models[k] == m) model_id = k;
  SurrogateSlot& sl = slot(ctx, model_id);
  const int K = ctx->gate_members;
  TGX_REQUIRE((int)sl.members.size() >= K, "surrogate: model %d has %zu installed members, %d requested", model_id,
              sl.members.size(), K);
  for (int k = 0; k < K; ++k) TGX_REQUIRE(sl.members[k], "surrogate: member %d of model %d missing", k, model_id);
  const int H = sl.members[0]->H;
  for (int k = 1; k < K; ++k) TGX_REQUIRE(sl.members[k]->H == H, "surrogate members differ in width");
  if (!selected) TGX_TRY(tgx_block_select(ctx, m));
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_TRY(tgx_soap_descriptors(ctx, m, TGX_PATH_MIXED, false));
  const int Mp = (int)tgx_round_up(ns, 128), Dp = m->Dpad;
  TGX_TRY(ensemble_pack(ctx, sl, K, Dp));
  const EnsemblePack& pk = sl.pack;
  const int KH = K * H;
  TGX_REQUIRE(m->qbf.bytes >= sizeof(__bf16) * (size_t)Mp * Dp, "surrogate: BF16 descriptors missing");
  TGX_TRY(m->qss32.ensure(sizeof(float) * (size_t)Mp * Dp));
  TGX_TRY(m->e_site.ensure(sizeof(double) * (size_t)Mp));
  // the block's work space: s1 (FP32 Mp x K H), h1b, a2b, a1b (BF16 Mp x K H), the second
  // layer's energy partials (FP64 K x Mp x H/128)
  const size_t fh = (size_t)Mp * KH;
  const int nt2 = H / 128;
  const size_t bytes = sizeof(float) * fh + sizeof(__bf16) * 3 * fh + sizeof(double) * (size_t)K * Mp * nt2 + 8 * 256;
  TGX_TRY(m->mlp_ws.ensure(bytes));
  char* p = m->mlp_ws.as<char>();
  auto take = [&](size_t n) { char* r = p; p += tgx_round_up((int64_t)n, 256); return r; };
  float* s1 = (float*)take(sizeof(float) * fh);
  __bf16* h1b = (__bf16*)take(sizeof(__bf16) * fh);
  __bf16* a2b = (__bf16*)take(sizeof(__bf16) * fh);
  __bf16* a1b = (__bf16*)take(sizeof(__bf16) * fh);
  double* e2part = (double*)take(sizeof(double) * (size_t)K * Mp * nt2);

  tgx_gemm_desc l1;
  l1.epi = TGX_EPI_SILU;
  l1.M = Mp; l1.N = KH; l1.K = Dp;
  l1.A = m->qbf.as<__bf16>(); l1.lda = Dp;
  l1.B = pk.W1cat.as<__bf16>(); l1.ldb = Dp;
  l1.C = h1b; l1.ldc = KH;
  l1.bias = pk.b1.as<float>();
  l1.aux = s1; l1.ldaux = KH;
  TGX_TRY(tgx_gemm_bf16(ctx, "mlp_l1", l1));

  // second layer with its activation fused: a2 = w3 silu'(z2 + b2) and the
  // per-tile partial sums of w3 silu(z2 + b2)
  tgx_gemm_desc l2;
  l2.epi = TGX_EPI_ACT2;
  l2.M = Mp; l2.N = H; l2.K = H; l2.batch = K;
  l2.A = h1b; l2.lda = KH; l2.sA = H;
  l2.B = pk.W2s.as<__bf16>(); l2.ldb = H; l2.sB = (int64_t)H * H;
  l2.C = a2b; l2.ldc = KH; l2.sC = H;
  l2.bias = pk.b2.as<float>(); l2.sBias = H;
  l2.w3 = pk.w3.as<float>(); l2.sW3 = H;
  l2.epart = e2part; l2.sEpart = (int64_t)Mp * nt2;
  TGX_TRY(tgx_gemm_bf16(ctx, "mlp_l2", l2));

  tgx_gemm_desc b2;
  b2.epi = TGX_EPI_SCALE_AUX;
  b2.M = Mp; b2.N = H; b2.K = H; b2.batch = K;
  b2.A = a2b; b2.lda = KH; b2.sA = H;
  b2.B = pk.W2Ts.as<__bf16>(); b2.ldb = H; b2.sB = (int64_t)H * H;
  b2.C = a1b; b2.ldc = KH; b2.sC = H;
  b2.aux = s1; b2.ldaux = KH; b2.sAux = H;
  b2.zscale = pk.zscale.as<float>();
  TGX_TRY(tgx_gemm_bf16(ctx, "mlp_back2", b2));

  tgx_gemm_desc b1;
  b1.M = Mp; b1.N = Dp; b1.K = KH;
  b1.A = a1b; b1.lda = KH;
  b1.B = pk.W1Tcat.as<__bf16>(); b1.ldb = KH;
  b1.C = m->qss32.p; b1.ldc = Dp;
  TGX_TRY(tgx_gemm_bf16(ctx, "mlp_back1", b1));

  TGX_TRY(m->gate_stat.ensure(sizeof(int) * 2));
  TGX_TRY(m->gate_list.ensure(sizeof(int) * (size_t)Mp));
  TGX_TRY(tgx_zero(ctx, m->gate_stat.p, sizeof(int) * 2));
  TGX_LAUNCH(ctx, "mlp_out", k_surrogate_out, g1(ns), dim3(256), 0, ns, K, (const double*)e2part, Mp, nt2,
             pk.b3.as<float>(), pk.shift.as<double>(),
             pk.scale.as<double>(), m->e_site.as<double>(), (float)ctx->gate_threshold, m->gate_stat.as<int>(),
             m->gate_list.as<int>());
  return TGX_OK;
}

// Second phase of a block: the ng gated sites (listed on the device by
// tgx_surrogate_eval) take the FP64 GAP energy and gradient, then the adjoint
// kernel turns the gradients into pair forces.
int tgx_surrogate_finish(tgx_ctx* ctx, tgx_model_t* m, int ng) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  const int Mp = (int)tgx_round_up(ns, 128), Dp = m->Dpad;
  if (ng > 0) {
    const int ngp = (int)tgx_round_up(ng, 128);
    TGX_TRY(m->q64.ensure(sizeof(double) * (size_t)ngp * Dp));
    // FP64 descriptors and GAP of the gated sites only
    dbuf& pn = m->gate_pn;
    TGX_TRY(pn.ensure(sizeof(double) * (size_t)ngp));
    TGX_TRY(tgx_soap_descriptors_sites(ctx, m, m->gate_list.as<int>(), ng, m->q64.as<double>(), pn.as<double>()));
    dbuf& e_keep = m->gate_e;
    TGX_TRY(e_keep.ensure(sizeof(double) * (size_t)Mp));
    TGX_CHECK_HIP(hipMemcpyAsync(e_keep.p, m->e_site.p, sizeof(double) * ns, hipMemcpyDeviceToDevice, ctx->stream));
    m->n_sites_b = ng;
    int st = tgx_gap_contract(ctx, m, TGX_PATH_FP64_ADJOINT, true);
    m->n_sites_b = ns;
    TGX_TRY(st);
    // e_site now holds the gated sites' GAP energies in its first ng rows
    TGX_LAUNCH(ctx, "gate_replace", k_gate_replace, dim3(ng), dim3(256), 0, ng, m->gate_list.as<int>(), Dp,
               m->e_site.as<double>(), m->qss64.as<double>(), e_keep.as<double>(), m->qss32.as<float>());
    TGX_CHECK_HIP(hipMemcpyAsync(m->e_site.p, e_keep.p, sizeof(double) * ns, hipMemcpyDeviceToDevice, ctx->stream));
  }
  // descriptors the adjoint kernel reads must be the FP32 ones of every site
  return tgx_adjoint_pair_forces(ctx, m, TGX_PATH_SURROGATE);
}
