// Wavefront-per-site SOAP kernels (descriptor and adjoint forces).
//
// One wavefront owns one central site and synchronises only with itself, so a
// workgroup's four wavefronts run four sites independently. Within a site the
// lanes take the neighbour pairs (up to CH per pass):
//   * each lane computes its pair's radial projections R_n (n in the pair's
//     species block) from the Gaussian at the fixed quadrature nodes, and
//     streams all real spherical harmonics from one recurrence in m and l;
//   * the expansion c(n, lm) = sum_j R_n(j) A_l(j) Y_lm(j) is then a small
//     product of the staged R and W = A Y, with lanes over (n, lm) outputs;
//   * the adjoint contracts H(n, lm) = d(g . p)/dc with dc/dr_j on the fly:
//     f_j = sum_lm [ (A' x + A y) Y u + A x (grad Y - l Y u)/r ], with
//     x = sum_n R_n H(n, lm), y = sum_n R_n' H(n, lm), so no per-pair table
//     of harmonics is kept.
// Same conventions and results as the workgroup-per-site kernels (the plain
// dq/dr kernel keeps that form).
#include "tgx_internal.h"

namespace {

constexpr int WAVES = 4;

template <class T> __device__ __forceinline__ T w_exp(T x);
template <> __device__ __forceinline__ double w_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float w_exp<float>(float x) { return __expf(x); }
template <class T> __device__ __forceinline__ T w_pow(T x, T y);
template <> __device__ __forceinline__ double w_pow<double>(double x, double y) { return pow(x, y); }
template <> __device__ __forceinline__ float w_pow<float>(float x, float y) { return powf(x, y); }
template <class T> __device__ __forceinline__ T w_sqrt(T x);
template <> __device__ __forceinline__ double w_sqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float w_sqrt<float>(float x) { return sqrtf(x); }

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T> __device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

struct WaveArgs {
  const int64_t* site_beg;
  const int* pr_spec;
  const double* pr_r;
  const double* pr_xyz;
  const int64_t* pr_gidx;
};

// Per-workgroup tables staged in LDS: harmonic constants, nodes, B transposed to [q][n].
template <class T> struct Tables {
  T* ynorm;  // [NLM]
  T* xq;     // [S][MAX_Q]
  T* Bt;     // [S][MAX_Q][MAX_NB]
  static __host__ __device__ int size(int S, int NLM) { return NLM + S * TGX_MAX_Q + S * TGX_MAX_Q * TGX_MAX_NB; }
  __device__ void load(const SoapParams& P, T* base) {
    ynorm = base;
    xq = ynorm + P.NLM;
    Bt = xq + P.n_species * TGX_MAX_Q;
    for (int i = threadIdx.x; i < P.NLM; i += blockDim.x) ynorm[i] = (T)P.ynorm[i];
    for (int i = threadIdx.x; i < P.n_species * TGX_MAX_Q; i += blockDim.x) xq[i] = (T)P.xq[i];
    for (int i = threadIdx.x; i < P.n_species * TGX_MAX_Q * TGX_MAX_NB; i += blockDim.x) {
      const int s = i / (TGX_MAX_Q * TGX_MAX_NB), rem = i - s * TGX_MAX_Q * TGX_MAX_NB;
      const int q = rem / TGX_MAX_NB, n = rem - q * TGX_MAX_NB;
      Bt[i] = (T)P.B[((size_t)s * TGX_MAX_NB + n) * TGX_MAX_Q + q];
    }
  }
};

// Radial projections of one pair (lane): R[n] and, when GRAD, dR[n]/dr.
template <class T, bool GRAD>
__device__ __forceinline__ void lane_radial(const SoapParams& P, const Tables<T>& tb, int s, T r, T (&R)[TGX_MAX_NB],
                                            T (&dR)[TGX_MAX_NB]) {
  const SpeciesParams& sp = P.sp[s];
  const T inv_rc = (T)sp.inv_rc;
  const T x = r * inv_rc;
  const T sig = (T)sp.sig_r * ((T)1 + (T)sp.sig_r_sc * r) * inv_rc;
  const T dsig = (T)(sp.sig_r * sp.sig_r_sc) * inv_rc;
  const T inv_s2 = (T)1 / (sig * sig);
#pragma unroll
  for (int n = 0; n < TGX_MAX_NB; ++n) { R[n] = 0; dR[n] = 0; }
  const T* xq = tb.xq + s * TGX_MAX_Q;
  const T* Bt = tb.Bt + s * TGX_MAX_Q * TGX_MAX_NB;
  const int nb = sp.nb;
  for (int q = 0; q < sp.nq; ++q) {
    const T dx = xq[q] - x;
    const T g = w_exp<T>(-(T)0.5 * dx * dx * inv_s2);
    T dg = 0;
    if (GRAD) dg = g * (dx * inv_s2 * inv_rc + dx * dx * inv_s2 / sig * dsig);
    const T* b = Bt + q * TGX_MAX_NB;
#pragma unroll
    for (int n = 0; n < TGX_MAX_NB; ++n) {
      if (n < nb) {
        R[n] += b[n] * g;
        if (GRAD) dR[n] += b[n] * dg;
      }
    }
  }
  const T a = (T)sp.amp_a;
  const T Pp = (T)1 + (T)2 * x * x * x - (T)3 * x * x;
  const T dPp = ((T)6 * x * x - (T)6 * x) * inv_rc;
  T amp, damp;
  if (a == (T)0) {
    amp = (T)1 / sig;
    damp = -dsig / (sig * sig);
  } else if (Pp <= (T)1.0e-10) {
    amp = 0;
    damp = 0;
  } else {
    const T Pa = w_pow<T>(Pp, a);
    amp = Pa / sig;
    damp = a * w_pow<T>(Pp, a - (T)1) * dPp / sig - Pa * dsig / (sig * sig);
  }
  amp *= (T)sp.gs;
  damp *= (T)sp.gs;
#pragma unroll
  for (int n = 0; n < TGX_MAX_NB; ++n) {
    const T I = R[n];
    R[n] = amp * I;
    if (GRAD) dR[n] = damp * I + amp * dR[n];
  }
}

// Angular damping recurrence state: A_l = exp(-l(l+1) k).
template <class T> struct Damp {
  T e2, kk, dk;
  __device__ void init(const SpeciesParams& sp, T r) {
    const T st = (T)sp.sig_t * ((T)1 + (T)sp.sig_t_sc * r);
    const T dst = (T)(sp.sig_t * sp.sig_t_sc);
    kk = st * st / ((T)2 * r * r);
    dk = st * st / (r * r * r) - st * dst / (r * r);  // -dk/dr
    const T e1 = w_exp<T>(-kk);
    e2 = e1 * e1;
  }
};

// Streams every (l, m) of the real harmonics at u (and the gradient of the
// solid harmonic when GRAD) to f(l, lm, Y, gx, gy, gz), m-major.
template <class T, bool GRAD, class F>
__device__ __forceinline__ void ylm_stream(int L, const T* ynorm, T ux, T uy, T uz, F&& f) {
  T c = 1, cx = 0, cy = 0, s = 0, sx = 0, sy = 0;
  for (int m = 0; m <= L; ++m) {
    if (m > 0) {
      const T nc = c * ux - s * uy, ns = c * uy + s * ux;
      if (GRAD) {
        const T ncx = cx * ux + c - sx * uy, ncy = cy * ux - sy * uy - s;
        const T nsx = cx * uy + sx * ux + s, nsy = cy * uy + c + sy * ux;
        cx = ncx; cy = ncy; sx = nsx; sy = nsy;
      }
      c = nc;
      s = ns;
    }
    T q = 1, qx = 0, qy = 0, qz = 0, p = 0, px = 0, py = 0, pz = 0;
    for (int l = m; l <= L; ++l) {
      if (l == m + 1) {
        p = q; px = qx; py = qy; pz = qz;
        q = (T)(2 * m + 1) * uz;
        qx = 0; qy = 0; qz = (T)(2 * m + 1);
      } else if (l > m + 1) {
        const T a = (T)(2 * l - 1), b = (T)(l + m - 1), inv = (T)1 / (T)(l - m);
        const T nq = (a * uz * q - b * p) * inv;
        if (GRAD) {
          const T nqx = (a * uz * qx - b * ((T)2 * ux * p + px)) * inv;
          const T nqy = (a * uz * qy - b * ((T)2 * uy * p + py)) * inv;
          const T nqz = (a * (q + uz * qz) - b * ((T)2 * uz * p + pz)) * inv;
          px = qx; py = qy; pz = qz;
          qx = nqx; qy = nqy; qz = nqz;
        }
        p = q;
        q = nq;
      }
      const int base = l * l + l;
      const T N = ynorm[base + m];
      if (m == 0) {
        f(l, base, N * q, N * qx, N * qy, N * qz);
      } else {
        f(l, base + m, N * q * c, N * (qx * c + q * cx), N * (qy * c + q * cy), N * qz * c);
        f(l, base - m, N * q * s, N * (qx * s + q * sx), N * (qy * s + q * sy), N * qz * s);
      }
    }
  }
}

// Expansion coefficients of one site into C (LDS, [N][NLM]) by one wavefront.
// Pass over chunks of CH pairs: pair lanes stage Rf[j][n] (stacked radial index,
// zero outside the pair's species block) and W[j][lm] = A_l Y_lm in LDS; then
// lanes over lm accumulate C[n][lm] += sum_j Rf[j][n] W[j][lm] with the n
// column in registers and Rf rows read as LDS broadcasts.
template <class T, int CH, int NMAX>
__device__ void wave_coefficients(const SoapParams& P, const Tables<T>& tb, const WaveArgs& a, int64_t b0, int nn, T* C,
                                  T* W, T* Rf) {
  const int lane = threadIdx.x & 63;
  const int NLM = P.NLM, N = P.N, L = P.L;
  T acc[2][NMAX];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int n = 0; n < NMAX; ++n) acc[h][n] = 0;
  {
    const int s = a.pr_spec[b0];
    if (s >= 0) {
      // central atom: l = m = 0 of its own block
#pragma unroll
      for (int n = 0; n < NMAX; ++n) {
        const int nl = n - P.sp[s].off;
        if (lane == 0 && nl >= 0 && nl < P.sp[s].nb) acc[0][n] = (T)P.cen[s * TGX_MAX_NB + nl];
      }
    }
  }
  for (int k0 = 1; k0 < nn; k0 += CH) {
    const int cnt = min(CH, nn - k0);
    wave_sync();
    if (lane < cnt) {
      const int64_t p = b0 + k0 + lane;
      int s = a.pr_spec[p];
      const double rr = a.pr_r[p];
      if (s >= 0 && !(rr < P.sp[s].rc)) s = -1;
      T* rf = Rf + lane * NMAX;
#pragma unroll
      for (int n = 0; n < NMAX; ++n) rf[n] = 0;
      if (s >= 0) {
        const T r = (T)rr, inv = (T)1 / r;
        const T ux = (T)a.pr_xyz[3 * p] * inv, uy = (T)a.pr_xyz[3 * p + 1] * inv, uz = (T)a.pr_xyz[3 * p + 2] * inv;
        T R[TGX_MAX_NB], dR[TGX_MAX_NB];
        lane_radial<T, false>(P, tb, s, r, R, dR);
        const int off = P.sp[s].off, nb = P.sp[s].nb;
#pragma unroll
        for (int n = 0; n < TGX_MAX_NB; ++n)
          if (n < nb) rf[off + n] = R[n];
        Damp<T> dmp;
        dmp.init(P.sp[s], r);
        T Atab[TGX_MAX_L + 1];
        T Al = 1, pw = 1;
#pragma unroll
        for (int l = 0; l <= TGX_MAX_L; ++l) {
          if (l > 0) { pw *= dmp.e2; Al *= pw; }
          Atab[l] = Al;
        }
        T* w = W + lane * NLM;
        ylm_stream<T, false>(L, tb.ynorm, ux, uy, uz, [&](int l, int lm, T Y, T, T, T) { w[lm] = Atab[l] * Y; });
      } else {
        T* w = W + lane * NLM;
        for (int lm = 0; lm < NLM; ++lm) w[lm] = 0;
      }
    }
    wave_sync();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lm = lane + 64 * h;
      if (lm < NLM) {
        for (int j = 0; j < cnt; ++j) {
          const T wv = W[j * NLM + lm];
          const T* rf = Rf + j * NMAX;
#pragma unroll
          for (int n = 0; n < NMAX; ++n) acc[h][n] += rf[n] * wv;
        }
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lm = lane + 64 * h;
    if (lm < NLM)
#pragma unroll
      for (int n = 0; n < NMAX; ++n)
        if (n < N) C[n * NLM + lm] = acc[h][n];
  }
  wave_sync();
}

// Power spectrum of C into X[0..D), returns |p| (X normalised to q in place).
template <class T>
__device__ T wave_power(const SoapParams& P, const T* C, T* X) {
  const int lane = threadIdx.x & 63;
  const int NLM = P.NLM;
  T acc = 0;
  for (int d = lane; d < P.D; d += 64) {
    const int code = P.dtab[d];
    const int n = code & 0xff, np = (code >> 8) & 0xff, l = code >> 16;
    const T* c1 = C + n * NLM + l * l;
    const T* c2 = C + np * NLM + l * l;
    T s0 = 0, s1 = 0;
#pragma unroll
    for (int m = 0; m < 2 * TGX_MAX_L + 1; m += 2) {
      if (m < 2 * l + 1) s0 += c1[m] * c2[m];
      if (m + 1 < 2 * l + 1) s1 += c1[m + 1] * c2[m + 1];
    }
    T s = s0 + s1;
    if (n != np) s *= (T)M_SQRT2;
    X[d] = s;
    acc += s * s;
  }
  const T nrm = w_sqrt<T>(wave_sum<T>(acc));
  const T inv = nrm > (T)0 ? (T)1 / nrm : (T)0;
  for (int d = lane; d < P.D; d += 64) X[d] *= inv;
  wave_sync();
  return nrm;
}

template <class T> struct WaveLayout {
  // per-wavefront LDS (elements of T):
  //   descriptor: C [N NLM] | U = max(W [CH NLM] + Rf [CH NMAX], X [D])
  //   adjoint:    C | H [N NLM] | U = max(W + Rf, X, pair chunk - N NLM)
  //   where the pair chunk (Xa, Xb [CH NLM], R, dR [CH NMAX]) overlays C and U.
  static __host__ __device__ int coef_chunk(int CH, int NLM, int NMAX) { return CH * NLM + CH * NMAX; }
  static __host__ __device__ int pair_chunk(int CH, int NLM, int NMAX) { return 2 * CH * NLM + 2 * CH * NMAX; }
  static __host__ __device__ int per_wave(int CH, int NMAX, const SoapParams& P, bool adjoint) {
    const int cn = P.N * P.NLM;
    int u = coef_chunk(CH, P.NLM, NMAX);
    if (P.D > u) u = P.D;
    if (!adjoint) return cn + u;
    const int pc = pair_chunk(CH, P.NLM, NMAX) - cn;
    if (pc > u) u = pc;
    return 2 * cn + u;
  }
};

template <class T, int CH, int NMAX>
__global__ __launch_bounds__(64 * WAVES) void k_desc_wave(SoapParams P, WaveArgs a, int n_sites, T* __restrict__ q_out,
                                                          T* __restrict__ pnorm_out, const int* __restrict__ site_list) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  Tables<T> tb;
  tb.load(P, sm);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int orow = blockIdx.x * WAVES + w;
  if (orow >= n_sites) return;
  const int site = site_list ? site_list[orow] : orow;
  T* base = sm + Tables<T>::size(P.n_species, P.NLM) + w * WaveLayout<T>::per_wave(CH, NMAX, P, false);
  T* C = base;
  T* U = C + P.N * P.NLM;
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  wave_coefficients<T, CH, NMAX>(P, tb, a, b0, nn, C, U, U + CH * P.NLM);
  T* X = U;  // reuses the chunk scratch
  const T nrm = wave_power<T>(P, C, X);
  T* qrow = q_out + (size_t)orow * P.Dpad;
  for (int d = lane; d < P.Dpad; d += 64) qrow[d] = d < P.D ? X[d] : (T)0;
  if (lane == 0) pnorm_out[orow] = nrm;
}

template <class T, int CH, int NMAX>
__global__ __launch_bounds__(64 * WAVES) void k_adjoint_wave(SoapParams P, WaveArgs a, int n_sites,
                                                             const T* __restrict__ q_in, const T* __restrict__ pnorm_in,
                                                             const T* __restrict__ qss, double* __restrict__ fglob) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  Tables<T> tb;
  tb.load(P, sm);
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int site = blockIdx.x * WAVES + w;
  if (site >= n_sites) return;
  const int NLM = P.NLM, N = P.N, L = P.L, D = P.D, CN = N * NLM;
  T* base = sm + Tables<T>::size(P.n_species, NLM) + w * WaveLayout<T>::per_wave(CH, NMAX, P, true);
  T* C = base;
  T* H = C + CN;
  T* U = H + CN;
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  wave_coefficients<T, CH, NMAX>(P, tb, a, b0, nn, C, U, U + CH * NLM);
  // g = (Qss - (Qss . q) q) / |p|
  T* X = U;
  const T* qrow = q_in + (size_t)site * P.Dpad;
  const T* srow = qss + (size_t)site * P.Dpad;
  T acc = 0;
  for (int d = lane; d < D; d += 64) acc += srow[d] * qrow[d];
  const T sq = wave_sum<T>(acc);
  const T nrm = pnorm_in[site];
  const T inv_n = nrm > (T)0 ? (T)1 / nrm : (T)0;
  for (int d = lane; d < D; d += 64) X[d] = (srow[d] - sq * qrow[d]) * inv_n;
  wave_sync();
  // H(n, lm) = sum_n' S(n, n', l) C(n', lm): lanes over lm, C column in registers
  const int L1 = L + 1;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lm = lane + 64 * h;
    if (lm < NLM) {
      int l = (int)sqrtf((float)lm + 0.5f);
      if (l * l > lm) --l;
      T cc[NMAX];
#pragma unroll
      for (int n = 0; n < NMAX; ++n) cc[n] = n < N ? C[n * NLM + lm] : (T)0;
#pragma unroll
      for (int n = 0; n < NMAX; ++n) {
        if (n >= N) break;
        T hsum = 0;
#pragma unroll
        for (int np = 0; np < NMAX; ++np) {
          if (np >= N) break;
          const int lo = n < np ? n : np, hi = n < np ? np : n;
          const int pidx = lo * N - lo * (lo - 1) / 2 + (hi - lo);
          hsum += (n == np ? (T)2 : (T)M_SQRT2) * X[pidx * L1 + l] * cc[np];
        }
        H[n * NLM + lm] = hsum;
      }
    }
  }
  wave_sync();
  // pair chunks: Xa[j][lm] = sum_n R_n(j) H(n, lm), Xb[j][lm] = sum_n R_n'(j) H(n, lm)
  T* Xa = C;                       // overlays C and the start of U (C is dead now)
  T* Xb = Xa + CH * NLM;
  T* Rr = Xb + CH * NLM;           // [CH][NMAX] stacked radial values, zero outside the block
  T* dRr = Rr + CH * NMAX;
  T hreg[2][NMAX];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lm = lane + 64 * h;
#pragma unroll
    for (int n = 0; n < NMAX; ++n) hreg[h][n] = (lm < NLM && n < N) ? H[n * NLM + lm] : (T)0;
  }
  T s0 = 0, s1 = 0, s2 = 0;
  for (int k0 = 1; k0 < nn; k0 += CH) {
    const int cnt = min(CH, nn - k0);
    // 1) pair lanes: radial values and derivatives
    int s = -1;
    T r = 0, ux = 0, uy = 0, uz = 0;
    int64_t p = 0;
    wave_sync();
    if (lane < cnt) {
      p = b0 + k0 + lane;
      s = a.pr_spec[p];
      const double rr = a.pr_r[p];
      if (s >= 0 && !(rr < P.sp[s].rc)) s = -1;
      T* rr_ = Rr + lane * NMAX;
      T* dr_ = dRr + lane * NMAX;
#pragma unroll
      for (int n = 0; n < NMAX; ++n) { rr_[n] = 0; dr_[n] = 0; }
      if (s >= 0) {
        r = (T)rr;
        const T inv = (T)1 / r;
        ux = (T)a.pr_xyz[3 * p] * inv; uy = (T)a.pr_xyz[3 * p + 1] * inv; uz = (T)a.pr_xyz[3 * p + 2] * inv;
        T R[TGX_MAX_NB], dR[TGX_MAX_NB];
        lane_radial<T, true>(P, tb, s, r, R, dR);
        const int off = P.sp[s].off, nb = P.sp[s].nb;
#pragma unroll
        for (int n = 0; n < TGX_MAX_NB; ++n)
          if (n < nb) { rr_[off + n] = R[n]; dr_[off + n] = dR[n]; }
      }
    }
    wave_sync();
    // 2) lm lanes: Xa, Xb for every pair of the chunk
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lm = lane + 64 * h;
      if (lm < NLM)
        for (int j = 0; j < cnt; ++j) {
          const T* rj = Rr + j * NMAX;
          const T* dj = dRr + j * NMAX;
          T xa = 0, xb = 0;
#pragma unroll
          for (int n = 0; n < NMAX; ++n) {
            xa += rj[n] * hreg[h][n];
            xb += dj[n] * hreg[h][n];
          }
          Xa[j * NLM + lm] = xa;
          Xb[j * NLM + lm] = xb;
        }
    }
    wave_sync();
    // 3) pair lanes: contract with the harmonics and their gradients
    T f0 = 0, f1 = 0, f2 = 0;
    if (lane < cnt && s >= 0) {
      const T inv_r = (T)1 / r;
      Damp<T> dmp;
      dmp.init(P.sp[s], r);
      T Atab[TGX_MAX_L + 1], dAtab[TGX_MAX_L + 1];
      T Al = 1, pw = 1;
#pragma unroll
      for (int l = 0; l <= TGX_MAX_L; ++l) {
        if (l > 0) { pw *= dmp.e2; Al *= pw; }
        Atab[l] = Al;
        dAtab[l] = Al * (T)(l * (l + 1)) * dmp.dk;
      }
      const T* xa = Xa + lane * NLM;
      const T* xb = Xb + lane * NLM;
      ylm_stream<T, true>(L, tb.ynorm, ux, uy, uz, [&](int l, int lm, T Y, T gx, T gy, T gz) {
        const T x = xa[lm], y = xb[lm];
        const T A = Atab[l], dA = dAtab[l];
        const T rad = dA * x + A * y;
        const T ang = A * x * inv_r;
        const T lf = (T)l;
        f0 += rad * Y * ux + ang * (gx - lf * Y * ux);
        f1 += rad * Y * uy + ang * (gy - lf * Y * uy);
        f2 += rad * Y * uz + ang * (gz - lf * Y * uz);
      });
    }
    if (lane < cnt) {
      const int64_t gi = a.pr_gidx[p];
      fglob[3 * gi] += (double)f0;
      fglob[3 * gi + 1] += (double)f1;
      fglob[3 * gi + 2] += (double)f2;
    }
    s0 += f0;
    s1 += f1;
    s2 += f2;
  }
  const double d0 = wave_sum<double>((double)s0), d1 = wave_sum<double>((double)s1), d2 = wave_sum<double>((double)s2);
  if (lane == 0) {
    const int64_t gi = a.pr_gidx[b0];
    fglob[3 * gi] -= d0;
    fglob[3 * gi + 1] -= d1;
    fglob[3 * gi + 2] -= d2;
  }
}

}  // namespace

static WaveArgs wave_args(tgx_model_t* m) {
  WaveArgs a;
  a.site_beg = m->site_beg.as<int64_t>();
  a.pr_spec = m->pr_spec.as<int>();
  a.pr_r = m->pr_r.as<double>();
  a.pr_xyz = m->pr_xyz.as<double>();
  a.pr_gidx = m->pr_gidx.as<int64_t>();
  return a;
}

template <class T> static size_t wave_lds(const SoapParams& P, int CH, int NMAX, bool adjoint) {
  return sizeof(T) * ((size_t)Tables<T>::size(P.n_species, P.NLM) +
                      (size_t)WAVES * WaveLayout<T>::per_wave(CH, NMAX, P, adjoint)) + 64;
}

// CH pairs per pass: 16. NMAX (stacked radial functions held per lane) is the
// smallest of 8, 16, 32 that covers N.
template <class T, int NMAX>
static int launch_desc(tgx_ctx* ctx, tgx_model_t* m, const int* d_sites, int n_rows, void* q_out, void* pn_out,
                       const char* name) {
  WaveArgs a = wave_args(m);
  const size_t lds = wave_lds<T>(m->P, 16, NMAX, false);
  TGX_REQUIRE(lds <= 160 * 1024, "descriptor kernel needs %zu bytes of LDS", lds);
  TGX_LAUNCH(ctx, name, (k_desc_wave<T, 16, NMAX>), dim3((n_rows + WAVES - 1) / WAVES), dim3(64 * WAVES), lds, m->P,
             a, n_rows, (T*)q_out, (T*)pn_out, d_sites);
  return TGX_OK;
}

template <class T, int NMAX>
static int launch_adj(tgx_ctx* ctx, tgx_model_t* m, const T* q, const T* pn, const T* qss, const char* name) {
  WaveArgs a = wave_args(m);
  const int ns = m->n_sites_b;
  const size_t lds = wave_lds<T>(m->P, 16, NMAX, true);
  TGX_REQUIRE(lds <= 160 * 1024, "adjoint kernel needs %zu bytes of LDS", lds);
  TGX_LAUNCH(ctx, name, (k_adjoint_wave<T, 16, NMAX>), dim3((ns + WAVES - 1) / WAVES), dim3(64 * WAVES), lds, m->P, a,
             ns, q, pn, qss, ctx->fglob.as<double>());
  return TGX_OK;
}

int tgx_desc_wave(tgx_ctx* ctx, tgx_model_t* m, bool fp64, const int* d_sites, int n_rows, void* q_out, void* pn_out) {
  if (n_rows == 0) return TGX_OK;
  const int N = m->P.N;
  if (fp64) {
    const char* nm = d_sites ? "soap_desc_list_f64" : "soap_desc_f64";
    if (N <= 8) return launch_desc<double, 8>(ctx, m, d_sites, n_rows, q_out, pn_out, nm);
    if (N <= 16) return launch_desc<double, 16>(ctx, m, d_sites, n_rows, q_out, pn_out, nm);
    return launch_desc<double, 32>(ctx, m, d_sites, n_rows, q_out, pn_out, nm);
  }
  if (N <= 8) return launch_desc<float, 8>(ctx, m, d_sites, n_rows, q_out, pn_out, "soap_desc_f32");
  if (N <= 16) return launch_desc<float, 16>(ctx, m, d_sites, n_rows, q_out, pn_out, "soap_desc_f32");
  return launch_desc<float, 32>(ctx, m, d_sites, n_rows, q_out, pn_out, "soap_desc_f32");
}

int tgx_adjoint_wave(tgx_ctx* ctx, tgx_model_t* m, bool fp64) {
  if (m->n_sites_b == 0) return TGX_OK;
  const int N = m->P.N;
  if (fp64) {
    // a compressed block contracts the full-size gradient with the full descriptors
    const double* q = (m->compress ? m->qf64 : m->q64).as<double>();
    const double* g = (m->compress ? m->qssf64 : m->qss64).as<double>();
    const double* pn = m->pnorm.as<double>();
    if (N <= 8) return launch_adj<double, 8>(ctx, m, q, pn, g, "soap_adjoint_f64");
    if (N <= 16) return launch_adj<double, 16>(ctx, m, q, pn, g, "soap_adjoint_f64");
    return launch_adj<double, 32>(ctx, m, q, pn, g, "soap_adjoint_f64");
  }
  const float *q = m->q32.as<float>(), *pn = m->pnorm.as<float>(), *g = m->qss32.as<float>();
  if (N <= 8) return launch_adj<float, 8>(ctx, m, q, pn, g, "soap_adjoint_f32");
  if (N <= 16) return launch_adj<float, 16>(ctx, m, q, pn, g, "soap_adjoint_f32");
  return launch_adj<float, 32>(ctx, m, q, pn, g, "soap_adjoint_f32");
}
