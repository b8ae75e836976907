// SOAP descriptor kernels: radial (poly3/poly3gauss by fixed Gauss-Legendre
// projection) and real-spherical-harmonic expansion, power spectrum,
// normalisation, and the two force routes:
//   plain   — dq_i/dr_j written for every pair, (3, D, P) like soap_cart_der
//             (src/gap_interface.f90:312-317), then contracted by k_plain_force;
//   adjoint — dE/dq back-propagated to dE/dc_nlm per site and contracted with
//             dc_nlm/dr_j recomputed per pair; dq/dr is never stored.
// The conventions (basis, density, amplitude, angular damping, ordering,
// multiplicity sqrt(2), normalisation, self pair = -sum) are the ones the
// FP64 oracle states; soap_turbo's own source is not available to follow.
//
// One workgroup (256 threads, 4 wavefronts) per central site. The site's
// expansion coefficients c(n, lm) live in LDS; lanes of a wavefront split the
// quadrature nodes, the (l, m) harmonics and the (n, l) channel sums of a pair;
// the power spectrum and its derivative are split over threads by component.
#include "tgx_internal.h"

int tgx_desc_wave(tgx_ctx* ctx, tgx_model_t* m, bool fp64, const int* d_sites, int n_rows, void* q_out, void* pn_out);
int tgx_adjoint_wave(tgx_ctx* ctx, tgx_model_t* m, bool fp64);
int tgx_desc_f32(tgx_ctx* ctx, tgx_model_t* m);
int tgx_adjoint_f32(tgx_ctx* ctx, tgx_model_t* m);

namespace {

constexpr int WG = 256;

template <class T> __device__ __forceinline__ T t_exp(T x);
template <> __device__ __forceinline__ double t_exp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float t_exp<float>(float x) { return __expf(x); }
template <class T> __device__ __forceinline__ T t_sqrt(T x);
template <> __device__ __forceinline__ double t_sqrt<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float t_sqrt<float>(float x) { return sqrtf(x); }
template <class T> __device__ __forceinline__ T t_pow(T x, T y);
template <> __device__ __forceinline__ double t_pow<double>(double x, double y) { return pow(x, y); }
template <> __device__ __forceinline__ float t_pow<float>(float x, float y) { return powf(x, y); }

// Real spherical harmonic Y_lm(u) (orthonormal, Condon-Shortley phase) and the
// gradient of the homogeneous solid harmonic r^l Y_lm at the unit vector u.
// nrm is the tabulated constant sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!)
// (-1)^|m| (2|m|-1)!! (times sqrt 2 for m != 0).
template <class T>
__device__ void ylm_one(int l, int m, T nrm, T ux, T uy, T uz, T& Y, T& gx, T& gy, T& gz) {
  const int am = m < 0 ? -m : m;
  // (x + i y)^am and its x, y derivatives
  T c = 1, cx = 0, cy = 0, s = 0, sx = 0, sy = 0;
  for (int k = 0; k < am; ++k) {
    T nc = c * ux - s * uy, ncx = cx * ux + c - sx * uy, ncy = cy * ux - sy * uy - s;
    T ns = c * uy + s * ux, nsx = cx * uy + sx * ux + s, nsy = cy * uy + c + sy * ux;
    c = nc; cx = ncx; cy = ncy; s = ns; sx = nsx; sy = nsy;
  }
  // Q_l^am(z, r^2) with Q_am^am = 1 (the constant is in nrm)
  T q = 1, qx = 0, qy = 0, qz = 0;        // Q_{l'}
  T p = 0, px = 0, py = 0, pz = 0;        // Q_{l'-1}
  if (l > am) {
    p = q; px = qx; py = qy; pz = qz;
    q = (T)(2 * am + 1) * uz; qx = 0; qy = 0; qz = (T)(2 * am + 1);
    for (int lp = am + 2; lp <= l; ++lp) {
      const T a = (T)(2 * lp - 1), b = (T)(lp + am - 1), inv = (T)1 / (T)(lp - am);
      // r^2 = 1 at the unit vector, d(r^2) = 2u
      T nq = (a * uz * q - b * p) * inv;
      T nqx = (a * uz * qx - b * ((T)2 * ux * p + px)) * inv;
      T nqy = (a * uz * qy - b * ((T)2 * uy * p + py)) * inv;
      T nqz = (a * (q + uz * qz) - b * ((T)2 * uz * p + pz)) * inv;
      p = q; px = qx; py = qy; pz = qz;
      q = nq; qx = nqx; qy = nqy; qz = nqz;
    }
  }
  const T N = nrm;
  if (m == 0) {
    Y = N * q; gx = N * qx; gy = N * qy; gz = N * qz;
  } else if (m > 0) {
    Y = N * q * c; gx = N * (qx * c + q * cx); gy = N * (qy * c + q * cy); gz = N * qz * c;
  } else {
    Y = N * q * s; gx = N * (qx * s + q * sx); gy = N * (qy * s + q * sy); gz = N * qz * s;
  }
}

__device__ __forceinline__ int lm_l(int lm) {
  int l = (int)sqrtf((float)lm + 0.5f);
  return l * l > lm ? l - 1 : l;
}

// Scratch of one pair in LDS.
template <class T> struct PairScratch {
  T* G;   // [MAX_Q]
  T* dG;  // [MAX_Q]
  T* R;   // [MAX_NB]
  T* dR;  // [MAX_NB]
  T* A;   // [L+1]
  T* dA;  // [L+1]
  T* Y;   // [NLM]
  T* gY;  // [3 NLM]  gradient of the solid harmonic
  static constexpr int size(int L, int NLM) { return 2 * TGX_MAX_Q + 2 * TGX_MAX_NB + 2 * (L + 1) + 4 * NLM; }
  __device__ void bind(T* base, int L, int NLM) {
    G = base; dG = G + TGX_MAX_Q; R = dG + TGX_MAX_Q; dR = R + TGX_MAX_NB; A = dR + TGX_MAX_NB; dA = A + (L + 1);
    Y = dA + (L + 1); gY = Y + NLM;
  }
};

// Gaussians at the quadrature nodes (threads t < nq of nthr) and the angular
// damping A_l (thread nthr - 1) of one pair.
template <class T>
__device__ __forceinline__ void pair_gauss(const SoapParams& P, int s, T r, int t, int nthr, PairScratch<T>& ps,
                                           bool need_grad) {
  const SpeciesParams& sp = P.sp[s];
  const T inv_rc = (T)sp.inv_rc;
  const T x = r * inv_rc;
  const T sig = (T)sp.sig_r * ((T)1 + (T)sp.sig_r_sc * r) * inv_rc;
  const T dsig = (T)(sp.sig_r * sp.sig_r_sc) * inv_rc;
  const T inv_s2 = (T)1 / (sig * sig);
  const double* xq = P.xq + s * TGX_MAX_Q;
  for (int q = t; q < sp.nq; q += nthr) {
    const T dx = (T)xq[q] - x;
    const T g = t_exp<T>(-(T)0.5 * dx * dx * inv_s2);
    ps.G[q] = g;
    if (need_grad) ps.dG[q] = g * (dx * inv_s2 * inv_rc + dx * dx * inv_s2 / sig * dsig);
  }
  if (t == nthr - 1) {
    // A_l = exp(-l(l+1) k), k = st^2/(2 r^2): one exponential, A_l = A_{l-1} e^(2l) with e = exp(-k)
    const T st = (T)sp.sig_t * ((T)1 + (T)sp.sig_t_sc * r);
    const T dst = (T)(sp.sig_t * sp.sig_t_sc);
    const T kk = st * st / ((T)2 * r * r);
    const T dk = st * st / (r * r * r) - st * dst / (r * r);  // -dk/dr
    const T e1 = t_exp<T>(-kk), e2 = e1 * e1;
    T Al = 1, pw = 1;
    for (int l = 0; l <= P.L; ++l) {
      if (l > 0) { pw *= e2; Al *= pw; }
      const T ll = (T)(l * (l + 1));
      ps.A[l] = Al;
      ps.dA[l] = Al * ll * dk;
    }
  }
}

// Step 1 of a pair, by `nthr` threads with local index t: Gaussians at the
// quadrature nodes, harmonics, angular damping.
template <class T>
__device__ void pair_step1(const SoapParams& P, int s, T r, T ux, T uy, T uz, int t, int nthr, PairScratch<T>& ps,
                           bool need_grad) {
  pair_gauss<T>(P, s, r, t, nthr, ps, true);
  for (int lm = t; lm < P.NLM; lm += nthr) {
    const int l = lm_l(lm), m = lm - l * l - l;
    T Y, gx, gy, gz;
    ylm_one<T>(l, m, (T)P.ynorm[lm], ux, uy, uz, Y, gx, gy, gz);
    ps.Y[lm] = Y;
    if (need_grad) { ps.gY[3 * lm] = gx; ps.gY[3 * lm + 1] = gy; ps.gY[3 * lm + 2] = gz; }
  }
}

// Real harmonics of one pair by lanes over m, for a compile-time l_max LL:
// lane t = m + LL (t <= 2 LL) runs the l recurrence of its |m| and writes
// Y_lm, and with GRAD the gradient of the solid harmonic r^l Y_lm at the unit
// vector (xyz interleaved), for l = |m|..LL. Same recurrences and constants
// as ylm_one; rinv[k] = 1/k replaces the division. With PACK (and GRAD) the
// four values of lm go to Y[4 lm .. 4 lm + 3] = (Y, gx, gy, gz) instead.
template <int LL, bool GRAD, bool PACK = false>
__device__ __forceinline__ void harmonics_m(int t, const double* ynorm, const double* rinv, double ux, double uy,
                                            double uz, double* Y, double* gY) {
  static_assert(!PACK || GRAD, "the packed layout carries gradients");
  if (t > 2 * LL) return;
  // opaque to the optimiser: the per-lane recurrence constants are rebuilt per call, not hoisted
  // out of the caller's pair loop where they would stay live across the whole site
  asm volatile("" : "+v"(t));
  const int m = t - LL, am = m < 0 ? -m : m;
  double c = 1, cx = 0, cy = 0, s = 0, sx = 0, sy = 0;
#pragma unroll
  for (int k = 0; k < LL; ++k) {
    if (k < am) {
      const double nc = c * ux - s * uy, ns = c * uy + s * ux;
      if (GRAD) {
        const double ncx = cx * ux + c - sx * uy, ncy = cy * ux - sy * uy - s;
        const double nsx = cx * uy + sx * ux + s, nsy = cy * uy + c + sy * ux;
        cx = ncx; cy = ncy; sx = nsx; sy = nsy;
      }
      c = nc;
      s = ns;
    }
  }
  const double f = m < 0 ? s : c, fx = m < 0 ? sx : cx, fy = m < 0 ? sy : cy;
  double q = 1, qx = 0, qy = 0, qz = 0, p = 0, px = 0, py = 0, pz = 0;
#pragma unroll
  for (int l = 0; l <= LL; ++l) {
    if (l < am) continue;
    if (l == am + 1) {
      p = q; px = qx; py = qy; pz = qz;
      q = (double)(2 * am + 1) * uz;
      qx = 0; qy = 0; qz = (double)(2 * am + 1);
    } else if (l > am + 1) {
      const double a = (double)(2 * l - 1), b = (double)(l + am - 1), inv = rinv[l - am];
      const double nq = (a * uz * q - b * p) * inv;
      if (GRAD) {
        const double nqx = (a * uz * qx - b * (2.0 * ux * p + px)) * inv;
        const double nqy = (a * uz * qy - b * (2.0 * uy * p + py)) * inv;
        const double nqz = (a * (q + uz * qz) - b * (2.0 * uz * p + pz)) * inv;
        px = qx; py = qy; pz = qz;
        qx = nqx; qy = nqy; qz = nqz;
      }
      p = q;
      q = nq;
    }
    const int lm = l * l + l + m;
    const double Nn = ynorm[lm];
    if (PACK) {
      double2* y4 = reinterpret_cast<double2*>(Y + 4 * lm);
      y4[0] = make_double2(Nn * q * f, Nn * (qx * f + q * fx));
      y4[1] = make_double2(Nn * (qy * f + q * fy), Nn * qz * f);
    } else {
      Y[lm] = Nn * q * f;
      if (GRAD) {
        gY[3 * lm] = Nn * (qx * f + q * fx);
        gY[3 * lm + 1] = Nn * (qy * f + q * fy);
        gY[3 * lm + 2] = Nn * qz * f;
      }
    }
  }
}

// Step 2: radial functions of the pair's species block and their r-derivatives.
template <class T>
__device__ void pair_step2(const SoapParams& P, int s, T r, int t, int nthr, PairScratch<T>& ps) {
  const SpeciesParams& sp = P.sp[s];
  const T inv_rc = (T)sp.inv_rc;
  const T x = r * inv_rc;
  const T sig = (T)sp.sig_r * ((T)1 + (T)sp.sig_r_sc * r) * inv_rc;
  const T dsig = (T)(sp.sig_r * sp.sig_r_sc) * inv_rc;
  const T a = (T)sp.amp_a;
  const T Pp = (T)1 + (T)2 * x * x * x - (T)3 * x * x;
  const T dPp = ((T)6 * x * x - (T)6 * x) * inv_rc;
  T amp, damp;
  if (a == (T)0) {
    amp = (T)1 / sig;
    damp = -dsig / (sig * sig);
  } else if (Pp <= (T)1.0e-10) {
    amp = 0; damp = 0;
  } else {
    const T Pa = t_pow<T>(Pp, a);
    amp = Pa / sig;
    damp = a * t_pow<T>(Pp, a - (T)1) * dPp / sig - Pa * dsig / (sig * sig);
  }
  amp *= (T)sp.gs;
  damp *= (T)sp.gs;
  const double* B = P.B + (size_t)s * TGX_MAX_NB * TGX_MAX_Q;
  if (nthr == 64) {
    // 64 lanes = nb groups x (64/nb) node slices (nb <= 16 divides into 4-lane slices at least)
    const int parts = sp.nb <= 8 ? 8 : 4;
    const int n = t / parts, part = t - n * parts;
    T I = 0, dI = 0;
    if (n < sp.nb) {
      const double* Bn = B + n * TGX_MAX_Q;
      for (int q = part; q < sp.nq; q += parts) {
        const T b = (T)Bn[q];
        I += b * ps.G[q];
        dI += b * ps.dG[q];
      }
    }
    for (int o = 1; o < parts; o <<= 1) {
      I += __shfl_xor(I, o);
      dI += __shfl_xor(dI, o);
    }
    if (n < sp.nb && part == 0) {
      ps.R[n] = amp * I;
      ps.dR[n] = damp * I + amp * dI;
    }
    return;
  }
  for (int n = t; n < sp.nb; n += nthr) {
    T I = 0, dI = 0;
    const double* Bn = B + n * TGX_MAX_Q;
    for (int q = 0; q < sp.nq; ++q) {
      const T b = (T)Bn[q];
      I += b * ps.G[q];
      dI += b * ps.dG[q];
    }
    ps.R[n] = amp * I;
    ps.dR[n] = damp * I + amp * dI;
  }
}

struct SiteArgs {
  const int* site_atom;
  const int64_t* site_beg;
  const int* pr_spec;
  const double* pr_r;
  const double* pr_xyz;
  const int64_t* pr_gidx;
};

// Phase A: expansion coefficients c(n, lm) of one site into LDS.
template <class T>
__device__ void site_coefficients(const SoapParams& P, const SiteArgs& a, int64_t b0, int nn, T* c, T* wscr) {
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int NLM = P.NLM, L = P.L;
  for (int e = tid; e < P.N * NLM; e += WG) c[e] = 0;
  PairScratch<T> ps;
  ps.bind(wscr + w * PairScratch<T>::size(L, NLM), L, NLM);
  __syncthreads();
  // central atom (self pair, first of the site): its own species block, l = 0
  {
    const int s = a.pr_spec[b0];
    if (s >= 0)
      for (int n = tid; n < P.sp[s].nb; n += WG) c[(P.sp[s].off + n) * NLM] += (T)P.cen[s * TGX_MAX_NB + n];
  }
  __syncthreads();
  for (int g = 1; g < nn; g += 4) {
    const int k = g + w;
    int s = -1;
    T r = 0, ux = 0, uy = 0, uz = 0;
    if (k < nn) {
      s = a.pr_spec[b0 + k];
      r = (T)a.pr_r[b0 + k];
      if (s >= 0 && !(a.pr_r[b0 + k] < P.sp[s].rc)) s = -1;
      if (s >= 0) {
        const T inv = (T)1 / r;
        ux = (T)a.pr_xyz[3 * (b0 + k)] * inv;
        uy = (T)a.pr_xyz[3 * (b0 + k) + 1] * inv;
        uz = (T)a.pr_xyz[3 * (b0 + k) + 2] * inv;
      }
    }
    if (s >= 0) pair_step1<T>(P, s, r, ux, uy, uz, lane, 64, ps, false);
    __syncthreads();
    if (s >= 0) pair_step2<T>(P, s, r, lane, 64, ps);
    __syncthreads();
    // accumulate the (up to) four pairs of this group in pair order
    for (int ww = 0; ww < 4; ++ww) {
      const int kk = g + ww;
      if (kk >= nn) break;
      int ss = a.pr_spec[b0 + kk];
      if (ss < 0 || !(a.pr_r[b0 + kk] < P.sp[ss].rc)) continue;
      PairScratch<T> po;
      po.bind(wscr + ww * PairScratch<T>::size(L, NLM), L, NLM);
      const int nb = P.sp[ss].nb, off = P.sp[ss].off;
      for (int e = tid; e < nb * NLM; e += WG) {
        const int n = e / NLM, lm = e - n * NLM;
        c[(off + n) * NLM + lm] += po.R[n] * po.A[lm_l(lm)] * po.Y[lm];
      }
    }
    __syncthreads();
  }
}

// Phase B: power spectrum into X[0..D), |p| returned to every thread; X is
// normalised in place to q.
template <class T>
__device__ T power_spectrum(const SoapParams& P, const T* c, T* X, T* red) {
  const int tid = threadIdx.x;
  const int NLM = P.NLM;
  T acc = 0;
  for (int d = tid; d < P.D; d += WG) {
    const int code = P.dtab[d];
    const int n = code & 0xff, np = (code >> 8) & 0xff, l = code >> 16;
    const T* c1 = c + n * NLM + l * l;
    const T* c2 = c + np * NLM + l * l;
    T s = 0;
    for (int m = 0; m < 2 * l + 1; ++m) s += c1[m] * c2[m];
    if (n != np) s *= (T)M_SQRT2;
    X[d] = s;
    acc += s * s;
  }
  red[tid] = acc;
  __syncthreads();
  for (int o = WG / 2; o > 0; o >>= 1) {
    if (tid < o) red[tid] += red[tid + o];
    __syncthreads();
  }
  const T nrm = t_sqrt<T>(red[0]);
  __syncthreads();
  const T inv = nrm > (T)0 ? (T)1 / nrm : (T)0;
  for (int d = tid; d < P.D; d += WG) X[d] *= inv;
  __syncthreads();
  return nrm;
}

// H(n, lm) = sum_n' S(n, n', l) c(n', lm) with S = sqrt2 V(n,n',l) off the
// diagonal and 2 V(n,n,l) on it: the gradient of V . p with respect to c.
template <class T>
__device__ void contract_channel(const SoapParams& P, const T* V, const T* c, T* H) {
  const int NLM = P.NLM, N = P.N, L1 = P.L + 1;
  for (int e = threadIdx.x; e < N * NLM; e += WG) {
    const int n = e / NLM, lm = e - n * NLM, l = lm_l(lm);
    T h = 0;
    for (int np = 0; np < N; ++np) {
      const int a = n < np ? n : np, b = n < np ? np : n;
      const int pidx = a * N - a * (a - 1) / 2 + (b - a);
      const T v = V[pidx * L1 + l];
      h += (n == np ? (T)2 : (T)M_SQRT2) * v * c[np * NLM + lm];
    }
    H[e] = h;
  }
}

// Mode 0: descriptors only. Mode 1: descriptors and dq/dr (plain path, FP64).
template <class T, int MODE, int KMAX>
__global__ __launch_bounds__(WG) void k_soap_site(SoapParams P, SiteArgs a, int n_sites, T* __restrict__ q_out,
                                                  T* __restrict__ pnorm_out, double* __restrict__ der_out,
                                                  const int* __restrict__ site_list) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  if ((int)blockIdx.x >= n_sites) return;
  // with a site list (descriptors only), row blockIdx.x of the output holds site_list[blockIdx.x]
  const int site = site_list ? site_list[blockIdx.x] : (int)blockIdx.x;
  const int orow = blockIdx.x;
  const int tid = threadIdx.x;
  const int NLM = P.NLM, N = P.N, L = P.L, L1 = L + 1, D = P.D;
  T* c = sm;                                   // N*NLM
  T* X = c + N * NLM;                          // D (p then q)
  T* red = X + D;                              // WG
  T* H = red + WG;                             // N*NLM (mode 1)
  T* wscr = H + (MODE == 1 ? N * NLM : 0);     // 4 pair scratches / phase D scratch
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  site_coefficients<T>(P, a, b0, nn, c, wscr);
  const T nrm = power_spectrum<T>(P, c, X, red);
  T* qrow = q_out + (size_t)orow * P.Dpad;
  for (int d = tid; d < P.Dpad; d += WG) qrow[d] = d < D ? X[d] : (T)0;
  if (tid == 0) pnorm_out[orow] = nrm;
  if (MODE == 0) return;
  // ---- dq/dr for every pair of the site (MODE 1) ----
  contract_channel<T>(P, X, c, H);
  __syncthreads();
  PairScratch<T> ps;
  ps.bind(wscr, L, NLM);
  T* s_nl = wscr + PairScratch<T>::size(L, NLM);  // N*L1
  T* vg_nl = s_nl + N * L1;                       // 3*N*L1 (becomes v)
  T* E_nl = vg_nl + 3 * N * L1;                   // N*L1
  T* Dd_nl = E_nl + N * L1;                       // N*L1
  T* sH = Dd_nl + N * L1;                         // MAX_NB*L1
  T* vgH = sH + TGX_MAX_NB * L1;                  // 3*MAX_NB*L1
  T* qd = vgH + 3 * TGX_MAX_NB * L1;              // 3 + pad
  // component ownership: d = tid + WG*k, k < KMAX (KMAX*WG >= D)
  T self_acc[KMAX][3];
  T qreg[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    self_acc[k][0] = self_acc[k][1] = self_acc[k][2] = 0;
    const int d = tid + WG * k;
    qreg[k] = d < D ? X[d] : (T)0;
  }
  const T inv_n = nrm > (T)0 ? (T)1 / nrm : (T)0;
  for (int k = 1; k < nn; ++k) {
    const int64_t p = b0 + k;
    int s = a.pr_spec[p];
    const double rr = a.pr_r[p];
    if (s >= 0 && !(rr < P.sp[s].rc)) s = -1;
    double* out = der_out + (size_t)p * 3 * D;
    if (s < 0 || nrm == (T)0) {
      // a pair the descriptor does not see: q does not depend on it
      for (int e = tid; e < 3 * D; e += WG) out[e] = 0.0;
      continue;
    }
    const T r = (T)rr, inv_r = (T)1 / r;
    const T ux = (T)a.pr_xyz[3 * p] * inv_r, uy = (T)a.pr_xyz[3 * p + 1] * inv_r, uz = (T)a.pr_xyz[3 * p + 2] * inv_r;
    pair_step1<T>(P, s, r, ux, uy, uz, tid, WG, ps, true);
    __syncthreads();
    const int nb = P.sp[s].nb, off = P.sp[s].off;
    if (tid < 64) {
      pair_step2<T>(P, s, r, tid, 64, ps);
    } else {
      // channel sums over m: s(n,l) = sum_m Y c, vg(n,l) = sum_m gY c, and the
      // same against H for the pair's block
      const int tasks = N * L1 + nb * L1;
      for (int task = tid - 64; task < tasks; task += WG - 64) {
        const bool withH = task >= N * L1;
        const int tt = withH ? task - N * L1 : task;
        const int n = tt / L1, l = tt - n * L1;
        const T* src = withH ? H + (off + n) * NLM : c + n * NLM;
        T s0 = 0, g0 = 0, g1 = 0, g2 = 0;
        for (int lm = l * l; lm < (l + 1) * (l + 1); ++lm) {
          const T v = src[lm];
          s0 += ps.Y[lm] * v;
          g0 += ps.gY[3 * lm] * v;
          g1 += ps.gY[3 * lm + 1] * v;
          g2 += ps.gY[3 * lm + 2] * v;
        }
        if (withH) {
          sH[tt] = s0; vgH[3 * tt] = g0; vgH[3 * tt + 1] = g1; vgH[3 * tt + 2] = g2;
        } else {
          s_nl[tt] = s0; vg_nl[3 * tt] = g0; vg_nl[3 * tt + 1] = g1; vg_nl[3 * tt + 2] = g2;
        }
      }
    }
    __syncthreads();
    // E = R A, Dd = (R A)', v = (vg - l s u)/r; zero outside the block
    for (int e = tid; e < N * L1; e += WG) {
      const int n = e / L1, l = e - n * L1;
      const T sv = s_nl[e];
      const T lf = (T)l;
      vg_nl[3 * e] = (vg_nl[3 * e] - lf * sv * ux) * inv_r;
      vg_nl[3 * e + 1] = (vg_nl[3 * e + 1] - lf * sv * uy) * inv_r;
      vg_nl[3 * e + 2] = (vg_nl[3 * e + 2] - lf * sv * uz) * inv_r;
      const int nl = n - off;
      if (nl >= 0 && nl < nb) {
        E_nl[e] = ps.R[nl] * ps.A[l];
        Dd_nl[e] = ps.dR[nl] * ps.A[l] + ps.R[nl] * ps.dA[l];
      } else {
        E_nl[e] = 0;
        Dd_nl[e] = 0;
      }
    }
    // q . dp: sum over the block's channels (one wavefront, fixed order)
    if (tid < 64) {
      T a0 = 0, a1 = 0, a2 = 0;
      for (int tt = tid; tt < nb * L1; tt += 64) {
        const int n = tt / L1, l = tt - n * L1;
        const T E = ps.R[n] * ps.A[l];
        const T Dd = ps.dR[n] * ps.A[l] + ps.R[n] * ps.dA[l];
        const T sh = sH[tt], lf = (T)l;
        a0 += Dd * sh * ux + E * (vgH[3 * tt] - lf * sh * ux) * inv_r;
        a1 += Dd * sh * uy + E * (vgH[3 * tt + 1] - lf * sh * uy) * inv_r;
        a2 += Dd * sh * uz + E * (vgH[3 * tt + 2] - lf * sh * uz) * inv_r;
      }
      for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o);
        a1 += __shfl_xor(a1, o);
        a2 += __shfl_xor(a2, o);
      }
      if (tid == 0) { qd[0] = a0; qd[1] = a1; qd[2] = a2; }
    }
    __syncthreads();
    const T q0 = qd[0], q1 = qd[1], q2 = qd[2];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
      const int d = tid + WG * k;
      if (d >= D) break;
      const int code = P.dtab[d];
      const int n = code & 0xff, np = (code >> 8) & 0xff, l = code >> 16;
      const int i1 = n * L1 + l, i2 = np * L1 + l;
      const T w = n == np ? (T)1 : (T)M_SQRT2;
      const T rad = Dd_nl[i1] * s_nl[i2] + Dd_nl[i2] * s_nl[i1];
      const T E1 = E_nl[i1], E2 = E_nl[i2];
      const T dp0 = w * (rad * ux + E1 * vg_nl[3 * i2] + E2 * vg_nl[3 * i1]);
      const T dp1 = w * (rad * uy + E1 * vg_nl[3 * i2 + 1] + E2 * vg_nl[3 * i1 + 1]);
      const T dp2 = w * (rad * uz + E1 * vg_nl[3 * i2 + 2] + E2 * vg_nl[3 * i1 + 2]);
      const T o0 = (dp0 - qreg[k] * q0) * inv_n;
      const T o1 = (dp1 - qreg[k] * q1) * inv_n;
      const T o2 = (dp2 - qreg[k] * q2) * inv_n;
      out[3 * d] = (double)o0;
      out[3 * d + 1] = (double)o1;
      out[3 * d + 2] = (double)o2;
      self_acc[k][0] -= o0;
      self_acc[k][1] -= o1;
      self_acc[k][2] -= o2;
    }
    __syncthreads();
  }
  double* out = der_out + (size_t)b0 * 3 * D;
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const int d = tid + WG * k;
    if (d >= D) break;
    out[3 * d] = (double)self_acc[k][0];
    out[3 * d + 1] = (double)self_acc[k][1];
    out[3 * d + 2] = (double)self_acc[k][2];
  }
}

// Workgroup barrier that orders LDS only: it waits for this wavefront's LDS
// operations, not for its global stores, so streamed output stays in flight
// across the barrier (no thread of a workgroup reads another's global output).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ void der_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-pair channel tables of the plain dq/dr kernel, one slot per wavefront:
//   s(n,l) = sum_m Y_lm c(n,lm)
//   v_k(n,l) = (sum_m dY_lm/du_k c(n,lm) - l s u_k) / r     (k = x, y, z)
//   E(n,l) = R_n A_l, Dd(n,l) = (R_n A_l)'   (zero outside the pair's block)
// and misc = { q.dp_x, q.dp_y, q.dp_z, u_x, u_y, u_z, valid }.
struct DerSlot {
  static __host__ __device__ int size(int NL, int L, int NLM) {
    return 6 * NL + 8 + PairScratch<double>::size(L, NLM);
  }
};

// One wavefront reduces pair p of the site to its channel tables in `slot`.
__device__ void der_pair_tables(const SoapParams& P, const SiteArgs& a, int64_t p, bool site_ok, const double* c,
                                const double* H, double* slot) {
  const int lane = threadIdx.x & 63;
  const int NLM = P.NLM, L = P.L, L1 = L + 1, NL = P.N * L1;
  double* s_t = slot;
  double* v_t = s_t + NL;
  double* E_t = v_t + 3 * NL;
  double* Dd_t = E_t + NL;
  double* misc = Dd_t + NL;
  int s = a.pr_spec[p];
  const double rr = a.pr_r[p];
  if (!site_ok || s < 0 || !(rr < P.sp[s].rc)) {
    if (lane == 0) misc[6] = 0.0;
    return;
  }
  PairScratch<double> ps;
  ps.bind(misc + 8, L, NLM);
  const double r = rr, inv_r = 1.0 / r;
  const double ux = a.pr_xyz[3 * p] * inv_r, uy = a.pr_xyz[3 * p + 1] * inv_r, uz = a.pr_xyz[3 * p + 2] * inv_r;
  pair_step1<double>(P, s, r, ux, uy, uz, lane, 64, ps, true);
  der_wave_sync();
  pair_step2<double>(P, s, r, lane, 64, ps);
  der_wave_sync();
  const int nb = P.sp[s].nb, off = P.sp[s].off;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int t = lane; t < NL; t += 64) {
    const int n = t / L1, l = t - n * L1;
    const int nl = n - off;
    const bool inb = nl >= 0 && nl < nb;
    const double* cr = c + n * NLM;
    const double* hr = H + n * NLM;
    double s0 = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    double h = 0.0, h0 = 0.0, h1 = 0.0, h2 = 0.0;
    for (int lm = l * l; lm < (l + 1) * (l + 1); ++lm) {
      const double Y = ps.Y[lm], gx = ps.gY[3 * lm], gy = ps.gY[3 * lm + 1], gz = ps.gY[3 * lm + 2];
      const double v = cr[lm];
      s0 += Y * v;
      g0 += gx * v;
      g1 += gy * v;
      g2 += gz * v;
      if (inb) {
        const double hv = hr[lm];
        h += Y * hv;
        h0 += gx * hv;
        h1 += gy * hv;
        h2 += gz * hv;
      }
    }
    const double lf = (double)l;
    s_t[t] = s0;
    v_t[t] = (g0 - lf * s0 * ux) * inv_r;
    v_t[NL + t] = (g1 - lf * s0 * uy) * inv_r;
    v_t[2 * NL + t] = (g2 - lf * s0 * uz) * inv_r;
    double E = 0.0, Dd = 0.0;
    if (inb) {
      E = ps.R[nl] * ps.A[l];
      Dd = ps.dR[nl] * ps.A[l] + ps.R[nl] * ps.dA[l];
      a0 += Dd * h * ux + E * (h0 - lf * h * ux) * inv_r;
      a1 += Dd * h * uy + E * (h1 - lf * h * uy) * inv_r;
      a2 += Dd * h * uz + E * (h2 - lf * h * uz) * inv_r;
    }
    E_t[t] = E;
    Dd_t[t] = Dd;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o);
    a1 += __shfl_xor(a1, o);
    a2 += __shfl_xor(a2, o);
  }
  if (lane == 0) {
    misc[0] = a0; misc[1] = a1; misc[2] = a2;
    misc[3] = ux; misc[4] = uy; misc[5] = uz;
    misc[6] = 1.0;
  }
}

// Plain-path dq/dr in FP64, (3, D) per pair like soap_cart_der
// (src/gap_interface.f90:312-317), the self pair holding -sum of the others
// (src/gap.f90:275-277). One workgroup per site. The site's coefficients c,
// its power spectrum and H = d(q . p)/dc stay in LDS; the pairs go four at a
// time: each wavefront reduces one pair to its channel tables, then the 256
// threads stream the four slabs. Thread t owns elements e = t + 256 j of a
// slab (e = 3 d + k, k the Cartesian component), so each store instruction
// writes 2 KB of consecutive doubles, and it keeps the self-pair sums of its
// elements in registers. J * 256 >= 3 D.
template <int J>
__global__ __launch_bounds__(WG, 2) void k_soap_der(SoapParams P, SiteArgs a, int n_sites, double* __restrict__ q_out,
                                                    double* __restrict__ pnorm_out, double* __restrict__ der_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* sm = reinterpret_cast<double*>(smem_raw);
  const int site = blockIdx.x;
  if (site >= n_sites) return;
  const int tid = threadIdx.x, w = tid >> 6;
  const int NLM = P.NLM, N = P.N, L1 = P.L + 1, D = P.D, D3 = 3 * D, NL = N * L1;
  double* c = sm;
  double* H = c + N * NLM;
  double* X = H + N * NLM;  // p, then q: kept for the whole site
  double* U = X + D;
  double* red = U;
  double* wscr = red + WG;
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  site_coefficients<double>(P, a, b0, nn, c, wscr);
  const double nrm = power_spectrum<double>(P, c, X, red);
  double* qrow = q_out + (size_t)site * P.Dpad;
  for (int d = tid; d < P.Dpad; d += WG) qrow[d] = d < D ? X[d] : 0.0;
  if (tid == 0) pnorm_out[site] = nrm;
  contract_channel<double>(P, X, c, H);
  // element codes: i1 = n L1 + l | i2 = n' L1 + l << 8 | k << 16 | (n == n') << 18 | d << 19
  int code[J];
  double self[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int e = tid + WG * j;
    code[j] = -1;
    self[j] = 0.0;
    if (e < D3) {
      const int d = e / 3, k = e - 3 * d;
      const int cd = P.dtab[d];
      const int n = cd & 0xff, np = (cd >> 8) & 0xff, l = cd >> 16;
      code[j] = (n * L1 + l) | ((np * L1 + l) << 8) | (k << 16) | ((n == np ? 1 : 0) << 18) | (d << 19);
    }
  }
  const double inv_n = nrm > 0.0 ? 1.0 / nrm : 0.0;
  const bool site_ok = nrm > 0.0;
  const int SL = DerSlot::size(NL, P.L, NLM);
  __syncthreads();  // red and the coefficient scratch are dead: the slots overlay them
  for (int g = 1; g < nn; g += 4) {
    if (g + w < nn) der_pair_tables(P, a, b0 + g + w, site_ok, c, H, U + w * SL);
    lds_barrier();
    const int cnt = min(4, nn - g);
    double* out = der_out + (size_t)(b0 + g) * D3 + tid;
    // element-major: an element's decoded indices are live only while its four pairs are written
#pragma unroll
    for (int j = 0; j < J; ++j) {
      int cj = code[j];
      if (cj < 0) continue;
      // opaque to the optimiser: the decoded fields are not hoisted out of the pair loop, where
      // they would stay live (16 elements' worth) across the table phase
      asm volatile("" : "+v"(cj));
      const int i1 = cj & 0xff, i2 = (cj >> 8) & 0xff, k = (cj >> 16) & 3;
      const double wgt = ((cj >> 18) & 1) ? 1.0 : M_SQRT2;
      const double qd = X[cj >> 19];
      for (int ww = 0; ww < cnt; ++ww) {
        const double* s_t = U + ww * SL;
        const double* v_t = s_t + NL;
        const double* E_t = v_t + 3 * NL;
        const double* Dd_t = E_t + NL;
        const double* misc = Dd_t + NL;
        double o = 0.0;  // zero for a pair the descriptor does not see: q does not depend on it
        if (misc[6] != 0.0) {
          const double rad = Dd_t[i1] * s_t[i2] + Dd_t[i2] * s_t[i1];
          const double dp = wgt * (rad * misc[3 + k] + E_t[i1] * v_t[k * NL + i2] + E_t[i2] * v_t[k * NL + i1]);
          o = (dp - qd * misc[k]) * inv_n;
        }
        out[(size_t)ww * D3 + WG * j] = o;
        self[j] -= o;
      }
    }
    lds_barrier();
  }
  double* out = der_out + (size_t)b0 * D3;
#pragma unroll
  for (int j = 0; j < J; ++j)
    if (code[j] >= 0) out[tid + WG * j] = self[j];
}

// ---- The same dq/dr for a compile-time l_max and at most two species of at
// most eight radial functions each. Every wavefront works on its own pairs in
// both pair phases; the site's pair data, the radial basis and the nodes sit in
// registers, so a pair's set-up does not wait on memory.

template <int LL> struct DerL {
  static constexpr int L1 = LL + 1, NLM = L1 * L1, NL = 16 * L1;
  static constexpr int PS = 2 * TGX_MAX_Q + 2 * TGX_MAX_NB + 2 * L1 + 4 * NLM;  // PairScratch
  static constexpr int SL = 6 * NL + 8 + PS;                                    // one pair slot
};

__device__ __forceinline__ double lane_read(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The pair-only factors of a pair's radial values and angular damping (what
// der_radial and der_gauss would otherwise form in every lane for every pair):
// the amplitude A(r) = global_scaling / s (1 + 2x^3 - 3x^2)^amplitude_scaling
// and dA/dr, and e = exp(-k) and -dk/dr of the damping exponent
// k = st^2 / (2 r^2). Zero for a pair outside the cutoff or unseen.
struct PairAmp {
  double amp, damp, ea, dk;
};

__device__ __forceinline__ PairAmp pair_amp(const SpeciesParams& sp, double r) {
  PairAmp o;
  const double inv_rc = sp.inv_rc;
  const double x = r * inv_rc;
  const double sig = sp.sig_r * (1.0 + sp.sig_r_sc * r) * inv_rc;
  const double dsig = (sp.sig_r * sp.sig_r_sc) * inv_rc;
  const double aa = sp.amp_a;
  const double Pp = 1.0 + 2.0 * x * x * x - 3.0 * x * x;
  const double dPp = (6.0 * x * x - 6.0 * x) * inv_rc;
  double amp, damp;
  if (aa == 0.0) {
    amp = 1.0 / sig;
    damp = -dsig / (sig * sig);
  } else if (Pp <= 1.0e-10) {
    amp = 0; damp = 0;
  } else {
    const double Pa = pow(Pp, aa);
    amp = Pa / sig;
    damp = aa * pow(Pp, aa - 1.0) * dPp / sig - Pa * dsig / (sig * sig);
  }
  o.amp = amp * sp.gs;
  o.damp = damp * sp.gs;
  const double st = sp.sig_t * (1.0 + sp.sig_t_sc * r);
  const double dst = sp.sig_t * sp.sig_t_sc;
  const double kk = st * st / (2.0 * r * r);
  o.dk = st * st / (r * r * r) - st * dst / (r * r);  // -dk/dr
  o.ea = exp(-kk);
  return o;
}

// Register copies, per lane: pair 1 + lane of the site with its PairAmp;
// B[s][n][q] at n = lane / 8, q = lane % 8 + 8 i; the node x_q of each species
// at q = lane. At most two species (the blocks this kernel serves).
struct DerRegs {
  int spec;
  double r, x, y, z;
  PairAmp pa;
  double B[2][8];
  double xq[2];
  __device__ void load(const SoapParams& P, const SiteArgs& a, int64_t b0, int nn, int lane) {
    const int k = 1 + lane;
    spec = -1;
    r = 1.0; x = y = z = 0.0;
    if (k < nn) {
      spec = a.pr_spec[b0 + k];
      r = a.pr_r[b0 + k];
      x = a.pr_xyz[3 * (b0 + k)];
      y = a.pr_xyz[3 * (b0 + k) + 1];
      z = a.pr_xyz[3 * (b0 + k) + 2];
    }
    pa = PairAmp{0.0, 0.0, 0.0, 0.0};
    if (spec >= 0 && r < (spec == 0 ? P.sp[0].rc : P.sp[1].rc)) pa = pair_amp(spec == 0 ? P.sp[0] : P.sp[1], r);
    const int n = lane >> 3, part = lane & 7;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const bool have = s < P.n_species;
#pragma unroll
      for (int i = 0; i < 8; ++i) B[s][i] = have ? P.B[((size_t)s * TGX_MAX_NB + n) * TGX_MAX_Q + part + 8 * i] : 0.0;
      xq[s] = have ? P.xq[s * TGX_MAX_Q + lane] : 0.0;
    }
  }
};

struct DerPair {
  int s;
  bool ok;
  double r, ux, uy, uz;
  PairAmp pa;
};

// Pair k >= 1 of the site (wavefront-uniform): from the lanes for k <= 64, else from memory.
__device__ __forceinline__ DerPair der_pair(const SoapParams& P, const SiteArgs& a, const DerRegs& R, int64_t b0,
                                            int k) {
  DerPair q;
  double x, y, z;
  const bool reg = k <= 64;
  if (reg) {
    q.s = __builtin_amdgcn_readlane(R.spec, k - 1);
    q.r = lane_read(R.r, k - 1);
    x = lane_read(R.x, k - 1);
    y = lane_read(R.y, k - 1);
    z = lane_read(R.z, k - 1);
    q.pa.amp = lane_read(R.pa.amp, k - 1);
    q.pa.damp = lane_read(R.pa.damp, k - 1);
    q.pa.ea = lane_read(R.pa.ea, k - 1);
    q.pa.dk = lane_read(R.pa.dk, k - 1);
  } else {
    const int64_t p = b0 + k;
    q.s = a.pr_spec[p];
    q.r = a.pr_r[p];
    x = a.pr_xyz[3 * p];
    y = a.pr_xyz[3 * p + 1];
    z = a.pr_xyz[3 * p + 2];
  }
  // wavefront-uniform: the species parameters then come from scalar loads
  q.s = __builtin_amdgcn_readfirstlane(q.s);
  q.r = lane_read(q.r, 0);
  x = lane_read(x, 0);
  y = lane_read(y, 0);
  z = lane_read(z, 0);
  q.ok = q.s >= 0 && q.r < P.sp[q.s].rc;
  if (!reg && q.ok) q.pa = pair_amp(P.sp[q.s], q.r);
  const double inv = 1.0 / q.r;
  q.ux = x * inv;
  q.uy = y * inv;
  q.uz = z * inv;
  return q;
}

// pair_gauss with the nodes from registers (lane q < nq) and the damping on
// lane 63 from the pair's e = exp(-k).
__device__ __forceinline__ void der_gauss(const SoapParams& P, const DerRegs& R, const DerPair& pr, int lane,
                                          PairScratch<double>& ps, bool need_grad) {
  const int s = pr.s;
  const double r = pr.r;
  const SpeciesParams& sp = P.sp[s];
  const double inv_rc = sp.inv_rc;
  const double x = r * inv_rc;
  const double sig = sp.sig_r * (1.0 + sp.sig_r_sc * r) * inv_rc;
  const double dsig = (sp.sig_r * sp.sig_r_sc) * inv_rc;
  const double inv_s2 = 1.0 / (sig * sig);
  if (lane < sp.nq) {
    const double dx = (s == 0 ? R.xq[0] : R.xq[1]) - x;
    const double g = exp(-0.5 * dx * dx * inv_s2);
    ps.G[lane] = g;
    if (need_grad) ps.dG[lane] = g * (dx * inv_s2 * inv_rc + dx * dx * inv_s2 / sig * dsig);
  }
  if (lane == 63) {
    const double dk = pr.pa.dk;
    const double e1 = pr.pa.ea, e2 = e1 * e1;
    double Al = 1, pw = 1;
    for (int l = 0; l <= P.L; ++l) {
      if (l > 0) { pw *= e2; Al *= pw; }
      ps.A[l] = Al;
      ps.dA[l] = Al * (double)(l * (l + 1)) * dk;
    }
  }
}

// pair_step2 for nb <= 8 (eight node slices per radial function) with B from
// registers and the pair's amplitude; the node values are read before they
// are summed, in the same order.
__device__ __forceinline__ void der_radial(const SoapParams& P, const DerRegs& R, const DerPair& pr, int lane,
                                           PairScratch<double>& ps, bool need_grad) {
  const SpeciesParams& sp = P.sp[pr.s];
  const double amp = pr.pa.amp, damp = pr.pa.damp;
  const int n = lane >> 3, part = lane & 7, nq = sp.nq;
  double g[8], dg[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int q = min(part + 8 * i, nq - 1);
    g[i] = ps.G[q];
    dg[i] = need_grad ? ps.dG[q] : 0.0;
  }
  double I = 0, dI = 0;
  if (n < sp.nb) {
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (part + 8 * i < nq) {
        const double b = pr.s == 0 ? R.B[0][i] : R.B[1][i];
        I += b * g[i];
        dI += b * dg[i];
      }
  }
  for (int o = 1; o < 8; o <<= 1) {
    I += __shfl_xor(I, o);
    if (need_grad) dI += __shfl_xor(dI, o);
  }
  if (n < sp.nb && part == 0) {
    ps.R[n] = amp * I;
    if (need_grad) ps.dR[n] = damp * I + amp * dI;
  }
}

// Channel tables of four pairs at once (slots 0..3 of U, pair data and
// scratch filled by their wavefronts): lanes (n, p) = (lane & 15, lane >> 4)
// take radial row n of pair p, and wavefront w the degrees of its set, so
// every (n, l) entry of every pair is formed once and each c / H read serves
// eight products. The pair's harmonics are packed (Y, gx, gy, gz) per lm.
// Entry (n, l) of a slot is six doubles (s, E, v_x, v_y, v_z, Dd) at
// 6 (n L1 + l), read back as three 16-byte loads:
//   s(n,l) = sum_m Y c(n,lm),  v_k(n,l) = (sum_m dY_k c(n,lm) - l s u_k) / r,
//   E(n,l) = R_n A_l,  Dd(n,l) = dR_n A_l + R_n dA_l  (rows of the pair's block),
// and the pair's q . dp_k partial over the set, sum_(n,l) Dd h u_k +
// E (h_k - l h u_k) / r with h = sum_m Y H(n,lm), h_k = sum_m dY_k H(n,lm),
// reduced over the 16 row lanes into qpart[(w * 4 + p) * 3 + k].
template <int LL>
__device__ __forceinline__ void der_tables_w(const SoapParams& P, int w, int lane, int N, double* U, int SL,
                                             const double* c, const double* H, double* qpart) {
  static_assert(LL == 8, "degree sets are laid out for l_max = 8");
  using G = DerL<LL>;
  constexpr int L1 = G::L1, NLM = G::NLM, NL = G::NL;
  const int n = lane & 15, p = lane >> 4;
  double* slot = U + p * SL;
  const double* misc = slot + 6 * NL;
  PairScratch<double> ps;
  ps.bind(const_cast<double*>(misc) + 8, LL, NLM);
  const bool ok = misc[0] != 0.0;
  const int off = (int)misc[1], nb = (int)misc[2];
  const double inv_r = 1.0 / misc[3], ux = misc[4], uy = misc[5], uz = misc[6];
  const int nl = n - off;
  const bool inb = ok && nl >= 0 && nl < nb;
  const bool row = n < N;
  const double* cr = c + n * NLM;
  const double* hr = H + n * NLM;
  double qx = 0.0, qy = 0.0, qz = 0.0;
  const int nsets = w == 0 ? 3 : 2;
  for (int li = 0; li < nsets; ++li) {
    const int l = li == 0 ? LL - w : (li == 1 ? (w == 0 ? 0 : w + 1) : 1);
    double s0 = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, h0 = 0.0, hx = 0.0, hy = 0.0, hz = 0.0;
    if (ok && row) {
#pragma unroll 3
      for (int lm = l * l; lm < (l + 1) * (l + 1); ++lm) {
        const double2* y4 = reinterpret_cast<const double2*>(ps.Y + 4 * lm);  // (Y, gx), (gy, gz)
        const double2 ya = y4[0], yb = y4[1];
        const double y = ya.x, dx = ya.y, dy = yb.x, dz = yb.y;
        const double cv = cr[lm], hv = hr[lm];
        s0 += y * cv;
        gx += dx * cv;
        gy += dy * cv;
        gz += dz * cv;
        h0 += y * hv;
        hx += dx * hv;
        hy += dy * hv;
        hz += dz * hv;
      }
    }
    const double lf = (double)l;
    const double E = inb ? ps.R[nl] * ps.A[l] : 0.0;
    const double Dd = inb ? ps.dR[nl] * ps.A[l] + ps.R[nl] * ps.dA[l] : 0.0;
    if (row) {
      double2* t = reinterpret_cast<double2*>(slot + 6 * (n * L1 + l));
      t[0] = make_double2(s0, E);
      t[1] = make_double2((gx - lf * s0 * ux) * inv_r, (gy - lf * s0 * uy) * inv_r);
      t[2] = make_double2((gz - lf * s0 * uz) * inv_r, Dd);
    }
    qx += Dd * h0 * ux + E * (hx - lf * h0 * ux) * inv_r;
    qy += Dd * h0 * uy + E * (hy - lf * h0 * uy) * inv_r;
    qz += Dd * h0 * uz + E * (hz - lf * h0 * uz) * inv_r;
  }
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) {
    qx += __shfl_xor(qx, o);
    qy += __shfl_xor(qy, o);
    qz += __shfl_xor(qz, o);
  }
  if (n == 0) {
    double* q = qpart + (w * 4 + p) * 3;
    q[0] = qx;
    q[1] = qy;
    q[2] = qz;
  }
}

// Plain-path dq/dr for l_max = LL (the output and element layout of
// k_soap_der). Coefficients: wavefront w expands pairs 1 + w, 5 + w, ... with
// lanes over (l, m) and the 16 rows in registers, and the four partial
// expansions are added in wavefront order. Pairs go in groups of four: each
// wavefront prepares one pair's Gaussians, harmonics and radial values, the
// four wavefronts then form the four pairs' channel tables together
// (der_tables_w), and each thread writes whole components d = tid + WG j (its
// three Cartesian entries, from 13 LDS reads) of every pair of the group,
// keeping the self-pair sums of its components in registers. J >= D / WG.
template <int LL, int J>
__global__ __launch_bounds__(WG, 2) void k_soap_der_l(SoapParams P, SiteArgs a, int n_sites,
                                                      double* __restrict__ q_out, double* __restrict__ pnorm_out,
                                                      double* __restrict__ der_out) {
  using G = DerL<LL>;
  constexpr int L1 = G::L1, NLM = G::NLM, NL = G::NL, SL = G::SL;
  static_assert(SL % 2 == 0 && (NLM + L1) % 2 == 0, "pair slots and tables need 16-byte alignment");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* sm = reinterpret_cast<double*>(smem_raw);
  const int site = blockIdx.x;
  if (site >= n_sites) return;
  const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int N = P.N, D = P.D, D3 = 3 * D, CN = N * NLM;
  double* ynorm = sm;                          // NLM
  double* rinv = ynorm + NLM;                  // L1
  double* c = rinv + L1;                       // N NLM
  double* H = c + CN;                          // N NLM
  double* X = H + CN;                          // D: p, then q
  double* U = X + D + (D & 1);                 // max(4 SL + 48, 4 N NLM, WG), 16-byte aligned
  const int64_t b0 = a.site_beg[site];
  const int nn = (int)(a.site_beg[site + 1] - b0);
  DerRegs RG;
  RG.load(P, a, b0, nn, lane);
  for (int i = tid; i < NLM; i += WG) ynorm[i] = P.ynorm[i];
  if (tid < L1) rinv[tid] = tid > 0 ? 1.0 / (double)tid : 0.0;
  lds_barrier();
  // ---- expansion coefficients ----
  {
    double acc[2][16];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int n = 0; n < 16; ++n) acc[h][n] = 0.0;
    PairScratch<double> ps;
    ps.bind(U + w * G::PS, LL, NLM);
    int lh[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) lh[h] = lm_l(lane + 64 * h);
    for (int k = 1 + w; k < nn; k += 4) {
      const DerPair pr = der_pair(P, a, RG, b0, k);
      if (!pr.ok) continue;
      der_gauss(P, RG, pr, lane, ps, false);
      harmonics_m<LL, false>(lane, ynorm, rinv, pr.ux, pr.uy, pr.uz, ps.Y, nullptr);
      der_wave_sync();
      der_radial(P, RG, pr, lane, ps, false);
      der_wave_sync();
      const int nb = P.sp[pr.s].nb, off = P.sp[pr.s].off;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int lm = lane + 64 * h;
        if (lm < NLM) {
          const double al = ps.A[lh[h]], yv = ps.Y[lm];
#pragma unroll
          for (int n = 0; n < 16; ++n) {
            const int nl = n - off;
            if (nl >= 0 && nl < nb) acc[h][n] += ps.R[nl] * al * yv;
          }
        }
      }
      der_wave_sync();
    }
    lds_barrier();  // the pair scratch is dead: the partial sums overlay it
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int lm = lane + 64 * h;
      if (lm < NLM)
#pragma unroll
        for (int n = 0; n < 16; ++n)
          if (n < N) U[w * CN + n * NLM + lm] = acc[h][n];
    }
  }
  lds_barrier();
  for (int e = tid; e < CN; e += WG) c[e] = ((U[e] + U[CN + e]) + U[2 * CN + e]) + U[3 * CN + e];
  lds_barrier();
  {
    // central atom (self pair): its own species block, l = 0
    const int s = a.pr_spec[b0];
    if (s >= 0)
      for (int n = tid; n < P.sp[s].nb; n += WG) c[(P.sp[s].off + n) * NLM] += P.cen[s * TGX_MAX_NB + n];
  }
  lds_barrier();
  const double nrm = power_spectrum<double>(P, c, X, U);
  double* qrow = q_out + (size_t)site * P.Dpad;
  for (int d = tid; d < P.Dpad; d += WG) qrow[d] = d < D ? X[d] : 0.0;
  if (tid == 0) pnorm_out[site] = nrm;
  contract_channel<double>(P, X, c, H);
  // components d = tid + WG j of this thread: row offsets of (n, l) and (n', l)
  // in the tables, n == n', d
  int code[J];
  double self[J][3];
#pragma unroll
  for (int j = 0; j < J; ++j) code[j] = P.dtab[min(tid + WG * j, D - 1)];  // all loads in flight at once
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int d = tid + WG * j;
    const int cd = code[j];
    const int n = cd & 0xff, np = (cd >> 8) & 0xff, l = cd >> 16;
    code[j] = d < D ? (n * L1 + l) | ((np * L1 + l) << 8) | ((n == np ? 1 : 0) << 16) | (d << 17) : -1;
    self[j][0] = self[j][1] = self[j][2] = 0.0;
  }
  const double inv_n = nrm > 0.0 ? 1.0 / nrm : 0.0;
  const bool site_ok = nrm > 0.0;
  double* qpart = U + 4 * SL;  // [wave][pair][k] partial q . dp
  lds_barrier();
  for (int g = 1; g < nn; g += 4) {
    // (a) wavefront w: pair g + w's Gaussians, harmonics with gradients and radial values
    {
      double* slot = U + w * SL;
      double* misc = slot + 6 * NL;
      PairScratch<double> ps;
      ps.bind(misc + 8, LL, NLM);
      bool ok = false;
      DerPair pr{};
      if (g + w < nn) {
        pr = der_pair(P, a, RG, b0, g + w);
        ok = pr.ok && site_ok;
      }
      if (ok) {
        der_gauss(P, RG, pr, lane, ps, true);
        harmonics_m<LL, true, true>(lane, ynorm, rinv, pr.ux, pr.uy, pr.uz, ps.Y, nullptr);
        der_wave_sync();
        der_radial(P, RG, pr, lane, ps, true);
      }
      if (lane == 0) {
        misc[0] = ok ? 1.0 : 0.0;
        misc[1] = ok ? (double)P.sp[pr.s].off : 0.0;
        misc[2] = ok ? (double)P.sp[pr.s].nb : 0.0;
        misc[3] = ok ? pr.r : 1.0;
        misc[4] = ok ? pr.ux : 0.0;
        misc[5] = ok ? pr.uy : 0.0;
        misc[6] = ok ? pr.uz : 0.0;
      }
    }
    lds_barrier();
    // (b) channel tables of the four pairs: lanes (n, pair), wavefront w takes
    // the degrees {8, 0, 1}, {7, 2}, {6, 3} or {5, 4} (20 or 21 (l, m) each)
    der_tables_w<LL>(P, w, lane, N, U, SL, c, H, qpart);
    lds_barrier();
    // (c) the slabs of the (up to) four pairs: thread owns components d, all three k
    const int cnt = min(4, nn - g);
    for (int pp = 0; pp < cnt; ++pp) {
      const double* slot = U + pp * SL;
      const double* misc = slot + 6 * NL;
      const double* qp = qpart + 3 * pp;
      const double q0 = ((qp[0] + qp[12]) + qp[24]) + qp[36];
      const double q1 = ((qp[1] + qp[13]) + qp[25]) + qp[37];
      const double q2 = ((qp[2] + qp[14]) + qp[26]) + qp[38];
      const double ux = misc[4], uy = misc[5], uz = misc[6];
      double* out = der_out + (size_t)(b0 + g + pp) * D3;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int cj = code[j];
        if (cj < 0) continue;
        const int i1 = cj & 0xff, i2 = (cj >> 8) & 0xff, d = cj >> 17;
        const double wgt = ((cj >> 16) & 1) ? 1.0 : M_SQRT2;
        const double qd = X[d];
        const double2* t1 = reinterpret_cast<const double2*>(slot + 6 * i1);
        const double2* t2 = reinterpret_cast<const double2*>(slot + 6 * i2);
        const double2 a1 = t1[0], b1 = t1[1], c1 = t1[2];  // (s, E), (v_x, v_y), (v_z, Dd)
        const double2 a2 = t2[0], b2 = t2[1], c2 = t2[2];
        const double rad = c1.y * a2.x + c2.y * a1.x;
        const double ox = (wgt * (rad * ux + a1.y * b2.x + a2.y * b1.x) - qd * q0) * inv_n;
        const double oy = (wgt * (rad * uy + a1.y * b2.y + a2.y * b1.y) - qd * q1) * inv_n;
        const double oz = (wgt * (rad * uz + a1.y * c2.x + a2.y * c1.x) - qd * q2) * inv_n;
        out[3 * d] = ox;
        out[3 * d + 1] = oy;
        out[3 * d + 2] = oz;
        self[j][0] -= ox;
        self[j][1] -= oy;
        self[j][2] -= oz;
      }
    }
    lds_barrier();
  }
  double* out = der_out + (size_t)b0 * D3;
#pragma unroll
  for (int j = 0; j < J; ++j)
    if (code[j] >= 0) {
      const int d = code[j] >> 17;
      out[3 * d] = self[j][0];
      out[3 * d + 1] = self[j][1];
      out[3 * d + 2] = self[j][2];
    }
}

// Plain-path pair force: f_p = Qss_i . soap_cart_der(:, :, p), one wavefront per
// pair, streamed over the (3, D) slab. This is the scatter of src/gap.f90:206-224.
__global__ __launch_bounds__(256) void k_plain_force(int n_sites, const int64_t* __restrict__ site_beg, int D,
                                                    int Dpad, const double* __restrict__ qss,
                                                    const double* __restrict__ der,
                                                    const int64_t* __restrict__ pr_gidx, int64_t n_pairs,
                                                    const int* __restrict__ pair_site, double* __restrict__ fglob) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p >= n_pairs) return;
  const int i = pair_site[p];
  if (i < 0) return;  // a pair no site claims (caller-supplied offsets)
  const double* s = qss + (size_t)i * Dpad;
  const double* g = der + (size_t)p * 3 * D;
  double a0 = 0, a1 = 0, a2 = 0;
  for (int d = lane; d < D; d += 64) {
    const double sv = s[d];
    a0 += sv * g[3 * d];
    a1 += sv * g[3 * d + 1];
    a2 += sv * g[3 * d + 2];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a0 += __shfl_xor(a0, o);
    a1 += __shfl_xor(a1, o);
    a2 += __shfl_xor(a2, o);
  }
  if (lane == 0) {
    const int64_t gi = pr_gidx ? pr_gidx[p] : p;
    fglob[3 * gi] += a0;
    fglob[3 * gi + 1] += a1;
    fglob[3 * gi + 2] += a2;
  }
}

// local_property_predict (src/local_properties.f90:174-297) from the block's
// GAP contraction: V_s = max(0, E_s) (the model's e0 is V0), and per pair
// V_der = dV_i/dq . dq/dr = -Qss_i . soap_cart_der(:, :, p) (Qss here is
// -dE/dq), zero for a centre whose V is 0. One wavefront per pair.
__global__ __launch_bounds__(256) void k_local_property(int n_sites, int D, int Dpad, const double* __restrict__ e_site,
                                                         const double* __restrict__ qss, const double* __restrict__ der,
                                                         int64_t n_pairs, const int* __restrict__ pair_site,
                                                         double* __restrict__ V, double* __restrict__ vder) {
  const int lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (p < n_sites && lane == 0) V[p] = e_site[p] < 0.0 ? 0.0 : e_site[p];
  if (p >= n_pairs || !vder) return;
  const int i = pair_site[p];
  double a0 = 0, a1 = 0, a2 = 0;
  if (e_site[i] > 0.0) {
    const double* s = qss + (size_t)i * Dpad;
    const double* g = der + (size_t)p * 3 * D;
    for (int d = lane; d < D; d += 64) {
      const double sv = s[d];
      a0 += sv * g[3 * d];
      a1 += sv * g[3 * d + 1];
      a2 += sv * g[3 * d + 2];
    }
    for (int o = 32; o > 0; o >>= 1) {
      a0 += __shfl_xor(a0, o);
      a1 += __shfl_xor(a1, o);
      a2 += __shfl_xor(a2, o);
    }
  }
  if (lane == 0) {
    vder[3 * p] = -a0;
    vder[3 * p + 1] = -a1;
    vder[3 * p + 2] = -a2;
  }
}

__global__ void k_pair_site(int n_sites, const int64_t* __restrict__ site_beg, int* __restrict__ pair_site) {
  const int i = blockIdx.x;
  if (i >= n_sites) return;
  for (int64_t p = site_beg[i] + threadIdx.x; p < site_beg[i + 1]; p += blockDim.x) pair_site[p] = i;
}

// k_pair_site for offsets the caller supplied: entries at or beyond n_pairs
// (offsets summing past it) are not written.
__global__ void k_pair_site_clip(int n_sites, const int64_t* __restrict__ site_beg, int64_t n_pairs,
                                 int* __restrict__ pair_site) {
  const int i = blockIdx.x;
  if (i >= n_sites) return;
  const int64_t e = site_beg[i + 1] < n_pairs ? site_beg[i + 1] : n_pairs;
  for (int64_t p = site_beg[i] + threadIdx.x; p < e; p += blockDim.x) pair_site[p] = i;
}

}  // namespace

// LDS bytes of the site kernels.
static size_t soap_lds_bytes(const SoapParams& P, int mode, size_t elt) {
  const int N = P.N, NLM = P.NLM, L1 = P.L + 1;
  const size_t ps = (size_t)(2 * TGX_MAX_Q + 2 * TGX_MAX_NB + 2 * L1 + 4 * NLM);
  size_t n = (size_t)N * NLM + P.D + WG;
  if (mode == 0) n += 4 * ps;
  if (mode == 1) {
    n += (size_t)N * NLM;
    size_t d_scr = ps + (size_t)N * L1 * 6 + (size_t)TGX_MAX_NB * L1 * 4 + 4;
    n += d_scr > 4 * ps ? d_scr : 4 * ps;
  }
  return n * elt;
}

static SiteArgs site_args(tgx_model_t* m) {
  SiteArgs a;
  a.site_atom = m->site_atom.as<int>();
  a.site_beg = m->site_beg.as<int64_t>();
  a.pr_spec = m->pr_spec.as<int>();
  a.pr_r = m->pr_r.as<double>();
  a.pr_xyz = m->pr_xyz.as<double>();
  a.pr_gidx = m->pr_gidx.as<int64_t>();
  return a;
}

static int check_lds(size_t bytes) {
  TGX_REQUIRE(bytes <= 160 * 1024, "SOAP kernel needs %zu bytes of LDS (more than 160 KiB): reduce n_max or l_max",
              bytes);
  return TGX_OK;
}

int tgx_soap_descriptors(tgx_ctx* ctx, tgx_model_t* m, int path, bool with_der) {
  m->der_whole = false;  // set again by a whole-block FP64 plain evaluation
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_REQUIRE(m->P.D <= 24 * WG, "descriptor size %d exceeds this build's limit %d", m->P.D, 24 * WG);
  const size_t nsp = (size_t)tgx_round_up(ns, 128);
  TGX_TRY(m->pnorm.ensure(sizeof(double) * nsp));
  SiteArgs a = site_args(m);
  TGX_REQUIRE(!m->compress || path == TGX_PATH_FP64_PLAIN || path == TGX_PATH_FP64_ADJOINT,
              "compressed SOAP blocks run on the FP64 paths only");
  if (path == TGX_PATH_FP64_PLAIN || path == TGX_PATH_FP64_ADJOINT) {
    // a compressed block's kernels write the full descriptors, compressed below
    dbuf& qout = m->compress ? m->qf64 : m->q64;
    dbuf& dout = m->compress ? m->derf : m->soap_der;
    TGX_TRY(qout.ensure(sizeof(double) * nsp * m->P.Dpad));
    if (with_der) {
      TGX_TRY(dout.ensure(sizeof(double) * (size_t)(m->n_pairs_b > 0 ? m->n_pairs_b : 1) * 3 * m->P.D));
      const SoapParams& P = m->P;
      const int NL = P.N * (P.L + 1);
      const int j_need = (3 * P.D + WG - 1) / WG;
      bool small_blocks = P.n_species <= 2;
      for (int sp = 0; sp < P.n_species; ++sp) small_blocks = small_blocks && P.sp[sp].nb <= 8 && P.sp[sp].nq <= 64;
      const int jd = (P.D + WG - 1) / WG;  // components per thread of k_soap_der_l
      if (P.L == 8 && P.N <= 16 && small_blocks && P.D < 4096) {
        using G = DerL<8>;
        const size_t cn = (size_t)P.N * P.NLM;
        size_t u = 4 * (size_t)G::SL + 48;  // four pair slots and the q . dp partials
        if (4 * cn > u) u = 4 * cn;
        if ((size_t)WG > u) u = WG;
        const size_t lds = sizeof(double) * ((size_t)G::NLM + G::L1 + 2 * cn + P.D + (P.D & 1) + u);
        TGX_TRY(check_lds(lds));
#define TGX_SOAP_DER(J)                                                                                     \
  TGX_LAUNCH(ctx, "soap_site_der_f64", (k_soap_der_l<8, J>), dim3(ns), dim3(WG), lds, m->P, a, ns, \
             qout.as<double>(), m->pnorm.as<double>(), dout.as<double>())
        if (jd <= 2) TGX_SOAP_DER(2);
        else if (jd <= 4) TGX_SOAP_DER(4);
        else if (jd <= 5) TGX_SOAP_DER(5);
        else if (jd <= 8) TGX_SOAP_DER(8);
        else TGX_SOAP_DER(16);
#undef TGX_SOAP_DER
      } else if (j_need <= 16 && NL <= 255 && P.D < 4096) {
        // coefficient, H and four pair slots; the slots overlay the spectrum and pair scratch
        const size_t head = 2 * (size_t)P.N * P.NLM + P.D;
        const size_t ps = (size_t)(2 * TGX_MAX_Q + 2 * TGX_MAX_NB + 2 * (P.L + 1) + 4 * P.NLM);
        const size_t u1 = (size_t)WG + 4 * ps;
        const size_t u2 = 4 * (size_t)(6 * NL + 8 + ps);
        const size_t lds = sizeof(double) * (head + (u1 > u2 ? u1 : u2));
        TGX_TRY(check_lds(lds));
#define TGX_SOAP_DER(J)                                                                                           \
  TGX_LAUNCH(ctx, "soap_site_der_f64", (k_soap_der<J>), dim3(ns), dim3(WG), lds, m->P, a, ns, qout.as<double>(), \
             m->pnorm.as<double>(), dout.as<double>())
        if (j_need <= 4) TGX_SOAP_DER(4);
        else if (j_need <= 8) TGX_SOAP_DER(8);
        else if (j_need <= 12) TGX_SOAP_DER(12);
        else TGX_SOAP_DER(16);
#undef TGX_SOAP_DER
      } else {
        const size_t lds = soap_lds_bytes(m->P, 1, sizeof(double));
        TGX_TRY(check_lds(lds));
        const int kmax = (m->P.D + WG - 1) / WG;
#define TGX_SOAP_DER(K)                                                                                 \
  TGX_LAUNCH(ctx, "soap_site_der_f64", (k_soap_site<double, 1, K>), dim3(ns), dim3(WG), lds, m->P, a, ns, \
             qout.as<double>(), m->pnorm.as<double>(), dout.as<double>(), (const int*)nullptr)
        if (kmax <= 8) TGX_SOAP_DER(8);
        else if (kmax <= 12) TGX_SOAP_DER(12);
        else TGX_SOAP_DER(24);
#undef TGX_SOAP_DER
      }
    } else {
      TGX_TRY(tgx_desc_wave(ctx, m, true, nullptr, ns, qout.p, m->pnorm.p));
    }
    if (m->compress) TGX_TRY(tgx_compress_descriptors(ctx, m, with_der));
  } else {
    TGX_TRY(m->q32.ensure(sizeof(float) * nsp * m->Dpad));
    if (m->P.N <= 16) {
      TGX_TRY(tgx_desc_f32(ctx, m));  // writes the BF16 copy as well
    } else {
      TGX_TRY(tgx_desc_wave(ctx, m, false, nullptr, ns, m->q32.p, m->pnorm.p));
      TGX_TRY(m->qbf.ensure(sizeof(__bf16) * nsp * m->Dpad));
      TGX_TRY(tgx_f32_to_bf16(ctx, m->q32.as<float>(), m->qbf.as<__bf16>(), (int64_t)nsp * m->Dpad));
    }
  }
  return TGX_OK;
}

int tgx_adjoint_pair_forces(tgx_ctx* ctx, tgx_model_t* m, int path) {
  const bool fp64 = path == TGX_PATH_FP64_ADJOINT || path == TGX_PATH_FP64_PLAIN;
  TGX_REQUIRE(!m->compress || fp64, "compressed SOAP blocks run on the FP64 paths only");
  if (m->compress) TGX_TRY(tgx_compress_gradient(ctx, m));
  if (!fp64 && m->P.N <= 16) return tgx_adjoint_f32(ctx, m);
  return tgx_adjoint_wave(ctx, m, fp64);
}

int tgx_plain_pair_forces(tgx_ctx* ctx, tgx_model_t* m) { return tgx_plain_pair_forces_q(ctx, m, m->qss64.as<double>()); }

int tgx_plain_pair_forces_q(tgx_ctx* ctx, tgx_model_t* m, const double* qss) {
  const int ns = m->n_sites_b;
  const int64_t np = m->n_pairs_b;
  if (ns == 0 || np == 0) return TGX_OK;
  TGX_TRY(m->pair_site.ensure(sizeof(int) * (size_t)np));
  int* pair_site = m->pair_site.as<int>();
  TGX_LAUNCH(ctx, "pair_site", k_pair_site, dim3(ns), dim3(64), 0, ns, m->site_beg.as<int64_t>(), pair_site);
  TGX_LAUNCH(ctx, "plain_force", k_plain_force, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, ns,
             m->site_beg.as<int64_t>(), m->D, m->Dpad, qss, m->soap_der.as<double>(), m->pr_gidx.as<int64_t>(), np,
             pair_site, ctx->fglob.as<double>());
  return TGX_OK;
}

// The plain pair forces on caller-supplied buffers: fpair (3, n_pairs) +=
// Qss_i . der(:, :, p) for the pairs p of every site i (offsets site_beg, which
// may sum past n_pairs: those entries are skipped), pair_site a work array of
// n_pairs ints. No allocation.
int tgx_plain_pair_forces_raw(tgx_ctx* ctx, int ns, const int64_t* site_beg, int64_t np, int D, int Dpad,
                              const double* qss, const double* der, int* pair_site, double* fpair) {
  if (ns == 0 || np == 0) return TGX_OK;
  TGX_LAUNCH(ctx, "pair_site", k_pair_site_clip, dim3(ns), dim3(64), 0, ns, site_beg, np, pair_site);
  TGX_LAUNCH(ctx, "plain_force", k_plain_force, dim3((unsigned)((np + 3) / 4)), dim3(256), 0, ns, site_beg, D, Dpad,
             qss, der, (const int64_t*)nullptr, np, (const int*)pair_site, fpair);
  return TGX_OK;
}

// V (n_sites_b) and, with derivatives, V_der (3, n_pairs_b) on the device
// (ctx->scratch) after the FP64 plain descriptors and contraction of the block.
int tgx_local_property_finish(tgx_ctx* ctx, tgx_model_t* m, bool with_der, double** d_V, double** d_vder) {
  const int ns = m->n_sites_b;
  const int64_t np = with_der ? m->n_pairs_b : 0;
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * ((size_t)ns + 3 * (size_t)np) + sizeof(int) * (size_t)np + 64));
  double* V = ctx->scratch.as<double>();
  double* vder = V + ns;
  int* pair_site = reinterpret_cast<int*>(vder + 3 * np);
  if (np > 0)
    TGX_LAUNCH(ctx, "pair_site", k_pair_site, dim3(ns), dim3(64), 0, ns, m->site_beg.as<int64_t>(), pair_site);
  const int64_t rows = np > ns ? np : ns;
  if (rows > 0)
    TGX_LAUNCH(ctx, "local_property", k_local_property, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, ns, m->D,
               m->Dpad, m->e_site.as<double>(), np > 0 ? m->qss64.as<double>() : nullptr,
               np > 0 ? m->soap_der.as<double>() : nullptr, np, pair_site, V, np > 0 ? vder : nullptr);
  *d_V = V;
  *d_vder = vder;
  return TGX_OK;
}

// FP64 descriptors of the listed sites of the current block, row r of q_out
// holding site d_sites[r] (used to re-evaluate gated sites with the GAP).
int tgx_soap_descriptors_sites(tgx_ctx* ctx, tgx_model_t* m, const int* d_sites, int n_list, double* q_out,
                               double* pnorm_out) {
  return tgx_desc_wave(ctx, m, true, d_sites, n_list, q_out, pnorm_out);
}
