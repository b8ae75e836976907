// Pair and triplet GAP families on the resident neighbour list:
//   distance_2b  src/gap.f90:400-563 (model: src/read_files.f90:722-742, 5238-5305)
//   core_pot     src/gap.f90:565-721 (model and spline: src/read_files.f90:5392-5510,
//                src/splines.f90:35-190)
//   angle_3b     src/gap.f90:724-1112 (model: src/read_files.f90:743-770, 5305-5384);
//                described above its kernel below
// Same pair selection as the reference: centre i of species s1 or s2 with a
// neighbour of the other one of the pair (or the same species when s1 == s2),
// r < rcut. Energies are per centre; the pair force acts on the centre (the
// reference's forces(:, i) +=), and the virial is the half-symmetrised
// -f (x) xyz summed over pairs, times 1/2. One wavefront per centre, lanes over
// its pairs; per-centre virial rows are summed on the host in atom order.
#include "tgx_internal.h"

namespace {

struct PairModelArgs {
  int kind;  // 0 distance_2b, 1 core_pot, 2 angle_3b
  int s1, s2;
  int n;     // sparse points (2b, 3b) or spline nodes (core_pot)
  double rcut, buffer, delta2, inv_sig2, e0;  // 2b (rcut, buffer, e0 also 3b)
  double yp1, ypn;                            // core_pot end slopes
  int sc, ktype;                              // 3b: centre species, kernel 0 "exp" / 1 "pol"
  double w1, w2, w3;                          // 3b: 1 / sigma_c^2
  const double* a;  // 2b: Qs   core_pot: x   3b: (Q1, Q2, Q3, 2 delta^2 alpha cutoff) per sparse point
  const double* b;  // 2b: alphas * cutoff   core_pot: V
  const double* c;  // core_pot: d2V/dx2
};

// spline and spline_der of src/splines.f90 at one r (x ascending): linear
// extrapolation below x[0] and above x[n-1], else the cubic on the interval
// whose upper node is the first x above r.
__device__ __forceinline__ void spline_eval(const PairModelArgs& m, double r, double& v, double& dv) {
  const double* x = m.a;
  const double* y = m.b;
  const double* y2 = m.c;
  const int n = m.n;
  if (r < x[0]) {
    v = y[0] + (r - x[0]) * m.yp1;
    dv = m.yp1;
    return;
  }
  if (r > x[n - 1]) {
    v = y[n - 1] + (r - x[n - 1]) * m.ypn;
    dv = m.ypn;
    return;
  }
  int j = 0;
  while (j < n - 2 && !(r < x[j + 1])) ++j;
  const double h = x[j + 1] - x[j];
  const double A = (x[j + 1] - r) / h, B = 1.0 - A;
  const double h26 = h * h / 6.0, h6 = h / 6.0;
  v = A * y[j] + B * y[j + 1] + (A * A * A - A) * h26 * y2[j] + (B * B * B - B) * h26 * y2[j + 1];
  dv = (-1.0 / h) * y[j] + (1.0 / h) * y[j + 1] + (1.0 - 3.0 * A * A) * h6 * y2[j] + (3.0 * B * B - 1.0) * h6 * y2[j + 1];
}

__global__ void k_pair_model(int n_sites, const int* __restrict__ n_neigh, const int64_t* __restrict__ off,
                             const int* __restrict__ species, const int* __restrict__ nspec,
                             const double* __restrict__ rjs, const double* __restrict__ xyz, PairModelArgs m,
                             double* __restrict__ e_out, double* __restrict__ f_out, double* __restrict__ w_out) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n_sites) return;
  const int si = species[i];
  double e = 0.0, f0 = 0.0, f1 = 0.0, f2 = 0.0;
  double w00 = 0.0, w01 = 0.0, w02 = 0.0, w11 = 0.0, w12 = 0.0, w22 = 0.0;
  if (si == m.s1 || si == m.s2) {
    const int64_t b = off[i];
    const int nn = n_neigh[i];
    const double pi = acos(-1.0);
    for (int t = 1 + lane; t < nn; t += 64) {
      const int64_t p = b + t;
      const int sj = nspec[p];
      if (!((si == m.s1 && sj == m.s2) || (si == m.s2 && sj == m.s1))) continue;
      const double r = rjs[p];
      if (!(r < m.rcut)) continue;
      const double x0 = xyz[3 * p], x1 = xyz[3 * p + 1], x2 = xyz[3 * p + 2];
      double g;  // pair force = g * xyz / r
      if (m.kind == 0) {
        double fc = 1.0, dfc = 0.0;
        if (!(r < m.rcut - m.buffer)) {
          const double arg = pi * (r - m.rcut + m.buffer) / m.buffer;
          fc = (cos(arg) + 1.0) / 2.0;
          dfc = pi / 2.0 / m.buffer * sin(arg);
        }
        g = 0.0;
        for (int s = 0; s < m.n; ++s) {
          const double d = r - m.a[s];
          const double ex = exp(-0.5 * d * d * m.inv_sig2);
          const double ac = m.b[s];
          e += m.delta2 * ac * fc * ex;
          g += -2.0 * m.delta2 * ac * ex * (d * m.inv_sig2 * fc + dfc);
        }
      } else {
        double v, dv;
        spline_eval(m, r, v, dv);
        e += 0.5 * v;
        g = dv;
      }
      const double fx = g * x0 / r, fy = g * x1 / r, fz = g * x2 / r;
      f0 += fx;
      f1 += fy;
      f2 += fz;
      w00 -= fx * x0;
      w11 -= fy * x1;
      w22 -= fz * x2;
      w01 -= 0.5 * (fx * x1 + fy * x0);
      w02 -= 0.5 * (fx * x2 + fz * x0);
      w12 -= 0.5 * (fy * x2 + fz * x1);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    e += __shfl_xor(e, o);
    f0 += __shfl_xor(f0, o);
    f1 += __shfl_xor(f1, o);
    f2 += __shfl_xor(f2, o);
    w00 += __shfl_xor(w00, o);
    w01 += __shfl_xor(w01, o);
    w02 += __shfl_xor(w02, o);
    w11 += __shfl_xor(w11, o);
    w12 += __shfl_xor(w12, o);
    w22 += __shfl_xor(w22, o);
  }
  if (lane == 0) {
    e_out[i] = (m.kind == 0 ? m.e0 : 0.0) + e;
    f_out[3 * (size_t)i] = f0;
    f_out[3 * (size_t)i + 1] = f1;
    f_out[3 * (size_t)i + 2] = f2;
    double* w = w_out + 6 * (size_t)i;
    w[0] = 0.5 * w00; w[1] = 0.5 * w01; w[2] = 0.5 * w02;
    w[3] = 0.5 * w11; w[4] = 0.5 * w12; w[5] = 0.5 * w22;
  }
}


// angle_3b. A centre i of species sc sees every unordered pair (j, k) of its
// neighbours with species (s1, s2) in either order and r_ij, r_ik < rcut; the
// triplet's descriptor q = (r_ij + r_ik, (r_ij - r_ik)^2, r_jk) enters
// E_i = e0 + fc(r_ij) fc(r_ik) sum_s p_s k(|(q - Qs_s) / sigma|), p_s = 2 delta^2
// alpha_s cutoff_s, k = exp(-r^2/2) ("exp") or (1 - r)^4 (4 r + 1) on r < 1
// ("pol", cov_pp with d = 3, q = 1). The triplet's energy depends on q only, so
// with K = sum_s p_s k_s and S_c = dK/dq_c the forces on i, j, k are closed
// forms in (K, S) and the pair geometry, the reference's force1/force2/force3.
//
// One wavefront per centre. The pairs that pass the species and cutoff test are
// compacted into LDS (at most MAXP; a centre with more is left to a second
// launch with a larger MAXP). Triplets (a < b over the compacted pairs) are
// spread over the lanes 64 at a time, each lane looping over the sparse points
// (uniform, scalar loads). The force on the far atom of each pair is then
// summed by the lane that owns the pair, in triplet order, from the 64 triplets
// staged in LDS: no atomics, the same sums on every run. Every list entry of
// every centre gets its pair force in fglob (zero for pairs outside the
// selection, the force on the centre itself in its self entry), which the
// deterministic reverse-pair gather turns into per-atom forces and the pair
// virial 1/2 (f x^T + x f^T) summed over entries, the reference's virial.
template <int MAXP, int KT>
__global__ __launch_bounds__(64) void k_angle_3b(const int* __restrict__ centres, const int* __restrict__ n_neigh,
                                                 const int64_t* __restrict__ off, const int* __restrict__ species,
                                                 const int* __restrict__ nspec, const double* __restrict__ rjs,
                                                 const double* __restrict__ xyz, PairModelArgs m,
                                                 const double4* __restrict__ qp, double* __restrict__ e_out,
                                                 double* __restrict__ fglob, int* __restrict__ big) {
  constexpr int NS = MAXP / 64;
  __shared__ double sx[MAXP], sy[MAXP], sz[MAXP], sr[MAXP];
  __shared__ int ssp[MAXP], sto[MAXP];
  __shared__ int ca[64], cb[64];
  __shared__ double c2[3][64], c3[3][64];
  const int lane = threadIdx.x;
  const int i = centres ? centres[blockIdx.x] : (int)blockIdx.x;
  const int64_t b = off[i];
  const int nn = n_neigh[i];
  const bool centre = species[i] == m.sc;

  int M = 0;
  if (centre) {
    for (int t0 = 1; t0 < nn; t0 += 64) {
      const int t = t0 + lane;
      bool el = false;
      int sj = 0;
      double r = 0.0;
      if (t < nn) {
        sj = nspec[b + t];
        r = rjs[b + t];
        el = (sj == m.s1 || sj == m.s2) && r < m.rcut;
      }
      const uint64_t mask = __ballot(el);
      if (el) {
        const int slot = M + __popcll(mask & ((1ull << lane) - 1ull));
        if (slot < MAXP) {
          sr[slot] = r;
          ssp[slot] = sj;
          sto[slot] = t;
          sx[slot] = xyz[3 * (b + t)];
          sy[slot] = xyz[3 * (b + t) + 1];
          sz[slot] = xyz[3 * (b + t) + 2];
        }
      }
      M += __popcll(mask);
    }
  }
  if (M > MAXP) {  // uniform: the whole wave leaves
    if (lane == 0) {
      if (!centres) big[2 + atomicAdd(&big[0], 1)] = i;
      else atomicOr(&big[1], 1);
    }
    return;
  }
  __syncthreads();

  const double pi = acos(-1.0);
  const int T = M * (M - 1) / 2;
  double e_acc = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
  double ax[NS], ay[NS], az[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) ax[k] = ay[k] = az[k] = 0.0;
  for (int t0 = 0; t0 < T; t0 += 64) {
    const int t = t0 + lane;
    int pa = -1, pb = -1;
    double f2x = 0.0, f2y = 0.0, f2z = 0.0, f3x = 0.0, f3y = 0.0, f3z = 0.0;
    if (t < T) {
      // t -> (a, b), a < b: row a starts at a (2M - a - 1) / 2
      auto row = [&](int x) { return x * (2 * M - x - 1) / 2; };
      const double c = 2.0 * M - 1.0;
      int a = (int)((c - sqrt(c * c - 8.0 * t)) * 0.5);
      a = a < 0 ? 0 : (a > M - 2 ? M - 2 : a);
      while (a < M - 2 && row(a + 1) <= t) ++a;
      while (a > 0 && row(a) > t) --a;
      const int k = t - row(a) + a + 1;
      const int sa = ssp[a], sk = ssp[k];
      if ((sa == m.s1 && sk == m.s2) || (sa == m.s2 && sk == m.s1)) {
        pa = a;
        pb = k;
        const double r12 = sr[a], r13 = sr[k];
        const double x12 = sx[a], y12 = sy[a], z12 = sz[a];
        const double x13 = sx[k], y13 = sy[k], z13 = sz[k];
        const double ux12 = x12 / r12, uy12 = y12 / r12, uz12 = z12 / r12;
        const double ux13 = x13 / r13, uy13 = y13 / r13, uz13 = z13 / r13;
        const double dx = x13 - x12, dy = y13 - y12, dz = z13 - z12;
        const double r23 = sqrt(dx * dx + dy * dy + dz * dz);
        const double ux23 = dx / r23, uy23 = dy / r23, uz23 = dz / r23;
        double fc12 = 1.0, gc12 = 0.0, fc13 = 1.0, gc13 = 0.0;
        if (!(r12 < m.rcut - m.buffer)) {
          const double arg = pi * (r12 - m.rcut + m.buffer) / m.buffer;
          fc12 = (cos(arg) + 1.0) / 2.0;
          gc12 = sin(arg) / 2.0 * pi / m.buffer;
        }
        if (!(r13 < m.rcut - m.buffer)) {
          const double arg = pi * (r13 - m.rcut + m.buffer) / m.buffer;
          fc13 = (cos(arg) + 1.0) / 2.0;
          gc13 = sin(arg) / 2.0 * pi / m.buffer;
        }
        const double q1 = r12 + r13, q2 = (r12 - r13) * (r12 - r13), q3 = r23;
        double K = 0.0, H1 = 0.0, H2 = 0.0, H3 = 0.0;  // H_c = sum_s kd_s (q_c - Q_sc)
#pragma unroll 4
        for (int s = 0; s < m.n; ++s) {
          const double4 Q = qp[s];
          const double d1 = q1 - Q.x, d2 = q2 - Q.y, d3 = q3 - Q.z;
          const double r2 = d1 * d1 * m.w1 + d2 * d2 * m.w2 + d3 * d3 * m.w3;
          double kv, kd;
          if (KT == 0) {
            kv = Q.w * exp(-0.5 * r2);
            kd = kv;  // -dk/d(r^2/2)
          } else if (r2 < 1.0) {
            const double r = sqrt(r2), om = 1.0 - r, om3 = om * om * om;
            kv = Q.w * om3 * om * (4.0 * r + 1.0);
            kd = 20.0 * Q.w * om3;  // -(dk/dr)/r
          } else {
            kv = kd = 0.0;
          }
          K += kv;
          H1 += kd * d1;
          H2 += kd * d2;
          H3 += kd * d3;
        }
        // S_c = -dK/dq_c
        const double S1 = H1 * m.w1, S2 = H2 * m.w2, S3 = H3 * m.w3;
        const double fcut = fc12 * fc13;
        const double A2 = 2.0 * (r12 - r13);
        // d fc / d x_i of each cutoff (the reference's dfcut12, dfcut13)
        const double dx12 = gc12 * ux12, dy12 = gc12 * uy12, dz12 = gc12 * uz12;
        const double dx13 = gc13 * ux13, dy13 = gc13 * uy13, dz13 = gc13 * uz13;
        e_acc += fcut * K;
        g0 -= K * (fc12 * dx13 + fc13 * dx12) + fcut * (S1 * (ux12 + ux13) + S2 * A2 * (ux12 - ux13));
        g1 -= K * (fc12 * dy13 + fc13 * dy12) + fcut * (S1 * (uy12 + uy13) + S2 * A2 * (uy12 - uy13));
        g2 -= K * (fc12 * dz13 + fc13 * dz12) + fcut * (S1 * (uz12 + uz13) + S2 * A2 * (uz12 - uz13));
        f2x = K * fc13 * dx12 - fcut * (-S1 * ux12 - S2 * A2 * ux12 + S3 * ux23);
        f2y = K * fc13 * dy12 - fcut * (-S1 * uy12 - S2 * A2 * uy12 + S3 * uy23);
        f2z = K * fc13 * dz12 - fcut * (-S1 * uz12 - S2 * A2 * uz12 + S3 * uz23);
        f3x = K * fc12 * dx13 - fcut * (-S1 * ux13 + S2 * A2 * ux13 - S3 * ux23);
        f3y = K * fc12 * dy13 - fcut * (-S1 * uy13 + S2 * A2 * uy13 - S3 * uy23);
        f3z = K * fc12 * dz13 - fcut * (-S1 * uz13 + S2 * A2 * uz13 - S3 * uz23);
      }
    }
    ca[lane] = pa;
    cb[lane] = pb;
    c2[0][lane] = f2x; c2[1][lane] = f2y; c2[2][lane] = f2z;
    c3[0][lane] = f3x; c3[1][lane] = f3y; c3[2][lane] = f3z;
    __syncthreads();
    const int cnt = T - t0 < 64 ? T - t0 : 64;
    for (int u = 0; u < cnt; ++u) {
      const int ua = ca[u], ub = cb[u];
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        const int slot = lane + 64 * k;
        if (ua == slot) {
          ax[k] += c2[0][u];
          ay[k] += c2[1][u];
          az[k] += c2[2][u];
        }
        if (ub == slot) {
          ax[k] += c3[0][u];
          ay[k] += c3[1][u];
          az[k] += c3[2][u];
        }
      }
    }
    __syncthreads();
  }

  for (int o = 32; o > 0; o >>= 1) {
    e_acc += __shfl_xor(e_acc, o);
    g0 += __shfl_xor(g0, o);
    g1 += __shfl_xor(g1, o);
    g2 += __shfl_xor(g2, o);
  }
  if (lane == 0) {
    e_out[i] = m.e0 + e_acc;
    fglob[3 * b] = g0;
    fglob[3 * b + 1] = g1;
    fglob[3 * b + 2] = g2;
  }
  // entries outside the selection (every entry of a non-centre) carry no force
  for (int t = 1 + lane; t < nn; t += 64) {
    const int sj = nspec[b + t];
    if (!(centre && (sj == m.s1 || sj == m.s2) && rjs[b + t] < m.rcut)) {
      fglob[3 * (b + t)] = 0.0;
      fglob[3 * (b + t) + 1] = 0.0;
      fglob[3 * (b + t) + 2] = 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int slot = lane + 64 * k;
    if (slot < M) {
      const int64_t p = b + sto[slot];
      fglob[3 * p] = ax[k];
      fglob[3 * p + 1] = ay[k];
      fglob[3 * p + 2] = az[k];
    }
  }
}
}  // namespace

struct tgx_pair_model {
  PairModelArgs m{};
  dbuf a, b, c;
};

static tgx_pair_model* pair_model(tgx_ctx* ctx, int id) {
  if (id < 0 || id >= (int)ctx->pair_models.size()) return nullptr;
  return ctx->pair_models[id];
}

void tgx_pair_models_release(tgx_ctx* ctx) {
  for (auto* pm : ctx->pair_models) {
    if (!pm) continue;
    pm->a.release();
    pm->b.release();
    pm->c.release();
    delete pm;
  }
  ctx->pair_models.clear();
}

static int upload(tgx_ctx* ctx, dbuf& d, const std::vector<double>& v) {
  TGX_TRY(d.ensure(sizeof(double) * (v.empty() ? 1 : v.size())));
  if (!v.empty())
    TGX_CHECK_HIP(hipMemcpyAsync(d.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, ctx->stream));
  return TGX_OK;
}

static int add_model(tgx_ctx* ctx, tgx_pair_model* pm, const std::vector<double>& a, const std::vector<double>& b,
                     const std::vector<double>& c, int* model_id) {
  int rc = upload(ctx, pm->a, a);
  if (rc == TGX_OK) rc = upload(ctx, pm->b, b);
  if (rc == TGX_OK) rc = upload(ctx, pm->c, c);
  if (rc == TGX_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) {
    tgx_set_error("pair model upload failed");
    rc = TGX_ERR_HIP;
  }
  if (rc != TGX_OK) {
    pm->a.release(); pm->b.release(); pm->c.release();
    delete pm;
    return rc;
  }
  pm->m.a = pm->a.as<double>();
  pm->m.b = pm->b.as<double>();
  pm->m.c = pm->c.as<double>();
  ctx->pair_models.push_back(pm);
  *model_id = (int)ctx->pair_models.size() - 1;
  return TGX_OK;
}

extern "C" int tgx_distance_2b_create(tgx_ctx* ctx, int species1, int species2, double rcut, double buffer,
                                      double delta, double sigma, double e0, int n_sparse, const double* Qs,
                                      const double* alphas, const double* cutoff, int* model_id) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(model_id && Qs && alphas && cutoff && n_sparse > 0, "tgx_distance_2b_create: empty sparse set");
  TGX_REQUIRE(rcut > 0.0 && buffer > 0.0 && buffer <= rcut && sigma > 0.0,
              "tgx_distance_2b_create: need rcut > 0, 0 < buffer <= rcut, sigma > 0");
  auto* pm = new tgx_pair_model();
  PairModelArgs& m = pm->m;
  m.kind = 0;
  m.s1 = species1;
  m.s2 = species2;
  m.n = n_sparse;
  m.rcut = rcut;
  m.buffer = buffer;
  m.delta2 = delta * delta;
  m.inv_sig2 = 1.0 / (sigma * sigma);
  m.e0 = e0;
  std::vector<double> q(Qs, Qs + n_sparse), ac(n_sparse), none;
  for (int s = 0; s < n_sparse; ++s) ac[s] = alphas[s] * cutoff[s];
  return add_model(ctx, pm, q, ac, none, model_id);
}

// Second derivatives of the interpolating cubic spline through (x, V) with end
// slopes yp1, ypn (natural end where a slope is >= 0.99e30), by the tridiagonal
// sweep of src/read_files.f90:5454-5497; a natural end's slope is then taken
// from the spline itself. Nodes beyond `cut` are dropped and those in
// [cut - buf, cut] damped by (cos(pi (x - cut + buf)/buf) + 1)/2
// (src/read_files.f90:5428-5441).
extern "C" int tgx_core_pot_create(tgx_ctx* ctx, int species1, int species2, int n, const double* x_in,
                                   const double* V_in, double yp1, double ypn, double cut, double buf,
                                   int* model_id) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(model_id && x_in && V_in && n >= 2, "tgx_core_pot_create: need at least two nodes");
  std::vector<double> x, V;
  const double pi = acos(-1.0);
  for (int i = 0; i < n; ++i) {
    if (!(x_in[i] <= cut)) continue;
    x.push_back(x_in[i]);
    V.push_back(x_in[i] <= cut - buf ? V_in[i] : V_in[i] * 0.5 * (cos(pi / buf * (x_in[i] - cut + buf)) + 1.0));
  }
  const int m = (int)x.size();
  TGX_REQUIRE(m >= 2, "tgx_core_pot_create: fewer than two nodes inside the cutoff");
  for (int i = 1; i < m; ++i)
    TGX_REQUIRE(x[i] > x[i - 1], "tgx_core_pot_create: nodes must increase (node %d)", i);
  std::vector<double> y2(m), u(m);
  if (yp1 > 0.99e30) {
    y2[0] = 0.0;
    u[0] = 0.0;
  } else {
    y2[0] = -0.5;
    u[0] = (3.0 / (x[1] - x[0])) * ((V[1] - V[0]) / (x[1] - x[0]) - yp1);
  }
  for (int i = 1; i < m - 1; ++i) {
    const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
    const double p = sig * y2[i - 1] + 2.0;
    y2[i] = (sig - 1.0) / p;
    u[i] = (V[i + 1] - V[i]) / (x[i + 1] - x[i]) - (V[i] - V[i - 1]) / (x[i] - x[i - 1]);
    u[i] = (6.0 * u[i] / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
  }
  double qn = 0.0, un = 0.0;
  if (!(ypn > 0.99e30)) {
    qn = 0.5;
    un = (3.0 / (x[m - 1] - x[m - 2])) * (ypn - (V[m - 1] - V[m - 2]) / (x[m - 1] - x[m - 2]));
  }
  y2[m - 1] = (un - qn * u[m - 2]) / (qn * y2[m - 2] + 1.0);
  for (int i = m - 2; i >= 0; --i) y2[i] = y2[i] * y2[i + 1] + u[i];
  // slope of the spline at an end node (spline_der at x[0] uses the first interval, at x[m-1] the last)
  auto slope = [&](int j, double r) {
    const double h = x[j + 1] - x[j];
    const double A = (x[j + 1] - r) / h, B = 1.0 - A;
    return (-1.0 / h) * V[j] + (1.0 / h) * V[j + 1] + (1.0 - 3.0 * A * A) * (h / 6.0) * y2[j] +
           (3.0 * B * B - 1.0) * (h / 6.0) * y2[j + 1];
  };
  if (yp1 > 0.99e30) yp1 = slope(0, x[0]);
  if (ypn > 0.99e30) ypn = slope(m - 2, x[m - 1]);
  auto* pm = new tgx_pair_model();
  PairModelArgs& a = pm->m;
  a.kind = 1;
  a.s1 = species1;
  a.s2 = species2;
  a.n = m;
  a.rcut = x[m - 1];  // rcut = maxval(x)
  a.yp1 = yp1;
  a.ypn = ypn;
  return add_model(ctx, pm, x, V, y2, model_id);
}

// Qs row-major (n_sparse, 3): the three descriptor components of each sparse
// point in the order the descriptor file lists them.
extern "C" int tgx_angle_3b_create(tgx_ctx* ctx, int species_center, int species1, int species2, double rcut,
                                   double buffer, double delta, const double* sigma, double e0,
                                   const char* kernel_type, int n_sparse, const double* Qs, const double* alphas,
                                   const double* cutoff, int* model_id) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(model_id && Qs && alphas && cutoff && sigma && n_sparse > 0, "tgx_angle_3b_create: empty sparse set");
  TGX_REQUIRE(kernel_type && (!strcmp(kernel_type, "exp") || !strcmp(kernel_type, "pol")),
              "tgx_angle_3b_create: kernel_type must be \"exp\" or \"pol\"");
  TGX_REQUIRE(rcut > 0.0 && buffer > 0.0 && buffer <= rcut && sigma[0] > 0.0 && sigma[1] > 0.0 && sigma[2] > 0.0,
              "tgx_angle_3b_create: need rcut > 0, 0 < buffer <= rcut, sigma > 0");
  auto* pm = new tgx_pair_model();
  PairModelArgs& m = pm->m;
  m.kind = 2;
  m.sc = species_center;
  m.s1 = species1;
  m.s2 = species2;
  m.ktype = strcmp(kernel_type, "exp") ? 1 : 0;
  m.n = n_sparse;
  m.rcut = rcut;
  m.buffer = buffer;
  m.e0 = e0;
  m.w1 = 1.0 / (sigma[0] * sigma[0]);
  m.w2 = 1.0 / (sigma[1] * sigma[1]);
  m.w3 = 1.0 / (sigma[2] * sigma[2]);
  std::vector<double> qp(4 * (size_t)n_sparse), none;
  for (int s = 0; s < n_sparse; ++s) {
    for (int c = 0; c < 3; ++c) qp[4 * s + c] = Qs[3 * s + c];
    qp[4 * s + 3] = 2.0 * delta * delta * alphas[s] * cutoff[s];
  }
  return add_model(ctx, pm, qp, none, none, model_id);
}

template <int KT>
static int angle_3b_launch(tgx_ctx* ctx, const tgx_pair_model* pm, double* e, double* fglob, int* big) {
  const int n = ctx->n_sites;
  const double4* qp = reinterpret_cast<const double4*>(pm->m.a);
  TGX_LAUNCH(ctx, "angle_3b", (k_angle_3b<128, KT>), dim3(n), dim3(64), 0, nullptr, ctx->n_neigh.as<int>(),
             ctx->nl_off.as<int64_t>(), ctx->species_sc.as<int>(), ctx->nspec.as<int>(), ctx->rjs.as<double>(),
             ctx->xyz.as<double>(), pm->m, qp, e, fglob, big);
  int h[2] = {0, 0};
  TGX_CHECK_HIP(hipMemcpyAsync(h, big, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (h[0] > 0) {  // centres with more than 128 selected pairs
    TGX_LAUNCH(ctx, "angle_3b_large", (k_angle_3b<1024, KT>), dim3(h[0]), dim3(64), 0, big + 2,
               ctx->n_neigh.as<int>(), ctx->nl_off.as<int64_t>(), ctx->species_sc.as<int>(), ctx->nspec.as<int>(),
               ctx->rjs.as<double>(), ctx->xyz.as<double>(), pm->m, qp, e, fglob, big);
    TGX_CHECK_HIP(hipMemcpyAsync(h, big, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    TGX_REQUIRE(!h[1], "angle_3b: a centre has more than 1024 neighbours of the selected species inside rcut");
  }
  return TGX_OK;
}

static int angle_3b_run(tgx_ctx* ctx, const tgx_pair_model* pm, double* energies, double* forces, double* virial) {
  const int n = ctx->n_sites;
  const int64_t P = ctx->n_pairs;
  const size_t nd = 4 * (size_t)n + 3 * (size_t)P;
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * nd + sizeof(int) * (2 + (size_t)n)));
  double* e = ctx->scratch.as<double>();
  double* F = e + n;
  double* fglob = F + 3 * (size_t)n;
  int* big = reinterpret_cast<int*>(e + nd);
  TGX_TRY(tgx_zero(ctx, big, 2 * sizeof(int)));
  TGX_TRY(pm->m.ktype == 0 ? angle_3b_launch<0>(ctx, pm, e, fglob, big) : angle_3b_launch<1>(ctx, pm, e, fglob, big));
  TGX_TRY(tgx_gather_pair_forces(ctx, fglob, F, virial));
  if (energies) TGX_CHECK_HIP(hipMemcpyAsync(energies, e, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  if (forces) TGX_CHECK_HIP(hipMemcpyAsync(forces, F, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_pair_energy_and_forces(tgx_ctx* ctx, int model_id, double* energies, double* forces,
                                          double* virial) {
  TGX_CTX_CHECK(ctx);
  tgx_pair_model* pm = pair_model(ctx, model_id);
  TGX_REQUIRE(pm, "tgx_pair_energy_and_forces: no pair model %d", model_id);
  TGX_REQUIRE(ctx->have_list, "tgx_pair_energy_and_forces: no neighbour list");
  if (pm->m.kind == 2) return angle_3b_run(ctx, pm, energies, forces, virial);
  const int n = ctx->n_sites;
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * 10 * (size_t)n));
  double* e = ctx->scratch.as<double>();
  double* f = e + n;
  double* w = f + 3 * (size_t)n;
  TGX_LAUNCH(ctx, "pair_model", k_pair_model, dim3((n + 3) / 4), dim3(256), 0, n, ctx->n_neigh.as<int>(),
             ctx->nl_off.as<int64_t>(), ctx->species_sc.as<int>(), ctx->nspec.as<int>(), ctx->rjs.as<double>(),
             ctx->xyz.as<double>(), pm->m, e, f, w);
  std::vector<double> wh(6 * (size_t)n);
  if (energies) TGX_CHECK_HIP(hipMemcpyAsync(energies, e, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
  if (forces) TGX_CHECK_HIP(hipMemcpyAsync(forces, f, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(wh.data(), w, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (virial) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < 6; ++k) s[k] += wh[6 * (size_t)i + k];
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int k = 0; k < 9; ++k) virial[k] = s[map[k]];
  }
  return TGX_OK;
}
