// Pair-potential GAP families on the resident neighbour list:
//   distance_2b  src/gap.f90:400-563 (model: src/read_files.f90:722-742, 5238-5305)
//   core_pot     src/gap.f90:565-721 (model and spline: src/read_files.f90:5392-5510,
//                src/splines.f90:35-190)
// Same pair selection as the reference: centre i of species s1 or s2 with a
// neighbour of the other one of the pair (or the same species when s1 == s2),
// r < rcut. Energies are per centre; the pair force acts on the centre (the
// reference's forces(:, i) +=), and the virial is the half-symmetrised
// -f (x) xyz summed over pairs, times 1/2. One wavefront per centre, lanes over
// its pairs; per-centre virial rows are summed on the host in atom order.
#include "tgx_internal.h"

namespace {

struct PairModelArgs {
  int kind;  // 0 distance_2b, 1 core_pot
  int s1, s2;
  int n;     // sparse points (2b) or spline nodes (core_pot)
  double rcut, buffer, delta2, inv_sig2, e0;  // 2b
  double yp1, ypn;                            // core_pot end slopes
  const double* a;  // 2b: Qs   core_pot: x
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

extern "C" int tgx_pair_energy_and_forces(tgx_ctx* ctx, int model_id, double* energies, double* forces,
                                          double* virial) {
  TGX_CTX_CHECK(ctx);
  tgx_pair_model* pm = pair_model(ctx, model_id);
  TGX_REQUIRE(pm, "tgx_pair_energy_and_forces: no pair model %d", model_id);
  TGX_REQUIRE(ctx->have_list, "tgx_pair_energy_and_forces: no neighbour list");
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
