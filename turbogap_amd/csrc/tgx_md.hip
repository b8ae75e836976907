// Molecular-dynamics step on the device, with the semantics of
// src/md.f90:226-417 (velocity_verlet with fixed components,
// berendsen_thermostat, remove_cm_vel, berendsen_barostat) and the temperature
// and pressure of src/turbogap_md.f90:281-295.
// Units: Å, fs, eV; masses in eV fs^2 / Å^2 (amu * 103.6426965268).
#include "tgx_internal.h"
#include "tgx_util.h"

#include <cmath>
#include <algorithm>
#include <vector>

namespace {

// velocities at t-dt -> t (not on the first step), then positions t -> t+dt.
// A fixed component (fix[j] != 0) gets velocity 0 after the first step and
// keeps its position; positions_prev and forces_prev are copied for every atom.
__global__ void k_vv(int n, double dt, double dt_prev, int first, const double* __restrict__ F,
                     double* __restrict__ Fprev, const double* __restrict__ mass, double* __restrict__ v,
                     double* __restrict__ x, double* __restrict__ xprev, const unsigned char* __restrict__ fix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = mass[i];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t j = 3 * (size_t)i + k;
    const bool fx = fix && fix[j];
    double vj = v[j];
    if (!first) vj = fx ? 0.0 : vj + 0.5 * (F[j] + Fprev[j]) / m * dt_prev;
    v[j] = vj;
    xprev[j] = x[j];
    Fprev[j] = F[j];
    if (!fx) x[j] = x[j] + vj * dt + 0.5 * F[j] / m * dt * dt;
  }
}

// x += M x for every atom (box_scaling's linear map, M = (gamma - I) / n_steps).
__global__ void k_linmap(int n, double m00, double m01, double m02, double m10, double m11, double m12, double m20,
                         double m21, double m22, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = x[3 * (size_t)i], b = x[3 * (size_t)i + 1], c = x[3 * (size_t)i + 2];
  x[3 * (size_t)i] = a + (m00 * a + m01 * b + m02 * c);
  x[3 * (size_t)i + 1] = b + (m10 * a + m11 * b + m12 * c);
  x[3 * (size_t)i + 2] = c + (m20 * a + m21 * b + m22 * c);
}

// Gradient descent, Barzilai-Borwein terms: positions_prev taken to the image of
// positions' convention (minimum image, orthorhombic box), then per atom
// (x - x_prev) . (F - F_prev) and |F - F_prev|^2.
__global__ void k_gd_terms(int n, double Lx, double Ly, double Lz, const double* __restrict__ x,
                           double* __restrict__ xprev, const double* __restrict__ F, const double* __restrict__ Fprev,
                           double* __restrict__ sxy, double* __restrict__ syy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double L[3] = {Lx, Ly, Lz};
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t j = 3 * (size_t)i + k;
    double d = x[j] - xprev[j];
    d = d - rint(d / L[k]) * L[k];
    const double xp = x[j] - d;
    xprev[j] = xp;
    const double y = F[j] - Fprev[j];
    a = a + (x[j] - xp) * y;
    b = b + y * y;
  }
  sxy[i] = a;
  syy[i] = b;
}

// |F_i|^2 per atom
__global__ void k_fsq(int n, const double* __restrict__ F, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = F[3 * (size_t)i], b = F[3 * (size_t)i + 1], c = F[3 * (size_t)i + 2];
  out[i] = a * a + b * b + c * c;
}

// positions_prev = x, forces_prev = F, then x = x_prev + gamma F_prev except
// fixed components.
__global__ void k_gd_update(int n3, double gamma, double* __restrict__ x, double* __restrict__ xprev,
                            const double* __restrict__ F, double* __restrict__ Fprev,
                            const unsigned char* __restrict__ fix) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n3) return;
  const double xj = x[j], fj = F[j];
  xprev[j] = xj;
  Fprev[j] = fj;
  if (!(fix && fix[j])) x[j] = xj + gamma * fj;
}

// variable_time_step: for candidate step c_k (block k), the largest squared
// displacement over atoms of v c + (0.5 F / m) c^2, evaluated as the reference
// writes it (no fused multiply-adds, so the comparisons with the target match).
__global__ __launch_bounds__(256) void k_dt_candidates(int n, const double* __restrict__ c,
                                                       const double* __restrict__ v, const double* __restrict__ F,
                                                       const double* __restrict__ mass, double* __restrict__ d2max) {
#pragma clang fp contract(off)
  __shared__ double red[256];
  const double ck = c[blockIdx.x];
  const double c2 = ck * ck;
  double best = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double m = mass[i];
    double s2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const size_t j = 3 * (size_t)i + k;
      const double x = v[j] * ck + 0.5 * F[j] / m * c2;
      s2 = s2 + x * x;
    }
    best = fmax(best, s2);
  }
  red[threadIdx.x] = best;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) d2max[blockIdx.x] = red[0];
}

// variable_time_step_adaptive. Its loop over atoms keeps the last velocity and
// force time (x/v, sqrt(2 x m/|F|)) when an atom has no velocity or no force, so
// atom i uses those of the last atom j <= i that had one. Pass 1: per block of
// 256 atoms, the last such index (or -1). Pass 2: each block takes the carry
// from the blocks before it, scans its own atoms (prefix max of the index), and
// writes the block's smallest step. Arithmetic as the reference writes it (no
// fused multiply-adds).
__device__ __forceinline__ double adt_vsq(const double* v, int i) {
#pragma clang fp contract(off)
  const double a = v[3 * (size_t)i], b = v[3 * (size_t)i + 1], c = v[3 * (size_t)i + 2];
  return a * a + b * b + c * c;
}

__global__ __launch_bounds__(256) void k_adt_last(int n, const double* __restrict__ v, const double* __restrict__ F,
                                                  int* __restrict__ last) {
  __shared__ int sv[256], sf[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  int lv = -1, lf = -1;
  if (i < n) {
    if (adt_vsq(v, i) > 0.0) lv = i;
    if (adt_vsq(F, i) > 0.0) lf = i;
  }
  sv[threadIdx.x] = lv;
  sf[threadIdx.x] = lf;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sv[threadIdx.x] = max(sv[threadIdx.x], sv[threadIdx.x + w]);
      sf[threadIdx.x] = max(sf[threadIdx.x], sf[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    last[2 * blockIdx.x] = sv[0];
    last[2 * blockIdx.x + 1] = sf[0];
  }
}

__global__ __launch_bounds__(256) void k_adt_min(int n, const double* __restrict__ v, const double* __restrict__ F,
                                                 const double* __restrict__ mass, const int* __restrict__ last,
                                                 double xmax, double emax, double dt_init,
                                                 double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ int sv[256], sf[256];
  __shared__ double sm[256];
  __shared__ int cv, cf;
  if (threadIdx.x == 0) {
    int a = -1, b = -1;
    for (int k = 0; k < (int)blockIdx.x; ++k) {
      a = max(a, last[2 * k]);
      b = max(b, last[2 * k + 1]);
    }
    cv = a;
    cf = b;
  }
  const int i = blockIdx.x * 256 + threadIdx.x;
  double vsq = 0.0, fsq = 0.0;
  if (i < n) {
    vsq = adt_vsq(v, i);
    fsq = adt_vsq(F, i);
  }
  sv[threadIdx.x] = (i < n && vsq > 0.0) ? i : -1;
  sf[threadIdx.x] = (i < n && fsq > 0.0) ? i : -1;
  __syncthreads();
  // inclusive prefix max over the block (Hillis-Steele)
  for (int o = 1; o < 256; o <<= 1) {
    const int a = (int)threadIdx.x >= o ? sv[threadIdx.x - o] : -1;
    const int b = (int)threadIdx.x >= o ? sf[threadIdx.x - o] : -1;
    __syncthreads();
    sv[threadIdx.x] = max(sv[threadIdx.x], a);
    sf[threadIdx.x] = max(sf[threadIdx.x], b);
    __syncthreads();
  }
  double dt = (double)1.0e20f;
  if (i < n) {
    const int jv = max(sv[threadIdx.x], cv), jf = max(sf[threadIdx.x], cf);
    const double dtv = jv >= 0 ? xmax / sqrt(adt_vsq(v, jv)) : dt_init;
    const double dtf = jf >= 0 ? sqrt(2.0 * xmax * mass[jf] / sqrt(adt_vsq(F, jf))) : dt_init;
    dt = fmin(dtv, dtf);
    if (emax > 0.0 && vsq * fsq > 0.0) dt = fmin(dt, emax / sqrt(vsq * fsq));
    const double m = mass[i];
    double s2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const size_t j = 3 * (size_t)i + k;
      const double x = v[j] * dt + 0.5 * F[j] / m * (dt * dt);
      s2 = s2 + x * x;
    }
    const double d = sqrt(s2);
    if (d > xmax) dt = dt * xmax / d;
  }
  sm[threadIdx.x] = dt;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// x_c *= f_c for every atom (the barostat's scaling, per Cartesian component).
__global__ void k_scale3(int n, double f0, double f1, double f2, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[3 * (size_t)i] *= f0;
  x[3 * (size_t)i + 1] *= f1;
  x[3 * (size_t)i + 2] *= f2;
}

__global__ void k_scale(int n3, double f, double* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) v[i] *= f;
}

// 0.5 m v^2 per atom and m v per atom (for the centre of mass)
__global__ void k_ke_terms(int n, const double* __restrict__ mass, const double* __restrict__ v, double* __restrict__ ke,
                           double* __restrict__ mv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double m = mass[i];
  const double a = v[3 * (size_t)i], b = v[3 * (size_t)i + 1], c = v[3 * (size_t)i + 2];
  ke[i] = 0.5 * m * (a * a + b * b + c * c);
  mv[i] = m * a;
  mv[n + i] = m * b;
  mv[2 * (size_t)n + i] = m * c;
}

// Centre-of-mass velocity without a host round trip: block b of CM_PARTS sums
// m v and m over atoms b*256 + t + k*CM_PARTS*256; k_cm_sub's blocks each add the
// partials in the same order and subtract sum(m v)/sum(m).
constexpr int CM_PARTS = 64;
__global__ __launch_bounds__(256) void k_cm_part(int n, const double* __restrict__ mass, const double* __restrict__ v,
                                                 double* __restrict__ part) {
  __shared__ double red[4][256];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += CM_PARTS * 256) {
    const double m = mass[i];
    a[0] += m * v[3 * (size_t)i];
    a[1] += m * v[3 * (size_t)i + 1];
    a[2] += m * v[3 * (size_t)i + 2];
    a[3] += m;
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) red[c][threadIdx.x] = a[c];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int c = 0; c < 4; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_cm_sub(int n, const double* __restrict__ part, double* __restrict__ v) {
  __shared__ double cm[4];
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int b = 0; b < CM_PARTS; ++b) s += part[4 * b + threadIdx.x];
    cm[threadIdx.x] = s;
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double M = cm[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) v[3 * (size_t)i + c] -= cm[c] / M;
}


__global__ void k_wrap(int n, double Lx, double Ly, double Lz, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double L[3] = {Lx, Ly, Lz};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double p = x[3 * (size_t)i + k];
    p = p - floor(p / L[k]) * L[k];
    if (p >= L[k]) p -= L[k];
    x[3 * (size_t)i + k] = p;
  }
}

// Displacement accumulated since the last list build (src/turbogap_md.f90:451-471),
// taken by minimum image so that wrapping into the box is not motion; flag[0] is
// raised when an atom has moved more than lim (half the neighbour buffer).
__global__ void k_disp(int n, double Lx, double Ly, double Lz, const double* __restrict__ x,
                       const double* __restrict__ xprev, double* __restrict__ disp, double lim2, int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double L[3] = {Lx, Ly, Lz};
  double s2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t j = 3 * (size_t)i + k;
    double d = x[j] - xprev[j];
    d = d - rint(d / L[k]) * L[k];
    const double a = disp[j] + d;
    disp[j] = a;
    s2 = s2 + a * a;
  }
  if (s2 > lim2) atomicOr(flag, 1);
}

}  // namespace

#define MD_CTX(ctx)                                                                \
  do {                                                                             \
    if (!(ctx)) { tgx_set_error("%s: null context", __func__); return TGX_ERR_ARG; } \
    TGX_CHECK_HIP(hipSetDevice((ctx)->device));                                    \
    TGX_REQUIRE((ctx)->n_sites > 0 && (ctx)->n_sites == (ctx)->n_sites_sc,          \
                "MD needs a system without supercell replicas");                   \
  } while (0)

// The MD state (velocities, masses, previous positions and forces) was set for
// the current system's atoms; the forces, when needed, are this system's.
#define MD_HAVE_STATE(ctx) ((ctx)->md_n == (ctx)->n_sites)
#define MD_HAVE_FORCES(ctx) ((ctx)->forces.p && (ctx)->results_n == (ctx)->n_sites)

extern "C" int tgx_md_set_state(tgx_ctx* ctx, const double* velocities, const double* masses) {
  MD_CTX(ctx);
  const size_t n = (size_t)ctx->n_sites;
  TGX_REQUIRE(velocities && masses, "tgx_md_set_state: null input");
  TGX_TRY(ctx->vel.ensure(sizeof(double) * 3 * n));
  TGX_TRY(ctx->mass.ensure(sizeof(double) * n));
  TGX_TRY(ctx->forces_prev.ensure(sizeof(double) * 3 * n));
  TGX_TRY(ctx->pos_prev.ensure(sizeof(double) * 3 * n));
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->vel.p, velocities, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->mass.p, masses, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
  TGX_TRY(tgx_zero(ctx, ctx->forces_prev.p, sizeof(double) * 3 * n));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  ctx->dt_prev = 0.0;
  ctx->md_n = ctx->n_sites;
  return TGX_OK;
}

extern "C" int tgx_md_get_state(tgx_ctx* ctx, double* positions, double* velocities) {
  MD_CTX(ctx);
  const size_t n = (size_t)ctx->n_sites;
  if (positions)
    TGX_CHECK_HIP(hipMemcpyAsync(positions, ctx->pos.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  if (velocities) {
    TGX_REQUIRE(MD_HAVE_STATE(ctx), "tgx_md_get_state: no velocities set for this system");
    TGX_CHECK_HIP(hipMemcpyAsync(velocities, ctx->vel.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, ctx->stream));
  }
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

extern "C" int tgx_velocity_verlet(tgx_ctx* ctx, double dt, int first_step) {
  MD_CTX(ctx);
  TGX_REQUIRE(MD_HAVE_STATE(ctx) && MD_HAVE_FORCES(ctx), "tgx_velocity_verlet: state or forces missing for this system");
  const int n = ctx->n_sites;
  TGX_LAUNCH(ctx, "md_verlet", k_vv, dim3((n + 255) / 256), dim3(256), 0, n, dt, ctx->dt_prev, first_step ? 1 : 0,
             ctx->forces.as<double>(), ctx->forces_prev.as<double>(), ctx->mass.as<double>(), ctx->vel.as<double>(),
             ctx->pos.as<double>(), ctx->pos_prev.as<double>(),
             ctx->md_fixed ? ctx->md_fix.as<unsigned char>() : (const unsigned char*)nullptr);
  ctx->dt_prev = dt;
  return TGX_OK;
}

extern "C" int tgx_md_set_fixed(tgx_ctx* ctx, const int* fix_atom) {
  MD_CTX(ctx);
  ctx->md_fixed = false;
  if (!fix_atom) return TGX_OK;
  const size_t n3 = 3 * (size_t)ctx->n_sites;
  std::vector<unsigned char> f(n3);
  bool any = false;
  for (size_t j = 0; j < n3; ++j) {
    f[j] = fix_atom[j] ? 1 : 0;
    any = any || f[j];
  }
  if (!any) return TGX_OK;
  TGX_TRY(ctx->md_fix.ensure(n3));
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->md_fix.p, f.data(), n3, hipMemcpyHostToDevice, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  ctx->md_fixed = true;
  return TGX_OK;
}

static double box_volume(const double* b) {
  return std::fabs(b[0] * (b[4] * b[8] - b[5] * b[7]) - b[1] * (b[3] * b[8] - b[5] * b[6]) +
                   b[2] * (b[3] * b[7] - b[4] * b[6]));
}

extern "C" int tgx_md_pressure(tgx_ctx* ctx, const double* virial, double temperature, double* pressure,
                               double* pressure_tensor) {
  MD_CTX(ctx);
  TGX_REQUIRE(virial, "tgx_md_pressure: null virial");
  const double kB = 8.6173303e-5, eVperA3tobar = 1602176.6208;
  const double v = box_volume(ctx->box);
  TGX_REQUIRE(v > 0.0, "tgx_md_pressure: degenerate box");
  const double kin = kB * (double)(ctx->n_sites - 1) * temperature;
  if (pressure) *pressure = (kin + (virial[0] + virial[4] + virial[8]) / 3.0) / v * eVperA3tobar;
  if (pressure_tensor) {
    for (int k = 0; k < 9; ++k) pressure_tensor[k] = virial[k] / v * eVperA3tobar;
    for (int c = 0; c < 3; ++c) pressure_tensor[4 * c] += kin / v * eVperA3tobar;
  }
  return TGX_OK;
}

extern "C" int tgx_berendsen_barostat(tgx_ctx* ctx, double P0, const double* P, const char* sym, double tau,
                                      double gamma, double dt, double* box_out) {
  MD_CTX(ctx);
  TGX_REQUIRE(P && sym && tau > 0.0 && gamma > 0.0, "tgx_berendsen_barostat: null input, tau <= 0 or gamma <= 0");
  double f[3];
  if (!strncmp(sym, "iso", 3)) {
    const double Piso = (P[0] + P[4] + P[8]) / 3.0;
    f[0] = f[1] = f[2] = std::pow(1.0 + dt / tau * 4.5e-5 / gamma * (Piso - P0), 1.0 / 3.0);
  } else if (!strncmp(sym, "diag", 4)) {
    for (int c = 0; c < 3; ++c) f[c] = std::pow(1.0 + dt / tau * 4.5e-5 / gamma * (P[4 * c] - P0), 1.0 / 3.0);
  } else {
    TGX_REQUIRE(false, "tgx_berendsen_barostat: barostat_sym must be \"iso\" or \"diag\"");
  }
  for (int c = 0; c < 3; ++c) TGX_REQUIRE(std::isfinite(f[c]) && f[c] > 0.0, "tgx_berendsen_barostat: scaling %g", f[c]);
  // lattice vectors a, b, c (rows of box): component c of each scales with f[c]
  for (int v = 0; v < 3; ++v)
    for (int c = 0; c < 3; ++c) ctx->box[3 * v + c] *= f[c];
  if (box_out)
    for (int k = 0; k < 9; ++k) box_out[k] = ctx->box[k];
  const int n = ctx->n_sites;
  TGX_LAUNCH(ctx, "md_barostat", k_scale3, dim3((n + 255) / 256), dim3(256), 0, n, f[0], f[1], f[2],
             ctx->pos.as<double>());
  return TGX_OK;
}

static int ke_and_cm(tgx_ctx* ctx, double* ekin, double cm[3], double* mtot) {
  const int n = ctx->n_sites;
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * 5 * (size_t)n));
  double* ke = ctx->scratch.as<double>();
  double* mv = ke + n;
  TGX_LAUNCH(ctx, "md_ke", k_ke_terms, dim3((n + 255) / 256), dim3(256), 0, n, ctx->mass.as<double>(),
             ctx->vel.as<double>(), ke, mv);
  TGX_TRY(tgx_device_sum(ctx, ke, n, ekin));
  if (cm) {
    for (int k = 0; k < 3; ++k) TGX_TRY(tgx_device_sum(ctx, mv + (size_t)k * n, n, &cm[k]));
    TGX_TRY(tgx_device_sum(ctx, ctx->mass.as<double>(), n, mtot));
  }
  return TGX_OK;
}

extern "C" int tgx_kinetic(tgx_ctx* ctx, double* e_kin, double* temperature) {
  MD_CTX(ctx);
  TGX_REQUIRE(MD_HAVE_STATE(ctx), "tgx_kinetic: no velocities for this system");
  double ek = 0.0;
  TGX_TRY(ke_and_cm(ctx, &ek, nullptr, nullptr));
  const double kB = 8.6173303e-5;
  if (e_kin) *e_kin = ek;
  if (temperature) *temperature = ctx->n_sites > 1 ? 2.0 / 3.0 / (double)(ctx->n_sites - 1) / kB * ek : 0.0;
  return TGX_OK;
}

extern "C" int tgx_berendsen(tgx_ctx* ctx, double T0, double T, double tau, double dt) {
  MD_CTX(ctx);
  TGX_REQUIRE(MD_HAVE_STATE(ctx) && tau > 0.0, "tgx_berendsen: no velocities for this system or tau <= 0");
  if (!(T > 0.0)) return TGX_OK;
  const double f = std::sqrt(1.0 + dt / tau * (T0 / T - 1.0));
  const int n3 = 3 * ctx->n_sites;
  TGX_LAUNCH(ctx, "md_berendsen", k_scale, dim3((n3 + 255) / 256), dim3(256), 0, n3, f, ctx->vel.as<double>());
  return TGX_OK;
}

// remove_cm_vel on the device stream, no synchronisation
static int remove_cm_device(tgx_ctx* ctx) {
  const int n = ctx->n_sites;
  TGX_TRY(ctx->md_cm.ensure(sizeof(double) * 4 * CM_PARTS));
  TGX_LAUNCH(ctx, "md_cm_part", k_cm_part, dim3(CM_PARTS), dim3(256), 0, n, ctx->mass.as<double>(),
             ctx->vel.as<double>(), ctx->md_cm.as<double>());
  TGX_LAUNCH(ctx, "md_cm_sub", k_cm_sub, dim3((n + 255) / 256), dim3(256), 0, n, ctx->md_cm.as<double>(),
             ctx->vel.as<double>());
  return TGX_OK;
}

extern "C" int tgx_remove_cm_vel(tgx_ctx* ctx) {
  MD_CTX(ctx);
  TGX_REQUIRE(MD_HAVE_STATE(ctx), "tgx_remove_cm_vel: no velocities for this system");
  return remove_cm_device(ctx);
}

extern "C" int tgx_wrap_pbc(tgx_ctx* ctx) {
  MD_CTX(ctx);
  const double* b = ctx->box;
  TGX_REQUIRE(b[1] == 0 && b[2] == 0 && b[3] == 0 && b[5] == 0 && b[6] == 0 && b[7] == 0,
              "tgx_wrap_pbc: orthorhombic boxes only");
  const int n = ctx->n_sites;
  TGX_LAUNCH(ctx, "md_wrap", k_wrap, dim3((n + 255) / 256), dim3(256), 0, n, b[0], b[4], b[8], ctx->pos.as<double>());
  return TGX_OK;
}

extern "C" int tgx_md_displacement(tgx_ctx* ctx, double neighbors_buffer, int reset, int* rebuild) {
  MD_CTX(ctx);
  TGX_REQUIRE(rebuild && neighbors_buffer >= 0.0, "tgx_md_displacement: null output or negative buffer");
  const double* b = ctx->box;
  TGX_REQUIRE(b[1] == 0 && b[2] == 0 && b[3] == 0 && b[5] == 0 && b[6] == 0 && b[7] == 0,
              "tgx_md_displacement: orthorhombic boxes only");
  TGX_REQUIRE(MD_HAVE_STATE(ctx), "tgx_md_displacement: no MD state for this system");
  const int n = ctx->n_sites;
  TGX_TRY(ctx->md_disp.ensure(sizeof(double) * 3 * (size_t)n));
  TGX_TRY(ctx->md_flag.ensure(sizeof(int)));
  if (!ctx->md_flag_host) TGX_CHECK_HIP(hipHostMalloc((void**)&ctx->md_flag_host, sizeof(int), hipHostMallocDefault));
  if (reset) {
    TGX_TRY(tgx_zero(ctx, ctx->md_disp.p, sizeof(double) * 3 * (size_t)n));
    *rebuild = 0;
    return TGX_OK;
  }
  TGX_TRY(tgx_zero(ctx, ctx->md_flag.p, sizeof(int)));
  const double lim = 0.5 * neighbors_buffer;
  TGX_LAUNCH(ctx, "md_disp", k_disp, dim3((n + 255) / 256), dim3(256), 0, n, b[0], b[4], b[8], ctx->pos.as<double>(),
             ctx->pos_prev.as<double>(), ctx->md_disp.as<double>(), lim * lim, ctx->md_flag.as<int>());
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->md_flag_host, ctx->md_flag.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *rebuild = *ctx->md_flag_host;
  return TGX_OK;
}

// One MD step after the forces, with one synchronisation: velocity Verlet (with
// fixed components), the per-atom kinetic energies and their fixed-order
// partial sums, and the neighbour-buffer displacement test, read back
// together; then, on the host's temperature, the Berendsen rescale and the
// wrap. The same kernels and sums as tgx_velocity_verlet, tgx_kinetic,
// tgx_berendsen, tgx_wrap_pbc and tgx_md_displacement called in that order
// (the displacement is taken before the wrap; by minimum image it is the
// same).
extern "C" int tgx_md_advance(tgx_ctx* ctx, double dt, int first_step, int remove_cm, double T0, double tau,
                              double neighbors_buffer, double* e_kin, double* temperature, int* rebuild) {
  MD_CTX(ctx);
  TGX_REQUIRE(MD_HAVE_STATE(ctx) && MD_HAVE_FORCES(ctx), "tgx_md_advance: state or forces missing for this system");
  TGX_REQUIRE(!(T0 > 0.0) || tau > 0.0, "tgx_md_advance: tau <= 0");
  const double* b = ctx->box;
  TGX_REQUIRE(b[1] == 0 && b[2] == 0 && b[3] == 0 && b[5] == 0 && b[6] == 0 && b[7] == 0,
              "tgx_md_advance: orthorhombic boxes only");
  const int n = ctx->n_sites;
  if (remove_cm) TGX_TRY(remove_cm_device(ctx));
  TGX_LAUNCH(ctx, "md_verlet", k_vv, dim3((n + 255) / 256), dim3(256), 0, n, dt, ctx->dt_prev, first_step ? 1 : 0,
             ctx->forces.as<double>(), ctx->forces_prev.as<double>(), ctx->mass.as<double>(), ctx->vel.as<double>(),
             ctx->pos.as<double>(), ctx->pos_prev.as<double>(),
             ctx->md_fixed ? ctx->md_fix.as<unsigned char>() : (const unsigned char*)nullptr);
  ctx->dt_prev = dt;
  // kinetic energy partials and the displacement flag in one block of device memory
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * (4 * (size_t)n + TGX_SUM_PARTS + 1)));
  double* ke = ctx->scratch.as<double>();
  double* mv = ke + n;
  double* part = ke + 4 * (size_t)n;
  int* flag = reinterpret_cast<int*>(part + TGX_SUM_PARTS);
  TGX_LAUNCH(ctx, "md_ke", k_ke_terms, dim3((n + 255) / 256), dim3(256), 0, n, ctx->mass.as<double>(),
             ctx->vel.as<double>(), ke, mv);
  TGX_TRY(tgx_device_sum_launch(ctx, ke, n, part));
  TGX_TRY(tgx_zero(ctx, flag, sizeof(double)));
  const bool buffered = neighbors_buffer > 0.0;
  if (buffered) {
    TGX_TRY(ctx->md_disp.ensure(sizeof(double) * 3 * (size_t)n));
    const double lim = 0.5 * neighbors_buffer;
    TGX_LAUNCH(ctx, "md_disp", k_disp, dim3((n + 255) / 256), dim3(256), 0, n, b[0], b[4], b[8], ctx->pos.as<double>(),
               ctx->pos_prev.as<double>(), ctx->md_disp.as<double>(), lim * lim, flag);
  }
  if (!ctx->red_host) {
    const size_t nred = 9 * 512 + (size_t)TGX_SUM_PARTS + 1;
    TGX_CHECK_HIP(hipHostMalloc((void**)&ctx->red_host, sizeof(double) * nred, hipHostMallocDefault));
  }
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->red_host, part, sizeof(double) * (TGX_SUM_PARTS + 1), hipMemcpyDeviceToHost,
                               ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  double ek = 0.0;
  for (int i = 0; i < TGX_SUM_PARTS; ++i) ek += ctx->red_host[i];
  int fl = 0;
  std::memcpy(&fl, ctx->red_host + TGX_SUM_PARTS, sizeof(int));
  const double kB = 8.6173303e-5;
  const double T = n > 1 ? 2.0 / 3.0 / (double)(n - 1) / kB * ek : 0.0;
  if (T0 > 0.0 && T > 0.0) {
    const double f = std::sqrt(1.0 + dt / tau * (T0 / T - 1.0));
    const int n3 = 3 * n;
    TGX_LAUNCH(ctx, "md_berendsen", k_scale, dim3((n3 + 255) / 256), dim3(256), 0, n3, f, ctx->vel.as<double>());
  }
  TGX_LAUNCH(ctx, "md_wrap", k_wrap, dim3((n + 255) / 256), dim3(256), 0, n, b[0], b[4], b[8], ctx->pos.as<double>());
  if (e_kin) *e_kin = ek;
  if (temperature) *temperature = T;
  if (rebuild) *rebuild = buffered ? fl : 0;
  return TGX_OK;
}

// Bussi-Donadio-Parrinello stochastic velocity rescaling (J. Chem. Phys. 126,
// 014101 (2007), eq. A7) as the reference evaluates it: sigma = ndeg kB T0 / 2,
// c = exp(-dt/tau) (0 when tau/dt <= 0.1), K' = K + (1 - c)(sigma (chi2 + g^2)/ndeg
// - K) + 2 g sqrt(K sigma/ndeg (1 - c) c). The Gaussian g and the chi-squared draw
// with ndeg - 1 degrees of freedom come from the caller.
static double bussi_resample(double kk, double T0, int ndeg, double tau, double dt, double g, double chi2) {
  const double kB = 8.6173303e-5;
  const double sigma = (double)(float)ndeg * kB * T0 / 2.0;
  const double taut = tau / dt;
  const double c = taut > 0.1 ? std::exp(-1.0 / taut) : 0.0;
  return kk + (1.0 - c) * (sigma * (chi2 + g * g) / (double)ndeg - kk) +
         2.0 * g * std::sqrt(kk * sigma / (double)ndeg * (1.0 - c) * c);
}

extern "C" int tgx_bussi_thermostat(tgx_ctx* ctx, double T0, double tau, double dt, double gauss, double chi2,
                                    double* e_kin_out) {
  MD_CTX(ctx);
  TGX_REQUIRE(MD_HAVE_STATE(ctx) && dt > 0.0 && ctx->n_sites > 1 && chi2 >= 0.0,
              "tgx_bussi_thermostat: no velocities for this system, dt <= 0, one atom or negative chi2");
  double ek = 0.0;
  TGX_TRY(ke_and_cm(ctx, &ek, nullptr, nullptr));
  const int n3 = 3 * ctx->n_sites;
  double knew = 0.0;
  if (ek > 0.0) {
    knew = bussi_resample(ek, T0, n3 - 3, tau, dt, gauss, chi2);
    TGX_REQUIRE(std::isfinite(knew) && knew >= 0.0, "tgx_bussi_thermostat: resampled kinetic energy %g", knew);
    TGX_LAUNCH(ctx, "md_bussi", k_scale, dim3((n3 + 255) / 256), dim3(256), 0, n3, std::sqrt(knew / ek),
               ctx->vel.as<double>());
  } else {
    TGX_TRY(tgx_zero(ctx, ctx->vel.p, sizeof(double) * (size_t)n3));
  }
  if (e_kin_out) *e_kin_out = knew;
  return TGX_OK;
}

extern "C" int tgx_gradient_descent(tgx_ctx* ctx, double energy, double max_opt_step, int first_step,
                                    double* gamma_out, int* restored) {
  MD_CTX(ctx);
  TGX_REQUIRE(MD_HAVE_STATE(ctx) && MD_HAVE_FORCES(ctx) && max_opt_step >= 0.0,
              "tgx_gradient_descent: no MD state or forces for this system, or negative max_opt_step");
  const double* b = ctx->box;
  TGX_REQUIRE(b[1] == 0 && b[2] == 0 && b[3] == 0 && b[5] == 0 && b[6] == 0 && b[7] == 0,
              "tgx_gradient_descent: orthorhombic boxes only");
  TGX_REQUIRE(first_step || ctx->gd_started, "tgx_gradient_descent: the first call must have first_step = 1");
  const int n = ctx->n_sites;
  const size_t n3 = 3 * (size_t)n;
  const dim3 g1((n + 255) / 256), g3((unsigned)((n3 + 255) / 256));
  double* x = ctx->pos.as<double>();
  double* F = ctx->forces.as<double>();
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * 2 * (size_t)n));
  double* t0 = ctx->scratch.as<double>();
  double* t1 = t0 + n;
  TGX_TRY(ctx->md_gd0.ensure(sizeof(double) * 2 * n3));
  double* x0 = ctx->md_gd0.as<double>();
  double* f0 = x0 + n3;
  TGX_TRY(tgx_zero(ctx, ctx->vel.p, sizeof(double) * n3));
  double gamma = 0.0;
  bool backtrack_now = false;
  if (first_step) {
    ctx->gd_started = true;
    ctx->gd_backtracking = true;
    TGX_CHECK_HIP(hipMemcpyAsync(x0, x, sizeof(double) * n3, hipMemcpyDeviceToDevice, ctx->stream));
    TGX_CHECK_HIP(hipMemcpyAsync(f0, F, sizeof(double) * n3, hipMemcpyDeviceToDevice, ctx->stream));
    ctx->gd_energy0 = energy;
    TGX_LAUNCH(ctx, "md_gd_fsq", k_fsq, g1, dim3(256), 0, n, (const double*)F, t0);
    std::vector<double> fsq(n);
    TGX_CHECK_HIP(hipMemcpyAsync(fsq.data(), t0, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    double fmax = 0.0;
    for (int i = 0; i < n; ++i) fmax = std::max(fmax, std::sqrt(fsq[i]));
    if (fmax == 0.0) gamma = 0.0;
    else if (ctx->gd_initialized) gamma = ctx->gd_gamma_back0;
    else gamma = max_opt_step / fmax;
  } else if (ctx->gd_backtracking) {
    // Armijo-Goldstein: accept the step, else go back to the start with half of it
    if (energy <= ctx->gd_energy0 - ctx->gd_gamma_prev * 0.5 * ctx->gd_m_prev) {
      ctx->gd_backtracking = false;
      ctx->gd_initialized = true;
      ctx->gd_gamma_back0 = ctx->gd_gamma_prev;
    } else {
      gamma = ctx->gd_gamma_prev * 0.5;
      TGX_CHECK_HIP(hipMemcpyAsync(x, x0, sizeof(double) * n3, hipMemcpyDeviceToDevice, ctx->stream));
      TGX_CHECK_HIP(hipMemcpyAsync(F, f0, sizeof(double) * n3, hipMemcpyDeviceToDevice, ctx->stream));
      backtrack_now = true;
    }
  }
  if (!first_step && !ctx->gd_backtracking) {
    // Barzilai-Borwein step |s.y| / |y|^2, s = x - x_prev, y = F - F_prev
    TGX_LAUNCH(ctx, "md_gd_terms", k_gd_terms, g1, dim3(256), 0, n, b[0], b[4], b[8], (const double*)x,
               ctx->pos_prev.as<double>(), (const double*)F, (const double*)ctx->forces_prev.as<double>(), t0, t1);
    double sxy = 0.0, syy = 0.0;
    TGX_TRY(tgx_device_sum(ctx, t0, n, &sxy));
    TGX_TRY(tgx_device_sum(ctx, t1, n, &syy));
    gamma = syy == 0.0 ? 0.0 : std::fabs(sxy / syy);
  }
  TGX_LAUNCH(ctx, "md_gd_update", k_gd_update, g3, dim3(256), 0, (int)n3, gamma, x, ctx->pos_prev.as<double>(),
             (const double*)F, ctx->forces_prev.as<double>(),
             ctx->md_fixed ? ctx->md_fix.as<unsigned char>() : (const unsigned char*)nullptr);
  TGX_LAUNCH(ctx, "md_gd_fsq", k_fsq, g1, dim3(256), 0, n, (const double*)F, t0);
  double m = 0.0;
  TGX_TRY(tgx_device_sum(ctx, t0, n, &m));
  ctx->gd_gamma_prev = gamma;
  ctx->gd_m_prev = m;
  if (gamma_out) *gamma_out = gamma;
  if (restored) *restored = backtrack_now ? 1 : 0;
  return TGX_OK;
}

extern "C" int tgx_variable_time_step(tgx_ctx* ctx, int init, double target_pos_step, double tau_dt, double dt0,
                                      double* dt) {
  MD_CTX(ctx);
  TGX_REQUIRE(dt && *dt > 0.0 && target_pos_step > 0.0 && MD_HAVE_STATE(ctx) && MD_HAVE_FORCES(ctx),
              "tgx_variable_time_step: no MD state or forces for this system, dt <= 0 or target_pos_step <= 0");
  constexpr int KC = 256;       // candidates per launch
  constexpr int MAX_ROUNDS = 64;  // 16384 steps of 1% from dt
  const int n = ctx->n_sites;
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * 2 * KC));
  double* dc = ctx->scratch.as<double>();
  double* dd = dc + KC;
  std::vector<double> c(KC), d2(KC);
  auto evaluate = [&](int nc) -> int {
    TGX_CHECK_HIP(hipMemcpyAsync(dc, c.data(), sizeof(double) * nc, hipMemcpyHostToDevice, ctx->stream));
    TGX_LAUNCH(ctx, "md_variable_dt", k_dt_candidates, dim3(nc), dim3(256), 0, n, (const double*)dc,
               ctx->vel.as<double>(), ctx->forces.as<double>(), ctx->mass.as<double>(), dd);
    TGX_CHECK_HIP(hipMemcpyAsync(d2.data(), dd, sizeof(double) * nc, hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nc; ++k)
      TGX_REQUIRE(std::isfinite(d2[k]), "tgx_variable_time_step: non-finite displacement at dt %g", c[k]);
    return TGX_OK;
  };
  // the displacement at dt sets the direction: shrink by 0.99 while too large,
  // else grow by 1.01 while too small (the reference's loop, candidates in batches)
  c[0] = *dt;
  TGX_TRY(evaluate(1));
  const bool too_large = std::sqrt(d2[0]) > target_pos_step;
  const double factor = too_large ? 0.99 : 1.01;
  double cur = *dt, new_dt = *dt;
  bool done = false;
  for (int round = 0; round < MAX_ROUNDS && !done; ++round) {
    for (int k = 0; k < KC; ++k) {
      c[k] = cur;
      cur = cur * factor;
    }
    TGX_TRY(evaluate(KC));
    for (int k = 0; k < KC && !done; ++k) {
      const double dmax = std::sqrt(d2[k]);
      const bool go_on = too_large ? dmax > target_pos_step : dmax < target_pos_step;
      if (!go_on) {
        new_dt = c[k];
        done = true;
      }
    }
  }
  TGX_REQUIRE(done, "tgx_variable_time_step: no step within 16384 steps of 1%% reaches target_pos_step %g",
              target_pos_step);
  if (init) {
    *dt = std::min(new_dt, dt0);
  } else {
    const double relaxed = *dt * tau_dt / (tau_dt + *dt - new_dt);
    *dt = std::min(relaxed, dt0);
  }
  return TGX_OK;
}

extern "C" int tgx_adaptive_time_step(tgx_ctx* ctx, int init, double tmin, double tmax, double xmax, double emax,
                                      double dt0, double* dt) {
  MD_CTX(ctx);
  TGX_REQUIRE(dt && MD_HAVE_STATE(ctx) && MD_HAVE_FORCES(ctx),
              "tgx_adaptive_time_step: no MD state or forces for this system");
  TGX_REQUIRE(xmax > 0.0 && tmax > 0.0 && tmin > 0.0 && tmax > tmin,
              "tgx_adaptive_time_step: xmax, tmin and tmax must be positive and tmax > tmin");
  const int n = ctx->n_sites;
  const int nb = (n + 255) / 256;
  TGX_TRY(ctx->scratch.ensure(sizeof(double) * (size_t)nb + sizeof(int) * 2 * (size_t)nb + 256));
  double* part = ctx->scratch.as<double>();
  int* last = reinterpret_cast<int*>(part + nb);
  TGX_LAUNCH(ctx, "md_adaptive_dt_last", k_adt_last, dim3(nb), dim3(256), 0, n, ctx->vel.as<double>(),
             ctx->forces.as<double>(), last);
  TGX_LAUNCH(ctx, "md_adaptive_dt", k_adt_min, dim3(nb), dim3(256), 0, n, ctx->vel.as<double>(),
             ctx->forces.as<double>(), ctx->mass.as<double>(), (const int*)last, xmax, emax, init ? dt0 : *dt, part);
  std::vector<double> h(nb);
  TGX_CHECK_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * nb, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  double dtmin = (double)1.0e20f;
  for (double x : h) dtmin = std::min(dtmin, x);
  // as the reference: its max(dtmin, tmin) is overwritten by min(dtmin, tmax),
  // so tmin does not bound the step from below
  double out = std::min(dtmin, tmax);
  if (init) out = std::min(out, dt0);
  *dt = out;
  return TGX_OK;
}

extern "C" int tgx_box_scaling(tgx_ctx* ctx, const double* gamma, int i_step, int n_steps, double* box_out) {
  MD_CTX(ctx);
  TGX_REQUIRE(gamma && n_steps > 0 && i_step >= 0, "tgx_box_scaling: null gamma or bad step counts");
  if (i_step == 0) {
    for (int k = 0; k < 9; ++k) ctx->md_box0[k] = ctx->box[k];
    ctx->md_box0_set = true;
  }
  TGX_REQUIRE(ctx->md_box0_set, "tgx_box_scaling: step 0 not taken (the starting box is saved there)");
  double M[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M[3 * r + c] = (gamma[3 * r + c] - (r == c ? 1.0 : 0.0)) / (double)n_steps;
  const int n = ctx->n_sites;
  TGX_LAUNCH(ctx, "md_box_scaling", k_linmap, dim3((n + 255) / 256), dim3(256), 0, n, M[0], M[1], M[2], M[3], M[4],
             M[5], M[6], M[7], M[8], ctx->pos.as<double>());
  // box vectors a, b, c (rows of box) = f . (a0, b0, c0), f = I + (gamma - I)(i_step + 1)/n_steps
  const double t = (double)(i_step + 1) / (double)n_steps;
  for (int v = 0; v < 3; ++v)
    for (int r = 0; r < 3; ++r) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) {
        const double f = (r == c ? 1.0 : 0.0) + (gamma[3 * r + c] - (r == c ? 1.0 : 0.0)) * t;
        s += f * ctx->md_box0[3 * v + c];
      }
      ctx->box[3 * v + r] = s;
    }
  if (box_out)
    for (int k = 0; k < 9; ++k) box_out[k] = ctx->box[k];
  return TGX_OK;
}
