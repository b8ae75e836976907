// SOAP compression of a soap_turbo block (compress_soap with a
// file_compress_soap; the matrix is read as src/read_files.f90:5162-5211
// reads it, and the block's descriptor size is the compressed one).
//
// The compressed descriptor is q_c = P p / |P p|, with p the full power
// spectrum: every entry (i, j, el) adds el p(j) to component i. The kernels here
// start from the normalised full q = p / |p| and its derivatives, which the
// descriptor kernels already produce:
//
//   q_c      = P q / |P q|                      (|P p| = |p| |P q|)
//   dq_c/dr  = (I - q_c q_c^T) P dq/dr / |P q|
//
// The second line holds because dp/dr = |p| dq/dr + q d|p|/dr and the
// projector removes P q, which is parallel to q_c. By the same identity an
// adjoint gradient g_c = -dE/dq_c becomes the gradient of the full descriptor
//
//   g = P^T (g_c - (g_c . q_c) q_c) / |P q|,
//
// which the full-size adjoint kernels contract with dq/dr unchanged. Every sum
// runs in a fixed order (the matrix entries in file order within a row).
#include "tgx_internal.h"

namespace {

constexpr int CW = 256;

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum over the 256 threads of the workgroup, the same value in every thread
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum64(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// q_c rows and |P q|: one workgroup per site, the unnormalised row in LDS
__global__ __launch_bounds__(CW) void k_compress_q(int Dfp, int Dc, int Dcp, const int* __restrict__ row,
                                                   const int* __restrict__ col, const double* __restrict__ el,
                                                   const double* __restrict__ qf, double* __restrict__ q,
                                                   double* __restrict__ cnorm) {
  extern __shared__ double y[];
  __shared__ double red[4];
  const int site = blockIdx.x;
  const double* qi = qf + (size_t)site * Dfp;
  double ss = 0.0;
  for (int i = threadIdx.x; i < Dc; i += CW) {
    double acc = 0.0;
    for (int k = row[i]; k < row[i + 1]; ++k) acc += el[k] * qi[col[k]];
    y[i] = acc;
    ss += acc * acc;
  }
  const double nrm = sqrt(block_sum(ss, red));
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  double* qo = q + (size_t)site * Dcp;
  for (int i = threadIdx.x; i < Dcp; i += CW) qo[i] = i < Dc ? y[i] * inv : 0.0;
  if (threadIdx.x == 0) cnorm[site] = nrm;
}

// dq_c/dr of every pair ([pair][component][xyz]): one workgroup per site, one
// wavefront per pair; the lane that writes a component rescales it
__global__ __launch_bounds__(CW) void k_compress_der(int Df, int Dc, int Dcp, const int64_t* __restrict__ site_beg,
                                                     const int* __restrict__ row, const int* __restrict__ col,
                                                     const double* __restrict__ el, const double* __restrict__ q,
                                                     const double* __restrict__ cnorm, const double* __restrict__ derf,
                                                     double* __restrict__ der) {
  const int site = blockIdx.x;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const double nrm = cnorm[site];
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  const double* qc = q + (size_t)site * Dcp;
  const int64_t b1 = site_beg[site + 1];
  for (int64_t p = site_beg[site] + w; p < b1; p += CW / 64) {
    const double* src = derf + (size_t)p * 3 * Df;
    double* dst = der + (size_t)p * 3 * Dc;
    double dx = 0.0, dy = 0.0, dz = 0.0;
    for (int i = lane; i < Dc; i += 64) {
      double tx = 0.0, ty = 0.0, tz = 0.0;
      for (int k = row[i]; k < row[i + 1]; ++k) {
        const double e = el[k];
        const double* s3 = src + 3 * (size_t)col[k];
        tx += e * s3[0];
        ty += e * s3[1];
        tz += e * s3[2];
      }
      dst[3 * i] = tx;
      dst[3 * i + 1] = ty;
      dst[3 * i + 2] = tz;
      dx += qc[i] * tx;
      dy += qc[i] * ty;
      dz += qc[i] * tz;
    }
    dx = wave_sum64(dx);
    dy = wave_sum64(dy);
    dz = wave_sum64(dz);
    for (int i = lane; i < Dc; i += 64) {
      dst[3 * i] = (dst[3 * i] - qc[i] * dx) * inv;
      dst[3 * i + 1] = (dst[3 * i + 1] - qc[i] * dy) * inv;
      dst[3 * i + 2] = (dst[3 * i + 2] - qc[i] * dz) * inv;
    }
  }
}

// gradient of the full descriptor from the compressed one: one workgroup per
// site, the projected compressed gradient in LDS
__global__ __launch_bounds__(CW) void k_compress_grad(int Df, int Dfp, int Dc, int Dcp, const int* __restrict__ tcol,
                                                      const int* __restrict__ trow, const double* __restrict__ tel,
                                                      const double* __restrict__ q, const double* __restrict__ cnorm,
                                                      const double* __restrict__ gc, double* __restrict__ gf) {
  extern __shared__ double h[];
  __shared__ double red[4];
  const int site = blockIdx.x;
  const double* qc = q + (size_t)site * Dcp;
  const double* gi = gc + (size_t)site * Dcp;
  double d = 0.0;
  for (int i = threadIdx.x; i < Dc; i += CW) d += gi[i] * qc[i];
  const double dot = block_sum(d, red);
  const double nrm = cnorm[site];
  const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
  for (int i = threadIdx.x; i < Dc; i += CW) h[i] = (gi[i] - dot * qc[i]) * inv;
  __syncthreads();
  double* go = gf + (size_t)site * Dfp;
  for (int j = threadIdx.x; j < Dfp; j += CW) {
    double acc = 0.0;
    if (j < Df)
      for (int k = tcol[j]; k < tcol[j + 1]; ++k) acc += tel[k] * h[trow[k]];
    go[j] = acc;
  }
}

int check_compressed_lds(const tgx_model_t* m) {
  TGX_REQUIRE((size_t)m->D * sizeof(double) <= 64 * 1024,
              "compressed descriptor size %d exceeds this build's limit %d", m->D, (int)(64 * 1024 / sizeof(double)));
  return TGX_OK;
}

}  // namespace

int tgx_compress_descriptors(tgx_ctx* ctx, tgx_model_t* m, bool with_der) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_TRY(check_compressed_lds(m));
  const size_t nsp = (size_t)tgx_round_up(ns, 128);
  TGX_TRY(m->q64.ensure(sizeof(double) * nsp * m->Dpad));
  TGX_TRY(m->cnorm.ensure(sizeof(double) * nsp));
  TGX_LAUNCH(ctx, "soap_compress", k_compress_q, dim3(ns), dim3(CW), sizeof(double) * (size_t)m->D, m->P.Dpad, m->D,
             m->Dpad, m->cp_row.as<int>(), m->cp_col.as<int>(), m->cp_el.as<double>(), m->qf64.as<double>(),
             m->q64.as<double>(), m->cnorm.as<double>());
  if (with_der && m->n_pairs_b > 0) {
    TGX_TRY(m->soap_der.ensure(sizeof(double) * 3 * (size_t)m->D * (size_t)m->n_pairs_b));
    TGX_LAUNCH(ctx, "soap_compress_der", k_compress_der, dim3(ns), dim3(CW), 0, m->P.D, m->D, m->Dpad,
               m->site_beg.as<int64_t>(), m->cp_row.as<int>(), m->cp_col.as<int>(), m->cp_el.as<double>(),
               m->q64.as<double>(), m->cnorm.as<double>(), m->derf.as<double>(), m->soap_der.as<double>());
  }
  return TGX_OK;
}

int tgx_compress_gradient(tgx_ctx* ctx, tgx_model_t* m) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_TRY(check_compressed_lds(m));
  const size_t nsp = (size_t)tgx_round_up(ns, 128);
  TGX_TRY(m->qssf64.ensure(sizeof(double) * nsp * m->P.Dpad));
  TGX_LAUNCH(ctx, "soap_compress_grad", k_compress_grad, dim3(ns), dim3(CW), sizeof(double) * (size_t)m->D, m->P.D,
             m->P.Dpad, m->D, m->Dpad, m->cpt_col.as<int>(), m->cpt_row.as<int>(), m->cpt_el.as<double>(),
             m->q64.as<double>(), m->cnorm.as<double>(), m->qss64.as<double>(), m->qssf64.as<double>());
  return TGX_OK;
}
