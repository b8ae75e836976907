// get_soap_energy_and_forces without device allocation, for callers that cannot
// call hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation). It has no counterpart in the reference.
//
// Inputs and outputs are device arrays the caller owns; every work buffer (the
// padded descriptors, K^(zeta-1) alpha, the energy partials, Qss, the pair
// offsets and pair forces, the per-atom gather index and the virial partials)
// lives in `scratch`, a file-scope device array of TGX_NOALLOC_SCRATCH_BYTES
// that each GPU the library runs on holds until the process ends. Its device
// address is taken once per device with hipGetSymbolAddress (tgx_scratch_setup,
// called by tgx_init with the device current); host code uses that address,
// never the array's name.
//
// Every call on a device shares that device's array, which keeps the last
// call's contents: what must start at zero is cleared by a fill kernel at the
// start of each call, and calls that overlap on one device (two contexts or two
// host threads at once) race on it.
//
// The launches are those of tgx_get_soap_energy_and_forces (the same GEMM,
// energy and pair-force kernels); the forces are then gathered on the device in
// the reference's pair order (src/gap.f90:204-224: forces(:, j2) += f for
// l = 1 .. n_pairs), so they equal that entry point's to the bit, and the virial
// is a fixed-order device sum.
#include <mutex>

#include "tgx_internal.h"
#include "tgx_util.h"

#define TGX_NOALLOC_SCRATCH_BYTES (64u << 20)
#define TGX_NOALLOC_MAX_DEVICES 64
#define TGX_NOALLOC_MAX_SITES 16384  // one-workgroup scans of the offsets

static __device__ __align__(256) unsigned char scratch[TGX_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch[TGX_NOALLOC_MAX_DEVICES] = {};

constexpr int VIR_PARTS = 256;

// q_pad[i][d] = soap[i][d] for i < ns, d < D, zero in the padding (rows up to Mp,
// columns up to Dp).
__global__ void k_pad_rows(int ns, int D, int Mp, int Dp, const double* __restrict__ soap, double* __restrict__ q) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)Mp * Dp) return;
  const int i = (int)(t / Dp), d = (int)(t % Dp);
  q[t] = (i < ns && d < D) ? soap[(size_t)i * D + d] : 0.0;
}

__global__ void k_fill_i32(int64_t n, int v, int* __restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = v;
}

// Home atom of every pair, j2 = mod(neighbors_list(l) - 1, n_sites0) (0-based),
// -1 for an index below 1; counts per atom.
__global__ void k_target_count(int64_t np, const int* __restrict__ nl, int n0, int* __restrict__ tgt,
                               int* __restrict__ cnt) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= np) return;
  const int v = nl[l];
  const int j = v >= 1 ? (v - 1) % n0 : -1;
  tgt[l] = j;
  if (j >= 0) atomicAdd(&cnt[j], 1);
}

// Every pair into its atom's segment (any order; sorted below).
__global__ void k_target_place(int64_t np, const int* __restrict__ tgt, const int64_t* __restrict__ off,
                               int* __restrict__ cnt, int* __restrict__ idx) {
  const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= np) return;
  const int j = tgt[l];
  if (j < 0) return;
  const int k = atomicSub(&cnt[j], 1) - 1;
  idx[off[j] + k] = (int)l;
}

// One thread per atom: its pairs sorted by pair index, then
// forces(:, j) = sum of their pair forces in that order.
__global__ void k_target_gather(int n0, const int64_t* __restrict__ off, int* __restrict__ idx,
                                const double* __restrict__ fpair, double* __restrict__ forces) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n0) return;
  const int64_t b = off[j], e = off[j + 1];
  for (int64_t a = b + 1; a < e; ++a) {
    const int v = idx[a];
    int64_t c = a - 1;
    while (c >= b && idx[c] > v) {
      idx[c + 1] = idx[c];
      --c;
    }
    idx[c + 1] = v;
  }
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int64_t a = b; a < e; ++a) {
    const double* f = fpair + 3 * (size_t)idx[a];
    f0 = f0 + f[0];
    f1 = f1 + f[1];
    f2 = f2 + f[2];
  }
  forces[3 * (size_t)j] = f0;
  forces[3 * (size_t)j + 1] = f1;
  forces[3 * (size_t)j + 2] = f2;
}

// Virial partials: part VIR_PARTS-way over contiguous pair ranges, each a fixed
// tree; virial(k1, k2) += 1/2 (f_k1 xyz_k2 + f_k2 xyz_k1), pairs of no atom skipped.
__global__ __launch_bounds__(256) void k_virial_part(int64_t np, const double* __restrict__ fpair,
                                                     const double* __restrict__ xyz, const int* __restrict__ tgt,
                                                     double* __restrict__ part) {
  __shared__ double sh[9][256];
  const int64_t per = (np + VIR_PARTS - 1) / VIR_PARTS;
  const int64_t b = (int64_t)blockIdx.x * per, e = b + per < np ? b + per : np;
  double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t l = b + threadIdx.x; l < e; l += 256) {
    if (tgt[l] < 0) continue;
    const double* f = fpair + 3 * (size_t)l;
    const double* x = xyz + 3 * (size_t)l;
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) w[3 * k2 + k1] += 0.5 * (f[k1] * x[k2] + f[k2] * x[k1]);
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] = w[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
#pragma unroll
      for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 9) part[(size_t)blockIdx.x * 9 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_virial_final(const double* __restrict__ part, double* __restrict__ virial) {
  const int k = threadIdx.x;
  if (k >= 9) return;
  double s = 0.0;
  for (int p = 0; p < VIR_PARTS; ++p) s += part[(size_t)p * 9 + k];
  virial[k] = s;
}

// Carves the scratch array into 256-byte aligned regions.
struct Carver {
  char* base;
  size_t used = 0;
  template <class T> T* take(size_t n) {
    T* p = reinterpret_cast<T*>(base + used);
    used += (n * sizeof(T) + 255) & ~(size_t)255;
    return p;
  }
};

}  // namespace

int tgx_scratch_setup(int device) {
  TGX_REQUIRE(device >= 0 && device < TGX_NOALLOC_MAX_DEVICES, "scratch setup: device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return TGX_OK;
  void* p = nullptr;
  TGX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  TGX_REQUIRE(p != nullptr, "scratch setup: no device address for device %d", device);
  g_scratch[device] = p;
  return TGX_OK;
}

void* tgx_scratch_ptr(int device) {
  if (device < 0 || device >= TGX_NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[device];
}

size_t tgx_scratch_bytes() { return TGX_NOALLOC_SCRATCH_BYTES; }

extern "C" int tgx_get_soap_energy_and_forces_noalloc(tgx_ctx* ctx, int n_sites, int n_soap, int64_t n_pairs,
                                                      const double* d_soap, const double* d_soap_der, double delta,
                                                      double zeta0, double e0, const int* d_n_neigh,
                                                      const int* d_neighbors_list, const double* d_xyz,
                                                      int n_sites0, int do_forces, double* d_energies,
                                                      double* d_forces, double* d_virial) {
  TGX_CTX_CHECK(ctx);
  TGX_REQUIRE(ctx->current >= 0 && ctx->current < (int)ctx->models.size() && ctx->models[ctx->current],
              "no soap_turbo block selected: call tgx_soap_backend_begin first");
  const tgx_model_t* m = ctx->models[ctx->current];
  TGX_REQUIRE(n_soap == m->D, "n_soap=%d does not match the model's descriptor size %d", n_soap, m->D);
  TGX_REQUIRE(std::fabs(delta - m->delta) <= 1e-12 * std::fabs(m->delta) &&
                  std::fabs(zeta0 - m->zeta) <= 1e-12 * std::fabs(m->zeta),
              "tgx_get_soap_energy_and_forces_noalloc: delta=%g, zeta0=%g differ from the block's delta=%g, zeta=%g",
              delta, zeta0, m->delta, m->zeta);
  TGX_REQUIRE(n_sites >= 0 && n_sites <= TGX_NOALLOC_MAX_SITES && (n_sites == 0 || (d_soap && d_energies)),
              "tgx_get_soap_energy_and_forces_noalloc: n_sites=%d (at most %d) or null soap/energies", n_sites,
              TGX_NOALLOC_MAX_SITES);
  if (do_forces)
    TGX_REQUIRE(n_pairs >= 0 && n_sites0 > 0 && n_sites0 <= TGX_NOALLOC_MAX_SITES && d_forces && d_virial &&
                    (n_pairs == 0 || (d_soap_der && d_n_neigh && d_neighbors_list && d_xyz)),
                "tgx_get_soap_energy_and_forces_noalloc: forces need soap_der, lists, xyz, outputs and "
                "0 < n_sites0 <= %d",
                TGX_NOALLOC_MAX_SITES);
  if (n_sites == 0 && !do_forces) return TGX_OK;
  void* base = tgx_scratch_ptr(ctx->device);
  TGX_REQUIRE(base, "tgx_get_soap_energy_and_forces_noalloc: scratch of device %d not set up", ctx->device);
  const int Mp = (int)tgx_round_up(n_sites > 0 ? n_sites : 1, 128);
  const int Sp = m->Spad, Dp = m->Dpad;
  const int64_t P = do_forces ? n_pairs : 0;
  const int n0 = do_forces ? n_sites0 : 0;
  // the whole layout first, checked against the array's size before any launch
  Carver cv{static_cast<char*>(base)};
  double* q = cv.take<double>((size_t)Mp * Dp);
  double* kd = cv.take<double>((size_t)Mp * Sp);
  double* epart = cv.take<double>((size_t)Mp * (Sp / 128));
  double* qss = do_forces ? cv.take<double>((size_t)Mp * Dp) : nullptr;
  int64_t* site_beg = cv.take<int64_t>((size_t)n_sites + 1);
  int64_t* toff = cv.take<int64_t>((size_t)n0 + 1);
  int* pair_site = cv.take<int>((size_t)P);
  int* tgt = cv.take<int>((size_t)P);
  int* tidx = cv.take<int>((size_t)P);
  int* cnt = cv.take<int>((size_t)n0);
  double* fpair = cv.take<double>(3 * (size_t)P);
  double* vpart = cv.take<double>(9 * (size_t)VIR_PARTS);
  dbuf scan_sums;  // a view (never ensured past its size: one-workgroup scans)
  scan_sums.p = cv.take<int64_t>(64);
  scan_sums.bytes = 64 * sizeof(int64_t);
  TGX_REQUIRE(cv.used <= TGX_NOALLOC_SCRATCH_BYTES,
              "tgx_get_soap_energy_and_forces_noalloc: %d sites, %lld pairs need %zu bytes of work space, more than "
              "the %u bytes this entry point has (split the sites into smaller calls)",
              n_sites, (long long)P, cv.used, (unsigned)TGX_NOALLOC_SCRATCH_BYTES);
  if (n_sites > 0) {
    const int64_t nq = (int64_t)Mp * Dp;
    TGX_LAUNCH(ctx, "na_pad_rows", k_pad_rows, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, n_sites, m->D, Mp, Dp,
               d_soap, q);
    TGX_TRY(tgx_gap_contract_f64(ctx, m, n_sites, q, kd, epart, d_energies, qss));
  }
  if (!do_forces) return TGX_OK;
  TGX_TRY(tgx_zero(ctx, d_forces, sizeof(double) * 3 * (size_t)n0));
  if (P == 0 || n_sites == 0) {
    TGX_TRY(tgx_zero(ctx, d_virial, sizeof(double) * 9));
    return TGX_OK;
  }
  // pair offsets of the sites, pair forces
  TGX_TRY(tgx_scan_i32(ctx, d_n_neigh, site_beg, n_sites, nullptr, &scan_sums));
  const unsigned gP = (unsigned)((P + 255) / 256);
  TGX_LAUNCH(ctx, "na_fill", k_fill_i32, dim3(gP), dim3(256), 0, P, -1, pair_site);
  TGX_TRY(tgx_zero(ctx, fpair, sizeof(double) * 3 * (size_t)P));
  TGX_TRY(tgx_plain_pair_forces_raw(ctx, n_sites, site_beg, P, m->D, Dp, qss, d_soap_der, pair_site, fpair));
  // gather by home atom in pair order, and the virial
  TGX_TRY(tgx_zero(ctx, cnt, sizeof(int) * (size_t)n0));
  TGX_LAUNCH(ctx, "na_target_count", k_target_count, dim3(gP), dim3(256), 0, P, d_neighbors_list, n0, tgt, cnt);
  TGX_TRY(tgx_scan_i32(ctx, cnt, toff, n0, nullptr, &scan_sums));
  TGX_LAUNCH(ctx, "na_target_place", k_target_place, dim3(gP), dim3(256), 0, P, (const int*)tgt,
             (const int64_t*)toff, cnt, tidx);
  TGX_LAUNCH(ctx, "na_target_gather", k_target_gather, dim3((n0 + 255) / 256), dim3(256), 0, n0,
             (const int64_t*)toff, tidx, (const double*)fpair, d_forces);
  TGX_LAUNCH(ctx, "na_virial_part", k_virial_part, dim3(VIR_PARTS), dim3(256), 0, P, (const double*)fpair, d_xyz,
             (const int*)tgt, vpart);
  TGX_LAUNCH(ctx, "na_virial_final", k_virial_final, dim3(1), dim3(64), 0, (const double*)vpart, d_virial);
  return TGX_OK;
}
