// Per-block marshal (src/gap_interface.f90:204-309), per-atom force gather and
// virial/energy reductions (src/gap.f90:206-224, src/turbogap.f90:1352-1373,
// 1941-1954).
//
// Forces are gathered, not scattered with atomics: every pair force lives at its
// global pair index, and atom j sums the entries of the pairs that point at it
// by walking its own list through the reverse-pair index, in list order. The
// result is the same bits on every run.
#include "tgx_internal.h"
#include "tgx_util.h"

namespace {

struct SpeciesMap {
  int n;
  int code[TGX_MAX_SPECIES];
  __device__ int index(int c) const {
    for (int s = 0; s < n; ++s)
      if (code[s] == c) return s;
    return -1;
  }
};

// A site is a centre of the block when its species is the block's central
// species (or the block takes every atom); a pair is kept when it lies inside
// the block's largest cutoff and the block sees the neighbour's species.
// Only atoms in [lo, hi) can be centres (a batch of the block's sites).
__global__ void k_select_count(int n, int lo, int hi, SpeciesMap sm, int central, double rcut,
                               const int* __restrict__ n_neigh, const int64_t* __restrict__ off,
                               const int* __restrict__ species, const double* __restrict__ rjs,
                               const int* __restrict__ nspec, int* __restrict__ flag, int* __restrict__ npairs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sc = sm.index(species[i]);
  const bool sel = i >= lo && i < hi && n_neigh[i] > 0 && sc >= 0 && (central == 0 || sc == central - 1);
  int cnt = 0;
  if (sel) {
    const int64_t b = off[i];
    for (int k = 0; k < n_neigh[i]; ++k)
      if (rjs[b + k] < rcut && sm.index(nspec[b + k]) >= 0) ++cnt;
  }
  flag[i] = sel ? 1 : 0;
  npairs[i] = cnt;
}

__global__ __launch_bounds__(256) void k_select_fill(int n, SpeciesMap sm, double rcut, const int* __restrict__ flag,
                                                   const int64_t* __restrict__ site_idx,
                                                   const int64_t* __restrict__ pair_off,
                                                   const int* __restrict__ n_neigh, const int64_t* __restrict__ off,
                                                   const int* __restrict__ nlist, const double* __restrict__ rjs,
                                                   const double* __restrict__ xyz, const int* __restrict__ nspec,
                                                   int* __restrict__ site_atom, int64_t* __restrict__ site_beg,
                                                   int64_t* __restrict__ pr_gidx, int* __restrict__ pr_nbr,
                                                   int* __restrict__ pr_spec, double* __restrict__ pr_r,
                                                   double* __restrict__ pr_xyz, int* __restrict__ seen) {
  // two atoms per wavefront, 32 lanes each (lists average about 22 entries);
  // each half's ballot bits are masked to its own lanes
  const int lane = threadIdx.x & 63, sub = lane & 31;
  const uint64_t half = (lane < 32 ? 0xffffffffull : 0xffffffff00000000ull);
  const int i = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 2 + (lane >> 5);
  if (i >= n || !flag[i]) return;
  const int64_t s = site_idx[i];
  const int64_t o = pair_off[i];
  if (sub == 0) { site_atom[s] = i; site_beg[s] = o; }
  const int64_t b = off[i];
  const int nn = n_neigh[i];
  int kept = 0;
  for (int base = 0; base < nn; base += 32) {
    const int k = base + sub;
    bool keep = false;
    int sp = -1;
    if (k < nn) {
      sp = sm.index(nspec[b + k]);
      keep = rjs[b + k] < rcut && sp >= 0;
    }
    const uint64_t mask = __ballot(keep) & half;
    if (keep) {
      const int64_t p = o + kept + __popcll(mask & ((1ull << lane) - 1ull));
      const int64_t g = b + k;
      const int j = (nlist[g] - 1) % n;
      pr_gidx[p] = g;
      pr_nbr[p] = j;
      pr_spec[p] = sp;
      pr_r[p] = rjs[g];
      pr_xyz[3 * p] = xyz[3 * g];
      pr_xyz[3 * p + 1] = xyz[3 * g + 1];
      pr_xyz[3 * p + 2] = xyz[3 * g + 2];
      if (seen) seen[j] = 1;
    }
    kept += __popcll(mask);
  }
}

__global__ void k_set_last(int64_t* __restrict__ site_beg, int ns, const int64_t* __restrict__ pair_off, int n) {
  site_beg[ns] = pair_off[n];
}

// F_j = sum over j's list entries q of fglob[rev[q]], in list order.
// GATHER_LANES lanes per atom: lane t sums entries t, t + GATHER_LANES, ... in
// order, then the lanes' sums combine by a fixed xor butterfly, so every run
// adds in the same order.
constexpr int GATHER_LANES = 16;

__global__ void k_gather(int n, const int* __restrict__ n_neigh, const int64_t* __restrict__ off,
                         const int64_t* __restrict__ rev, const double* __restrict__ fglob, double* __restrict__ F,
                         int* __restrict__ bad) {
  const int t = threadIdx.x % GATHER_LANES;
  const int j = (blockIdx.x * blockDim.x + threadIdx.x) / GATHER_LANES;
  const bool live = j < n;
  double f0 = 0, f1 = 0, f2 = 0;
  if (live) {
    const int64_t b = off[j];
    const int nn = n_neigh[j];
    for (int k = t; k < nn; k += GATHER_LANES) {
      const int64_t r = rev[b + k];
      if (r < 0) { atomicExch(bad, 1); continue; }
      f0 += fglob[3 * r];
      f1 += fglob[3 * r + 1];
      f2 += fglob[3 * r + 2];
    }
  }
#pragma unroll
  for (int o = GATHER_LANES / 2; o > 0; o >>= 1) {  // every lane of the wavefront takes part
    f0 += __shfl_xor(f0, o);
    f1 += __shfl_xor(f1, o);
    f2 += __shfl_xor(f2, o);
  }
  if (live && t == 0) {
    F[3 * j] = f0;
    F[3 * j + 1] = f1;
    F[3 * j + 2] = f2;
  }
}

// Virial partials: 0.5 (f x^T + x f^T) summed over a fixed slice of pairs per
// block (column-major 3x3).
__global__ __launch_bounds__(256) void k_virial(int64_t P, const double* __restrict__ fglob,
                                                const double* __restrict__ xyz, double* __restrict__ part) {
  __shared__ double sh[9][256];
  const int64_t per = (P + gridDim.x - 1) / gridDim.x;
  const int64_t b = (int64_t)blockIdx.x * per, e = b + per < P ? b + per : P;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t p = b + threadIdx.x; p < e; p += 256) {
    const double f[3] = {fglob[3 * p], fglob[3 * p + 1], fglob[3 * p + 2]};
    const double x[3] = {xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]};
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1)
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) v[3 * k2 + k1] += 0.5 * (f[k1] * x[k2] + f[k2] * x[k1]);
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o)
#pragma unroll
      for (int k = 0; k < 9; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 9) part[9 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_energy_scatter(int ns, const int* __restrict__ site_atom, const double* __restrict__ e_site,
                                 double* __restrict__ energies) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < ns) energies[site_atom[s]] += e_site[s];
}

}  // namespace

static SpeciesMap species_map(const tgx_model_t* m) {
  SpeciesMap sm{};
  sm.n = m->n_species;
  for (int s = 0; s < m->n_species; ++s) sm.code[s] = m->species_types[s];
  return sm;
}

// Block selection in two halves: the count and scans with the two totals copied
// to pinned host memory (no wait), then, once the caller has synchronised the
// stream, the allocation and the fill. Several blocks can have their counts in
// flight before one wait.
struct SelectView {
  int* flag;
  int* npairs;
  int64_t* site_idx;
  int64_t* pair_off;
};

static SelectView select_view(tgx_model_t* m, int n) {
  SelectView v;
  v.flag = m->sel_scratch.as<int>();
  v.npairs = v.flag + n;
  v.site_idx = reinterpret_cast<int64_t*>(m->sel_scratch.as<char>() + tgx_round_up(sizeof(int) * 2 * (size_t)n, 16));
  v.pair_off = v.site_idx + tgx_round_up(n + 1, 2);  // 16-byte aligned for the scan
  return v;
}

int tgx_block_select_count(tgx_ctx* ctx, tgx_model_t* m) {
  TGX_REQUIRE(ctx->have_list, "no neighbour list: call tgx_build_neighbors_list first");
  const int n = ctx->n_sites;
  // the block's own scratch: blocks may be selected while another block's kernels still run
  TGX_TRY(m->sel_scratch.ensure(sizeof(int) * 2 * (size_t)n + sizeof(int64_t) * 2 * (size_t)(n + 1) + 256));
  const SelectView v = select_view(m, n);
  SpeciesMap sm = species_map(m);
  const int lo = m->sel_lo > 0 ? m->sel_lo : 0, hi = (m->sel_hi >= 0 && m->sel_hi < n) ? m->sel_hi : n;
  TGX_LAUNCH(ctx, "select_count", k_select_count, dim3((n + 255) / 256), dim3(256), 0, n, lo, hi, sm,
             m->central_species, m->rcut_max, ctx->n_neigh.as<int>(), ctx->nl_off.as<int64_t>(),
             ctx->species_sc.as<int>(), ctx->rjs.as<double>(), ctx->nspec.as<int>(), v.flag, v.npairs);
  TGX_TRY(tgx_scan2_i32(ctx, v.flag, v.site_idx, v.npairs, v.pair_off, n, m->scan_tmp));
  if (!m->sel_host) TGX_CHECK_HIP(hipHostMalloc((void**)&m->sel_host, 2 * sizeof(int64_t), hipHostMallocDefault));
  TGX_CHECK_HIP(hipMemcpyAsync(m->sel_host, v.site_idx + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(m->sel_host + 1, v.pair_off + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  return TGX_OK;
}

int tgx_block_select_counts_host(tgx_ctx* ctx, tgx_model_t* m, std::vector<int>& flag, std::vector<int64_t>& pair_off) {
  const int n = ctx->n_sites;
  const SelectView v = select_view(m, n);
  flag.resize((size_t)n);
  pair_off.resize((size_t)n + 1);
  TGX_CHECK_HIP(hipMemcpyAsync(flag.data(), v.flag, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipMemcpyAsync(pair_off.data(), v.pair_off, sizeof(int64_t) * ((size_t)n + 1), hipMemcpyDeviceToHost,
                               ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return TGX_OK;
}

int tgx_block_select_fill(tgx_ctx* ctx, tgx_model_t* m) {
  const int n = ctx->n_sites;
  const SelectView v = select_view(m, n);
  SpeciesMap sm = species_map(m);
  const int64_t ns = m->sel_host[0], np = m->sel_host[1];
  m->n_sites_b = (int)ns;
  m->n_pairs_b = np;
  const size_t npz = (size_t)(np > 0 ? np : 1);
  TGX_TRY(m->site_atom.ensure(sizeof(int) * (size_t)(ns + 1)));
  TGX_TRY(m->site_beg.ensure(sizeof(int64_t) * (size_t)(ns + 1)));
  TGX_TRY(m->pr_gidx.ensure(sizeof(int64_t) * npz));
  TGX_TRY(m->pr_nbr.ensure(sizeof(int) * npz));
  TGX_TRY(m->pr_spec.ensure(sizeof(int) * npz));
  TGX_TRY(m->pr_r.ensure(sizeof(double) * npz));
  TGX_TRY(m->pr_xyz.ensure(sizeof(double) * 3 * npz));
  TGX_LAUNCH(ctx, "select_fill", k_select_fill, dim3((n + 7) / 8), dim3(256), 0, n, sm, m->rcut_max, v.flag,
             v.site_idx, v.pair_off, ctx->n_neigh.as<int>(), ctx->nl_off.as<int64_t>(), ctx->nlist.as<int>(),
             ctx->rjs.as<double>(), ctx->xyz.as<double>(), ctx->nspec.as<int>(), m->site_atom.as<int>(),
             m->site_beg.as<int64_t>(), m->pr_gidx.as<int64_t>(), m->pr_nbr.as<int>(), m->pr_spec.as<int>(),
             m->pr_r.as<double>(), m->pr_xyz.as<double>(), (int*)nullptr);
  TGX_LAUNCH(ctx, "select_last", k_set_last, dim3(1), dim3(1), 0, m->site_beg.as<int64_t>(), (int)ns, v.pair_off, n);
  return TGX_OK;
}

int tgx_block_select(tgx_ctx* ctx, tgx_model_t* m) {
  TGX_TRY(tgx_block_select_count(ctx, m));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return tgx_block_select_fill(ctx, m);
}

// Forces and virial from the accumulated pair forces; adds every model's site
// energies was done by the caller. energy_out/virial_out may be NULL.
int tgx_gather_forces(tgx_ctx* ctx, double* energy_out, double* virial_out) {
  TGX_TRY(ctx->forces.ensure(sizeof(double) * 3 * (size_t)ctx->n_sites));
  return tgx_gather_pair_forces(ctx, ctx->fglob.as<double>(), ctx->forces.as<double>(), virial_out,
                                ctx->energies.as<double>(), energy_out);
}

// Layout of ctx->red (and its pinned host copy): 512 blocks x 9 virial
// partials, TGX_SUM_PARTS energy partials, the missing-reverse flag; one
// readback and one synchronisation for all of them.
int tgx_gather_pair_forces(tgx_ctx* ctx, const double* fglob, double* F, double* virial_out,
                           const double* d_energies, double* energy_out) {
  const int n = ctx->n_sites;
  const int64_t P = ctx->n_pairs;
  const int nbk = 512;
  const size_t nred = 9 * (size_t)nbk + TGX_SUM_PARTS + 1;
  TGX_TRY(ctx->red.ensure(sizeof(double) * nred));
  if (!ctx->red_host) TGX_CHECK_HIP(hipHostMalloc((void**)&ctx->red_host, sizeof(double) * nred, hipHostMallocDefault));
  double* vpart = ctx->red.as<double>();
  double* epart = vpart + 9 * nbk;
  int* bad = reinterpret_cast<int*>(epart + TGX_SUM_PARTS);
  TGX_TRY(tgx_zero(ctx, bad, sizeof(double)));
  TGX_LAUNCH(ctx, "force_gather", k_gather, dim3((unsigned)(((int64_t)n * GATHER_LANES + 255) / 256)), dim3(256), 0,
             n, ctx->n_neigh.as<int>(),
             ctx->nl_off.as<int64_t>(), ctx->rev.as<int64_t>(), fglob, F, bad);
  TGX_LAUNCH(ctx, "virial", k_virial, dim3(nbk), dim3(256), 0, P, fglob, ctx->xyz.as<double>(), vpart);
  const bool with_e = energy_out && d_energies;
  if (with_e) TGX_TRY(tgx_device_sum_launch(ctx, d_energies, n, epart));
  TGX_CHECK_HIP(hipMemcpyAsync(ctx->red_host, vpart, sizeof(double) * nred, hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  const double* hv = ctx->red_host;
  int badh = 0;
  std::memcpy(&badh, hv + 9 * nbk + TGX_SUM_PARTS, sizeof(int));
  if (badh) {
    tgx_set_error("force gather: a pair has no reverse entry (an atom without its own list); build with do_list = all");
    return TGX_ERR_UNSUPPORTED;
  }
  if (virial_out) {
    for (int k = 0; k < 9; ++k) virial_out[k] = 0.0;
    for (int b = 0; b < nbk; ++b)
      for (int k = 0; k < 9; ++k) virial_out[k] += hv[9 * b + k];
  }
  if (with_e) {
    double e = 0.0;
    for (int i = 0; i < TGX_SUM_PARTS; ++i) e += hv[9 * nbk + i];
    *energy_out = e;
  }
  return TGX_OK;
}

int tgx_energy_scatter(tgx_ctx* ctx, tgx_model_t* m) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  TGX_LAUNCH(ctx, "energy_scatter", k_energy_scatter, dim3((ns + 255) / 256), dim3(256), 0, ns, m->site_atom.as<int>(),
             m->e_site.as<double>(), ctx->energies.as<double>());
  return TGX_OK;
}

// Atoms a block touches: its centres and every kept neighbour.
__global__ void k_mark_seen(int64_t np, const int* __restrict__ pr_nbr, int* __restrict__ seen) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < np) seen[pr_nbr[p]] = 1;
}

int tgx_block_seen(tgx_ctx* ctx, tgx_model_t* m, int* d_seen, bool clear) {
  if (clear) TGX_TRY(tgx_zero(ctx, d_seen, sizeof(int) * (size_t)ctx->n_sites));
  const int64_t np = m->n_pairs_b;
  if (np == 0) return TGX_OK;
  TGX_LAUNCH(ctx, "mark_seen", k_mark_seen, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, np, m->pr_nbr.as<int>(),
             d_seen);
  return TGX_OK;
}
