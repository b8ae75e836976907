// Neighbour list on the device: the orthorhombic cell-list build and the
// brute-force build of src/neighbors.f90:243-605, in the reference's order.
//
// Compiled with -ffp-contract=off: rjs and xyz must come from the same IEEE
// operation sequence as the reference's get_distance (src/neighbors.f90:90-101),
// so no multiply-add may be fused.
//
// Layout (device): positions (3, Nsc) as in the reference; cell_atoms holds each
// cell's atoms in descending index order (the order the reference's head/next
// linked list walks, src/neighbors.f90:398-408,445-466); the list is CSR with
// int64 offsets.
//
// One wavefront per central atom: lanes test 64 candidates at a time in the
// reference's visiting order and compact the hits with a ballot, so the output
// order is the reference's without any per-thread serial loop.
#include "tgx_internal.h"
#include "tgx_util.h"

namespace {

struct BoxInfo {
  double a[3], b[3], c[3];
  double L[3];
  double mat_inv[9];  // triclinic branch only, row-major
  int ortho;
};

__device__ __forceinline__ double f_modulo(double a, double p) {
  // fmod(a, p) is exactly a when |a| < p (a zero quotient): skip the division
  // for the in-box differences that are almost every call
  double r = fabs(a) < p ? a : fmod(a, p);
  if (r != 0.0) {
    if ((a < 0.0) != (p < 0.0)) r = r + p;
  } else {
    r = copysign(0.0, p);
  }
  return r;
}

__device__ __forceinline__ int i_modulo(int a, int p) {
  int r = a % p;
  if (r != 0 && ((r < 0) != (p < 0))) r += p;
  return r;
}

__device__ __forceinline__ double dot3(const double* u, const double* v) {
  double s = u[0] * v[0];
  s = s + u[1] * v[1];
  s = s + u[2] * v[2];
  return s;
}

// The orthorhombic branch of src/neighbors.f90:90-101.
__device__ __forceinline__ void get_distance_ortho(const double* pi, const double* pj, const BoxInfo& bx, double* dist,
                                                   double& d) {
  double d2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double dk = f_modulo(pj[k] - pi[k], bx.L[k]);
    if (dk > bx.L[k] / 2.0) dk = dk - bx.L[k];
    dist[k] = dk;
    d2 = d2 + dk * dk;
  }
  d = sqrt(d2);
}

// src/neighbors.f90:47-163 with PBC on in every direction.
__device__ void get_distance(const double* pi, const double* pj, const BoxInfo& bx, double* dist, double& d) {
  if (bx.ortho) {
    get_distance_ortho(pi, pj, bx, dist, d);
    return;
  }
  double dd[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  double v[3] = {-1.0 * dot3(dd, bx.a), -1.0 * dot3(dd, bx.b), -1.0 * dot3(dd, bx.c)};
  int ind[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = bx.mat_inv[3 * i] * v[0];
    s = s + bx.mat_inv[3 * i + 1] * v[1];
    s = s + bx.mat_inv[3 * i + 2] * v[2];
    ind[i] = (int)llround(s);
  }
  double res_opt = 1.0e10;
  double best[3] = {dd[0], dd[1], dd[2]};
  for (int i = ind[0] - 1; i <= ind[0] + 1; ++i)
    for (int j = ind[1] - 1; j <= ind[1] + 1; ++j)
      for (int k = ind[2] - 1; k <= ind[2] + 1; ++k) {
        double t[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) t[q] = dd[q] + (double)i * bx.a[q] + (double)j * bx.b[q] + (double)k * bx.c[q];
        double res = dot3(t, t);
        if (res < res_opt) {
          res_opt = res;
          best[0] = t[0]; best[1] = t[1]; best[2] = t[2];
        }
      }
  dist[0] = best[0]; dist[1] = best[1]; dist[2] = best[2];
  d = sqrt(dot3(best, best));
}

__device__ __forceinline__ void angles(const double* dist, double d, double& th, double& ph) {
  if (fabs(dist[2]) >= d) th = dist[2] > 0.0 ? 0.0 : acos(-1.0);
  else th = acos(dist[2] / d);
  ph = atan2(dist[1], dist[0]);
}

// Cell of every unit-cell atom (src/neighbors.f90:398-408) and its 0-based
// cell coordinates (the i2-1, j2-1, k2-1 of :441-443).
__global__ void k_bin(int n, const double* __restrict__ pos, BoxInfo bx, int mx, int my, int mz,
                      int* __restrict__ cell_of, int* __restrict__ cell_xyz, int* __restrict__ cell_cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double tol = 1.0e-10;
  const double ctr[3] = {bx.a[0] / 2.0, bx.b[1] / 2.0, bx.c[2] / 2.0};
  double p[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
  double dist[3], d;
  get_distance(ctr, p, bx, dist, d);
  for (int k = 0; k < 3; ++k) dist[k] = dist[k] + ctr[k];
  int cx = (int)(dist[0] / (bx.a[0] + tol) * (double)mx);
  int cy = (int)(dist[1] / (bx.b[1] + tol) * (double)my);
  int cz = (int)(dist[2] / (bx.c[2] + tol) * (double)mz);
  int cell = i_modulo(cx, mx) + i_modulo(cy, my) * mx + i_modulo(cz, mz) * my * mx;
  cell_of[i] = cell;
  cell_xyz[3 * i] = cx;
  cell_xyz[3 * i + 1] = cy;
  cell_xyz[3 * i + 2] = cz;
  atomicAdd(&cell_cnt[cell], 1);
}

__global__ void k_place(int n, const int* __restrict__ cell_of, const int64_t* __restrict__ cell_start,
                        int* __restrict__ fill, int* __restrict__ cell_atoms) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_of[i];
  int slot = atomicAdd(&fill[c], 1);
  cell_atoms[cell_start[c] + slot] = i;
}

// Descending atom index within each cell: the walk order of the linked list.
__global__ void k_sort_cells(int n_cells, const int64_t* __restrict__ cell_start, const int* __restrict__ cell_cnt,
                             int* __restrict__ cell_atoms) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  int* a = cell_atoms + cell_start[c];
  int m = cell_cnt[c];
  for (int i = 1; i < m; ++i) {
    int v = a[i], j = i - 1;
    while (j >= 0 && a[j] < v) { a[j + 1] = a[j]; --j; }
    a[j + 1] = v;
  }
}

// Positions in cell order: cell_pos[slot] = pos[cell_atoms[slot]], so the
// candidates of a cell are read as one contiguous run.
__global__ void k_cell_pos(int n, const int* __restrict__ cell_atoms, const double* __restrict__ pos,
                           double* __restrict__ cell_pos) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int k = cell_atoms[s];
  cell_pos[3 * (size_t)s] = pos[3 * (size_t)k];
  cell_pos[3 * (size_t)s + 1] = pos[3 * (size_t)k + 1];
  cell_pos[3 * (size_t)s + 2] = pos[3 * (size_t)k + 2];
}

// Workgroup b is dispatched to XCD b % 8. Handing each XCD a contiguous run of
// logical workgroups keeps spatially close atoms (when the run is walked in
// cell order) on one XCD's L2.
__device__ __forceinline__ int xcd_block() {
  const int G = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, local = b >> 3;
  const int base = G / 8, rem = G % 8;
  return xcd * base + (xcd < rem ? xcd : rem) + local;
}

__device__ __forceinline__ void wave_sync_nb() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int NBW = 4;  // wavefronts (central atoms) per workgroup

// Cell-list traversal in the reference's order for one central atom, one
// wavefront per atom. The visited cells (src/neighbors.f90:429-446) and the
// running count of their atoms go to LDS; lane t of a round takes candidate
// base + t, finds its cell by bisection and reads the candidate's index and
// position from the cell-ordered arrays. Hits are compacted with a ballot, so
// the output order is the reference's. mode 0 counts, mode 1 fills (index,
// distance, displacement, species; the angles come from k_angles on demand).
// The count also keeps the cell slots of an atom's first NB_HITS hits, in that
// order, in hit_slots[i][.]: the fill of an atom with at most NB_HITS
// neighbours then rewrites those hits (the same distance arithmetic on the same
// operands, so the same bits) instead of visiting its 27 cells again.
constexpr int NB_HITS = 64;

// FP32 screen of the count pass: squared distances surely inside (in2) and
// surely outside (out2) the cutoff, and the box sides in FP32.
struct CountScreen {
  int on;
  float in2, out2;
  float L[3];
};

template <int MODE>
__global__ __launch_bounds__(64 * NBW) void k_cell_neigh(int n, const double* __restrict__ pos, BoxInfo bx, int mx,
                                                       int my, int mz, double rcut, const int* __restrict__ do_list,
                                                       const int* __restrict__ cell_xyz,
                                                       const int64_t* __restrict__ cell_start,
                                                       const int* __restrict__ cell_cnt,
                                                       const int* __restrict__ cell_atoms,
                                                       const double* __restrict__ cell_pos,
                                                       const int* __restrict__ species, int* __restrict__ n_neigh,
                                                       const int64_t* __restrict__ off, int* __restrict__ nlist,
                                                       double* __restrict__ rjs, double* __restrict__ xyz,
                                                       int* __restrict__ nspec, int* __restrict__ hit_slots,
                                                       CountScreen screen) {
  __shared__ int s_beg[NBW][29];  // running counts of the (up to 27) visited cells, then their number
  __shared__ int64_t s_st[NBW][27];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // central atoms in cell order (slot -> atom): neighbouring waves share candidate cells
  const int slot_i = __builtin_amdgcn_readfirstlane(xcd_block() * NBW + w);
  if (slot_i >= n) return;
  const int i = __builtin_amdgcn_readfirstlane(cell_atoms[slot_i]);
  if (do_list && !do_list[i]) {
    if (MODE == 0 && lane == 0) n_neigh[i] = 0;
    return;
  }
  if (MODE == 1) {
    const int nh = n_neigh[i] - 1;
    if (nh <= NB_HITS) {
      // the count's hits, in its order (lane h writes list entry off[i] + 1 + h)
      const int64_t o = off[i];
      if (lane == 0) {
        nlist[o] = i + 1;
        rjs[o] = 0.0;
        xyz[3 * o] = 0.0; xyz[3 * o + 1] = 0.0; xyz[3 * o + 2] = 0.0;
        nspec[o] = species[i];
      }
      if (lane < nh) {
        const int slot = hit_slots[(size_t)i * NB_HITS + lane];
        const int k = cell_atoms[slot];
        const double pi[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
        const double pk[3] = {cell_pos[3 * (size_t)slot], cell_pos[3 * (size_t)slot + 1],
                              cell_pos[3 * (size_t)slot + 2]};
        double dist[3], d;
        get_distance_ortho(pi, pk, bx, dist, d);
        const int64_t p = o + 1 + lane;
        nlist[p] = k + 1;
        rjs[p] = d;
        xyz[3 * p] = dist[0]; xyz[3 * p + 1] = dist[1]; xyz[3 * p + 2] = dist[2];
        nspec[p] = species[k];
      }
      return;
    }
  }
  int* vbeg = s_beg[w];
  int64_t* vst = s_st[w];
  const int i2 = 1 + cell_xyz[3 * i], j2 = 1 + cell_xyz[3 * i + 1], k2 = 1 + cell_xyz[3 * i + 2];
  {
    // the 27 candidate cells in visiting order (k3 outer, then j3, then i3), one
    // lane each; skipped cells are compacted out with a ballot and the running
    // atom counts come from an inclusive scan over the lanes
    bool valid = false;
    int c = 0, cnt = 0;
    int64_t st = 0;
    if (lane < 27) {
      const int k3 = k2 - 1 + lane / 9, j3 = j2 - 1 + (lane / 3) % 3, i3 = i2 - 1 + lane % 3;
      valid = true;
      if (mz == 1 && k3 != 1) valid = false;
      if (mz == 2 && k2 == 1 && k3 == 0) valid = false;
      if (mz == 2 && k2 == 2 && k3 == 3) valid = false;
      if (my == 1 && j3 != 1) valid = false;
      if (my == 2 && j2 == 1 && j3 == 0) valid = false;
      if (my == 2 && j2 == 2 && j3 == 3) valid = false;
      if (mx == 1 && i3 != 1) valid = false;
      if (mx == 2 && i2 == 1 && i3 == 0) valid = false;
      if (mx == 2 && i2 == 2 && i3 == 3) valid = false;
      if (valid) {
        c = i_modulo(i3 - 1, mx) + i_modulo(j3 - 1, my) * mx + i_modulo(k3 - 1, mz) * mx * my;
        cnt = cell_cnt[c];
        st = cell_start[c];
      }
    }
    const uint64_t vm = __ballot(valid);
    const int pos = __popcll(vm & ((1ull << lane) - 1ull));
    int incl = cnt;  // inclusive scan of the counts over lanes 0..26
    for (int o = 1; o < 32; o <<= 1) {
      const int t = __shfl_up(incl, o);
      if (lane >= o) incl += t;
    }
    if (valid) {
      vbeg[pos] = incl - cnt;
      vst[pos] = st;
    }
    const int nvv = __popcll(vm);
    if (lane == 26) {
      vbeg[nvv] = incl;
      vbeg[28] = nvv;
    }
  }
  wave_sync_nb();
  const int nv = vbeg[28];
  const int total = vbeg[nv];
  const double pi[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
  int64_t out = 0;
  if (MODE == 1) {
    out = off[i];
    if (lane == 0) {
      nlist[out] = i + 1;
      rjs[out] = 0.0;
      xyz[3 * out] = 0.0; xyz[3 * out + 1] = 0.0; xyz[3 * out + 2] = 0.0;
      nspec[out] = species[i];
    }
    out += 1;
  }
  int hits = 0;
  for (int base = 0; base < total; base += 64) {
    const int cidx = base + lane;
    bool hit = false;
    int k = -1;
    int64_t slot = 0;
    double dist[3] = {0, 0, 0}, d = 0.0;
    if (cidx < total) {
      int lo = 0, hi = nv;  // vbeg[lo] <= cidx < vbeg[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (vbeg[mid] <= cidx) lo = mid;
        else hi = mid;
      }
      slot = vst[lo] + (cidx - vbeg[lo]);
      k = cell_atoms[slot];
      if (k != i) {
        const double pk[3] = {cell_pos[3 * slot], cell_pos[3 * slot + 1], cell_pos[3 * slot + 2]};
        // the count decides candidates far from the cutoff from an FP32 minimum-image
        // estimate (screen.on: at least three cells a side, so a component near L/2
        // is far outside, and sides <= 1000 A, so the estimate's error is a tenth of
        // the margin or less); the rest take the reference's FP64 sequence
        int verdict = -1;
        if (MODE == 0 && screen.on) {
          float d2f = 0.f;
          bool wide = false;
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double a = pk[c] - pi[c];
            wide = wide || fabs(a) >= bx.L[c];
            float r = (float)a;
            if (r < 0.f) r += screen.L[c];
            if (r > 0.5f * screen.L[c]) r -= screen.L[c];
            d2f += r * r;
          }
          if (!wide) verdict = d2f < screen.in2 ? 1 : (d2f > screen.out2 ? 0 : -1);
        }
        if (verdict >= 0) {
          hit = verdict == 1;
        } else {
          get_distance_ortho(pi, pk, bx, dist, d);  // the cell list is built for orthorhombic boxes only
          hit = d < rcut;
        }
      }
    }
    const uint64_t mask = __ballot(hit);
    if (MODE == 0 && hit) {
      const int h = hits + __popcll(mask & ((1ull << lane) - 1ull));
      if (h < NB_HITS) hit_slots[(size_t)i * NB_HITS + h] = (int)slot;
    }
    if (MODE == 1 && hit) {
      const int64_t p = out + hits + __popcll(mask & ((1ull << lane) - 1ull));
      nlist[p] = k + 1;
      rjs[p] = d;
      xyz[3 * p] = dist[0]; xyz[3 * p + 1] = dist[1]; xyz[3 * p + 2] = dist[2];
      nspec[p] = species[k];
    }
    hits += __popcll(mask);
  }
  if (MODE == 0 && lane == 0) n_neigh[i] = 1 + hits;
}

// thetas / phis of every pair from its displacement (src/neighbors.f90:552-591);
// the self pair (r = 0) gets 0, 0.
__global__ void k_angles(int64_t P, const double* __restrict__ rjs, const double* __restrict__ xyz,
                         double* __restrict__ thetas, double* __restrict__ phis) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double d = rjs[p];
  if (d == 0.0) {
    thetas[p] = 0.0;
    phis[p] = 0.0;
    return;
  }
  const double dist[3] = {xyz[3 * p], xyz[3 * p + 1], xyz[3 * p + 2]};
  double th, ph;
  angles(dist, d, th, ph);
  thetas[p] = th;
  phis[p] = ph;
}

// Brute-force build over the supercell (src/neighbors.f90:473-510): small boxes
// and triclinic cells.
template <int MODE>
__global__ __launch_bounds__(256) void k_brute_neigh(int n, int n_sc, const double* __restrict__ pos, BoxInfo bx,
                                                   double rcut, const int* __restrict__ do_list,
                                                   const int* __restrict__ species, int* __restrict__ n_neigh,
                                                   const int64_t* __restrict__ off, int* __restrict__ nlist,
                                                   double* __restrict__ rjs, double* __restrict__ xyz,
                                                   int* __restrict__ nspec) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;
  if (do_list && !do_list[i]) {
    if (MODE == 0 && lane == 0) n_neigh[i] = 0;
    return;
  }
  const double pi[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
  int64_t out = 0;
  if (MODE == 1) {
    out = off[i];
    if (lane == 0) {
      nlist[out] = i + 1;
      rjs[out] = 0.0;
      xyz[3 * out] = 0.0; xyz[3 * out + 1] = 0.0; xyz[3 * out + 2] = 0.0;
      nspec[out] = species[i];
    }
    out += 1;
  }
  int hits = 0;
  for (int base = 0; base < n_sc; base += 64) {
    const int j = base + lane;
    bool hit = false;
    double dist[3] = {0, 0, 0}, d = 0.0;
    if (j < n_sc && j != i) {
      const double pj[3] = {pos[3 * (size_t)j], pos[3 * (size_t)j + 1], pos[3 * (size_t)j + 2]};
      get_distance(pi, pj, bx, dist, d);
      hit = d < rcut;
    }
    const uint64_t mask = __ballot(hit);
    if (MODE == 1 && hit) {
      const int64_t p = out + hits + __popcll(mask & ((1ull << lane) - 1ull));
      nlist[p] = j + 1;
      rjs[p] = d;
      xyz[3 * p] = dist[0]; xyz[3 * p + 1] = dist[1]; xyz[3 * p + 2] = dist[2];
      nspec[p] = species[j];
    }
    hits += __popcll(mask);
  }
  if (MODE == 0 && lane == 0) n_neigh[i] = 1 + hits;
}

// For pair p = (i -> k), the index of the pair in j = fold(k)'s list that points
// back at i with the opposite displacement; the self pair maps to itself. -1
// when j has no list. Used by the deterministic force gather.
// REV_LANES lanes per central atom (64 / REV_LANES atoms per wavefront): lists
// average about 22 entries, so a whole wavefront per atom would leave most
// lanes idle.
constexpr int REV_LANES = 32;

__global__ void k_reverse(int n, const int* __restrict__ n_neigh, const int64_t* __restrict__ off,
                          const int* __restrict__ nlist, const double* __restrict__ xyz, const int* __restrict__ order,
                          int64_t* __restrict__ rev) {
  const int lane = threadIdx.x & 63, sub = lane % REV_LANES;
  const int w = (xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (64 / REV_LANES) + lane / REV_LANES;
  if (w >= n) return;
  const int i = order ? order[w] : w;  // cell order when the cell list was built
  const int64_t b = off[i];
  const int nn = n_neigh[i];
  for (int t = sub; t < nn; t += REV_LANES) {
    const int64_t p = b + t;
    if (t == 0) { rev[p] = p; continue; }
    int j = nlist[p] - 1;
    if (j >= n) j %= n;
    const int64_t bj = off[j];
    const int nj = n_neigh[j];
    // indices first, eight loads in flight; the displacement test only decides
    // between several images of the same atom (small boxes)
    int first = -1, nmatch = 0;
    for (int u0 = 1; u0 < nj; u0 += 8) {
      int v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = u0 + k < nj ? nlist[bj + u0 + k] : 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        int jq = v[k] - 1;
        if (jq >= n) jq %= n;
        if (u0 + k < nj && jq == i) {
          if (first < 0) first = u0 + k;
          ++nmatch;
        }
      }
    }
    int64_t best = first >= 0 ? bj + first : -1;
    if (nmatch > 1) {
      double bd = 0.0;
      best = -1;
      for (int u = 1; u < nj; ++u) {
        const int64_t q = bj + u;
        int jq = nlist[q] - 1;
        if (jq >= n) jq %= n;
        if (jq != i) continue;
        double e0 = xyz[3 * q] + xyz[3 * p], e1 = xyz[3 * q + 1] + xyz[3 * p + 1], e2 = xyz[3 * q + 2] + xyz[3 * p + 2];
        double dd = e0 * e0 + e1 * e1 + e2 * e2;
        if (best < 0 || dd < bd) { best = q; bd = dd; }
      }
    }
    rev[p] = best;
  }
}

// Distances of the resident list from the current positions, as
// build_neighbors_list recomputes them for rebuild_neighbors_list = .false.
// (src/neighbors.f90:554-585): the same pairs in the same order, the self
// entry left at zero, get_distance for every other pair. One wavefront per atom.
__global__ void k_refresh(int n, const int* __restrict__ n_neigh, const int64_t* __restrict__ off,
                          const int* __restrict__ nlist, const double* __restrict__ pos, BoxInfo bx,
                          double* __restrict__ rjs, double* __restrict__ xyz) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;
  const int64_t b = off[i];
  const int nn = n_neigh[i];
  const double pi[3] = {pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2]};
  for (int t = 1 + lane; t < nn; t += 64) {
    const int64_t p = b + t;
    const int j = nlist[p] - 1;
    const double pj[3] = {pos[3 * (size_t)j], pos[3 * (size_t)j + 1], pos[3 * (size_t)j + 2]};
    double dist[3], d;
    get_distance(pi, pj, bx, dist, d);
    rjs[p] = d;
    xyz[3 * p] = dist[0];
    xyz[3 * p + 1] = dist[1];
    xyz[3 * p + 2] = dist[2];
  }
}

}  // namespace

static void host_box_info(const double* box, BoxInfo& bx) {
  for (int k = 0; k < 3; ++k) { bx.a[k] = box[k]; bx.b[k] = box[3 + k]; bx.c[k] = box[6 + k]; }
  const double tol = 1.0e-6;
  bx.ortho = fabs(bx.a[1]) < tol && fabs(bx.a[2]) < tol && fabs(bx.b[0]) < tol && fabs(bx.b[2]) < tol &&
             fabs(bx.c[0]) < tol && fabs(bx.c[1]) < tol;
  bx.L[0] = bx.a[0]; bx.L[1] = bx.b[1]; bx.L[2] = bx.c[2];
  auto dot = [](const double* u, const double* v) {
    double s = u[0] * v[0];
    s = s + u[1] * v[1];
    s = s + u[2] * v[2];
    return s;
  };
  double m11 = dot(bx.a, bx.a), m12 = dot(bx.a, bx.b), m13 = dot(bx.a, bx.c);
  double m22 = dot(bx.b, bx.b), m23 = dot(bx.b, bx.c), m33 = dot(bx.c, bx.c);
  double md = -(m13 * m13) * m22 + 2.0 * m12 * m13 * m23 - m11 * (m23 * m23) - (m12 * m12) * m33 + m11 * m22 * m33;
  double* mi = bx.mat_inv;
  mi[0] = m22 * m33 - m23 * m23;
  mi[1] = m13 * m23 - m12 * m33;
  mi[2] = m12 * m23 - m13 * m22;
  mi[3] = mi[1];
  mi[4] = m11 * m33 - m13 * m13;
  mi[5] = m12 * m13 - m11 * m23;
  mi[6] = mi[2];
  mi[7] = mi[5];
  mi[8] = m11 * m22 - m12 * m12;
  for (int k = 0; k < 9; ++k) mi[k] = mi[k] / md;
}

int tgx_neighbors_run(tgx_ctx* ctx, double rcut_max, const int* indices, const int* do_list_h) {
  const int n = ctx->n_sites, n_sc = ctx->n_sites_sc;
  TGX_REQUIRE(n > 0, "build_neighbors_list: no system set");
  TGX_REQUIRE(rcut_max > 0.0, "build_neighbors_list: rcut_max must be positive");
  // src/neighbors.f90:319-330: a box whose off-diagonal entries are below 1e-6 is
  // treated as orthorhombic and those entries are zeroed.
  BoxInfo bx;
  host_box_info(ctx->box, bx);
  if (bx.ortho) {
    ctx->box[1] = ctx->box[2] = ctx->box[3] = ctx->box[5] = ctx->box[6] = ctx->box[7] = 0.0;
    host_box_info(ctx->box, bx);
  }
  const bool small = indices && (indices[0] > 1 || indices[1] > 1 || indices[2] > 1);
  const int* d_do = nullptr;
  if (do_list_h) {
    TGX_TRY(ctx->do_list.ensure(sizeof(int) * n));
    TGX_CHECK_HIP(hipMemcpyAsync(ctx->do_list.p, do_list_h, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    d_do = ctx->do_list.as<int>();
  }
  TGX_TRY(ctx->n_neigh.ensure(sizeof(int) * n));
  TGX_TRY(ctx->nl_off.ensure(sizeof(int64_t) * (n + 1)));
  const dim3 wblk(64 * NBW);
  const dim3 wgrid((n + NBW - 1) / NBW);
  int mx = 0, my = 0, mz = 0;
  const bool cells = bx.ortho && !small;
  if (cells) {
    mx = (int)(bx.a[0] / rcut_max);
    my = (int)(bx.b[1] / rcut_max);
    mz = (int)(bx.c[2] / rcut_max);
    TGX_REQUIRE(mx >= 1 && my >= 1 && mz >= 1, "build_neighbors_list: box shorter than rcut_max needs a supercell");
    const int64_t nc = (int64_t)mx * my * mz;
    TGX_TRY(ctx->cell_of.ensure(sizeof(int) * 4 * (size_t)n));
    TGX_TRY(ctx->cell_cnt.ensure(sizeof(int) * 2 * (size_t)nc));
    TGX_TRY(ctx->cell_start.ensure(sizeof(int64_t) * (size_t)(nc + 1)));
    TGX_TRY(ctx->cell_atoms.ensure(sizeof(int) * (size_t)n));
    int* cell_of = ctx->cell_of.as<int>();
    int* cell_xyz = cell_of + n;
    int* cnt = ctx->cell_cnt.as<int>();
    int* fill = cnt + nc;
    TGX_TRY(tgx_zero(ctx, cnt, sizeof(int) * 2 * (size_t)nc));
    TGX_LAUNCH(ctx, "nbr_bin", k_bin, dim3((n + 255) / 256), dim3(256), 0, n, ctx->pos.as<double>(), bx, mx, my, mz,
               cell_of, cell_xyz, cnt);
    TGX_TRY(tgx_scan_i32(ctx, cnt, ctx->cell_start.as<int64_t>(), nc, nullptr));
    TGX_LAUNCH(ctx, "nbr_place", k_place, dim3((n + 255) / 256), dim3(256), 0, n, cell_of,
               ctx->cell_start.as<int64_t>(), fill, ctx->cell_atoms.as<int>());
    TGX_LAUNCH(ctx, "nbr_sort", k_sort_cells, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, (int)nc,
               ctx->cell_start.as<int64_t>(), cnt, ctx->cell_atoms.as<int>());
    TGX_TRY(ctx->cell_pos.ensure(sizeof(double) * 3 * (size_t)n));
    TGX_LAUNCH(ctx, "nbr_cell_pos", k_cell_pos, dim3((n + 255) / 256), dim3(256), 0, n, ctx->cell_atoms.as<int>(),
               ctx->pos.as<double>(), ctx->cell_pos.as<double>());
    TGX_TRY(ctx->nbr_hits.ensure(sizeof(int) * (size_t)n * NB_HITS));
    // margin 1e-3 rcut (2e-3 rcut^2 in the squares): the FP32 estimate's error, a
    // few ulp of the side per component, is a tenth of it or less up to 1000 A
    CountScreen screen{0, 0.f, 0.f, {(float)bx.L[0], (float)bx.L[1], (float)bx.L[2]}};
    if (mx >= 3 && my >= 3 && mz >= 3 && bx.L[0] <= 1000.0 && bx.L[1] <= 1000.0 && bx.L[2] <= 1000.0) {
      screen.on = 1;
      screen.in2 = (float)((rcut_max * (1.0 - 1e-3)) * (rcut_max * (1.0 - 1e-3)));
      screen.out2 = (float)((rcut_max * (1.0 + 1e-3)) * (rcut_max * (1.0 + 1e-3)));
    }
    TGX_LAUNCH(ctx, "nbr_count", k_cell_neigh<0>, wgrid, wblk, 0, n, ctx->pos.as<double>(), bx, mx, my, mz, rcut_max,
               d_do, cell_xyz, ctx->cell_start.as<int64_t>(), cnt, ctx->cell_atoms.as<int>(),
               ctx->cell_pos.as<double>(), ctx->species_sc.as<int>(), ctx->n_neigh.as<int>(), nullptr, nullptr,
               nullptr, nullptr, nullptr, ctx->nbr_hits.as<int>(), screen);
  } else {
    TGX_LAUNCH(ctx, "nbr_count_brute", k_brute_neigh<0>, wgrid, wblk, 0, n, n_sc, ctx->pos.as<double>(), bx, rcut_max,
               d_do, ctx->species_sc.as<int>(), ctx->n_neigh.as<int>(), nullptr, nullptr, nullptr, nullptr, nullptr);
  }
  int64_t P = 0;
  TGX_TRY(tgx_scan_i32(ctx, ctx->n_neigh.as<int>(), ctx->nl_off.as<int64_t>(), n, &P));
  ctx->n_pairs = P;
  const size_t Pz = (size_t)(P > 0 ? P : 1);
  TGX_TRY(ctx->nlist.ensure(sizeof(int) * Pz));
  TGX_TRY(ctx->rjs.ensure(sizeof(double) * Pz));
  TGX_TRY(ctx->xyz.ensure(sizeof(double) * 3 * Pz));
  TGX_TRY(ctx->thetas.ensure(sizeof(double) * Pz));
  TGX_TRY(ctx->phis.ensure(sizeof(double) * Pz));
  TGX_TRY(ctx->nspec.ensure(sizeof(int) * Pz));
  TGX_TRY(ctx->rev.ensure(sizeof(int64_t) * Pz));
  if (cells) {
    int* cell_xyz = ctx->cell_of.as<int>() + n;
    TGX_LAUNCH(ctx, "nbr_fill", k_cell_neigh<1>, wgrid, wblk, 0, n, ctx->pos.as<double>(), bx, mx, my, mz, rcut_max,
               d_do, cell_xyz, ctx->cell_start.as<int64_t>(), ctx->cell_cnt.as<int>(), ctx->cell_atoms.as<int>(),
               ctx->cell_pos.as<double>(), ctx->species_sc.as<int>(), ctx->n_neigh.as<int>(),
               ctx->nl_off.as<int64_t>(), ctx->nlist.as<int>(), ctx->rjs.as<double>(), ctx->xyz.as<double>(),
               ctx->nspec.as<int>(), ctx->nbr_hits.as<int>(), CountScreen{0, 0.f, 0.f, {0.f, 0.f, 0.f}});
  } else {
    TGX_LAUNCH(ctx, "nbr_fill_brute", k_brute_neigh<1>, wgrid, wblk, 0, n, n_sc, ctx->pos.as<double>(), bx, rcut_max,
               d_do, ctx->species_sc.as<int>(), ctx->n_neigh.as<int>(), ctx->nl_off.as<int64_t>(),
               ctx->nlist.as<int>(), ctx->rjs.as<double>(), ctx->xyz.as<double>(), ctx->nspec.as<int>());
  }
  const int rev_waves = (n + 64 / REV_LANES - 1) / (64 / REV_LANES);
  TGX_LAUNCH(ctx, "nbr_reverse", k_reverse, dim3((rev_waves + NBW - 1) / NBW), wblk, 0, n, ctx->n_neigh.as<int>(),
             ctx->nl_off.as<int64_t>(),
             ctx->nlist.as<int>(), ctx->xyz.as<double>(), cells ? ctx->cell_atoms.as<int>() : (const int*)nullptr,
             ctx->rev.as<int64_t>());
  ctx->rcut_max = rcut_max;
  ctx->have_list = true;
  ctx->have_angles = false;
  return TGX_OK;
}

// rebuild_neighbors_list = .false.: new distances for the resident list. Lists
// of a supercell are rebuilt instead, as the reference's MD does
// (src/turbogap_md.f90:461-463), because the replicas are not moved.
int tgx_neighbors_refresh(tgx_ctx* ctx) {
  TGX_REQUIRE(ctx->have_list, "refresh_neighbors_list: no list built");
  TGX_REQUIRE(ctx->n_sites == ctx->n_sites_sc, "refresh_neighbors_list: a supercell list must be rebuilt");
  const int n = ctx->n_sites;
  BoxInfo bx;
  host_box_info(ctx->box, bx);
  TGX_LAUNCH(ctx, "nbr_refresh", k_refresh, dim3((n + NBW - 1) / NBW), dim3(64 * NBW), 0, n, ctx->n_neigh.as<int>(),
             ctx->nl_off.as<int64_t>(), ctx->nlist.as<int>(), ctx->pos.as<double>(), bx, ctx->rjs.as<double>(),
             ctx->xyz.as<double>());
  ctx->have_angles = false;
  return TGX_OK;
}

// thetas / phis of the resident list, computed once per build when asked for.
int tgx_neighbor_angles(tgx_ctx* ctx) {
  if (ctx->have_angles || ctx->n_pairs == 0) return TGX_OK;
  const int64_t P = ctx->n_pairs;
  TGX_LAUNCH(ctx, "nbr_angles", k_angles, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, P, ctx->rjs.as<double>(),
             ctx->xyz.as<double>(), ctx->thetas.as<double>(), ctx->phis.as<double>());
  ctx->have_angles = true;
  return TGX_OK;
}

// ---- supercell (host only) ------------------------------------------------------
// number_of_unit_cells_for_given_cutoff (src/neighbors.f90:175-236). The
// repetitions follow from the heights of the cell over its three face pairs:
// with unit normals n_ab, n_ca, n_bc, the matrix of their projections on
// (a, b, c) is inverted in closed form and applied to (2 rc, 2 rc, 2 rc); each
// count is that value rounded up. The operations are the reference's, in its
// order (the unit compiles without fused multiply-adds), so the counts agree
// with it also where 2 rc lands on a multiple of a height.
static void cross3(const double* u, const double* v, double* w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

static double dot3h(const double* u, const double* v) {
  double s = u[0] * v[0];
  s = s + u[1] * v[1];
  return s + u[2] * v[2];
}

static void unit_normal(const double* u, const double* v, double* n) {
  cross3(u, v, n);
  const double r = std::sqrt(dot3h(n, n));
  for (int k = 0; k < 3; ++k) n[k] = n[k] / r;
}

extern "C" int tgx_number_of_unit_cells(const double* box, double rcut, int* indices) {
  TGX_REQUIRE(box && indices && rcut > 0.0, "tgx_number_of_unit_cells: null box/indices or rcut <= 0");
  const double *a = box, *b = box + 3, *c = box + 6;
  double nab[3], nca[3], nbc[3];
  unit_normal(a, b, nab);
  unit_normal(c, a, nca);
  unit_normal(b, c, nbc);
  const double* nr[3] = {nab, nca, nbc};
  const double* ax[3] = {a, b, c};
  double M[3][3];
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) M[r][s] = dot3h(nr[r], ax[s]);
  TGX_REQUIRE(std::isfinite(M[0][2]) && std::isfinite(M[1][1]) && std::isfinite(M[2][0]),
              "tgx_number_of_unit_cells: degenerate box");
  const double det = -M[0][2] * M[1][1] * M[2][0] + M[0][1] * M[1][2] * M[2][0] + M[0][2] * M[1][0] * M[2][1] -
                     M[0][0] * M[1][2] * M[2][1] - M[0][1] * M[1][0] * M[2][2] + M[0][0] * M[1][1] * M[2][2];
  TGX_REQUIRE(det != 0.0 && std::isfinite(det), "tgx_number_of_unit_cells: degenerate box");
  // adjugate (transposed cofactors), then divided by the determinant
  double Mi[3][3];
  Mi[0][0] = M[1][1] * M[2][2] - M[1][2] * M[2][1];
  Mi[0][1] = M[0][2] * M[2][1] - M[0][1] * M[2][2];
  Mi[0][2] = M[0][1] * M[1][2] - M[0][2] * M[1][1];
  Mi[1][0] = M[1][2] * M[2][0] - M[1][0] * M[2][2];
  Mi[1][1] = M[0][0] * M[2][2] - M[0][2] * M[2][0];
  Mi[1][2] = M[0][2] * M[1][0] - M[0][0] * M[1][2];
  Mi[2][0] = M[1][0] * M[2][1] - M[1][1] * M[2][0];
  Mi[2][1] = M[0][1] * M[2][0] - M[0][0] * M[2][1];
  Mi[2][2] = M[0][0] * M[1][1] - M[0][1] * M[1][0];
  const double d = 2.0 * rcut;
  for (int r = 0; r < 3; ++r) {
    double v = (Mi[r][0] / det) * d;
    v = v + (Mi[r][1] / det) * d;
    v = std::fabs(v + (Mi[r][2] / det) * d);
    TGX_REQUIRE(v < 1.0e6, "tgx_number_of_unit_cells: %g repetitions along axis %d", v, r + 1);
    int k = (int)v;
    if (v > (double)k) ++k;
    indices[r] = k;
  }
  return TGX_OK;
}

// The supercell of src/read_files.f90:320-367: the cell repeated indices[0] x
// indices[1] x indices[2] times, copy (i, j, k) (i slowest) holding every atom
// shifted by i a + j b + k c, in atom order; the box vectors scaled by the
// counts. positions (3, n_sites) and positions_sc (3, n_sites * prod(indices))
// column-major; box and box_sc hold a, b, c.
extern "C" int tgx_build_supercell(int n_sites, const double* positions, const int* species, const double* box,
                                   const int* indices, double* positions_sc, int* species_sc, double* box_sc) {
  TGX_REQUIRE(n_sites > 0 && positions && species && box && indices && positions_sc && species_sc && box_sc,
              "tgx_build_supercell: bad arguments");
  TGX_REQUIRE(indices[0] >= 1 && indices[1] >= 1 && indices[2] >= 1, "tgx_build_supercell: indices must be >= 1");
  const double *a = box, *b = box + 3, *c = box + 6;
  size_t at = 0;
  for (int i = 0; i < indices[0]; ++i)
    for (int j = 0; j < indices[1]; ++j)
      for (int k = 0; k < indices[2]; ++k)
        for (int s = 0; s < n_sites; ++s, ++at) {
          for (int x = 0; x < 3; ++x) {
            double p = positions[3 * (size_t)s + x] + (double)i * a[x];
            p = p + (double)j * b[x];
            positions_sc[3 * at + x] = p + (double)k * c[x];
          }
          species_sc[at] = species[s];
        }
  for (int x = 0; x < 3; ++x) {
    box_sc[x] = (double)indices[0] * a[x];
    box_sc[3 + x] = (double)indices[1] * b[x];
    box_sc[6 + x] = (double)indices[2] * c[x];
  }
  return TGX_OK;
}
