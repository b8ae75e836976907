// GAP kernel contraction on the matrix cores (src/gap.f90:115-186):
//   K    = q . Qs^T                       (n_sites x n_sparse, k = D)
//   E_i  = delta^2 sum_s alpha_s K_is^zeta + e0
//   Kd   = -zeta delta^2 K^(zeta-1) o alpha
//   Qss  = Kd . Qs                        (n_sites x D, k = n_sparse)  = -dE/dq
// Both products are "NT" GEMMs (both operands contiguous along k): Qs is kept
// as [S][D] for the first and transposed [D][S] for the second, uploaded once.
//
// FP64: v_mfma_f64_16x16x4f64. BF16: v_mfma_f32_16x16x32_bf16 with FP32
// accumulation; the epilogue keeps the energy row sums in FP64.
// Workgroup tile 128x128, four wavefronts of 64x64 (4x4 MFMA tiles each).
// Tiles are handed out XCD-aware: the eight XCDs each walk a contiguous run of
// tiles in groups of eight tile rows, column by column inside a group, so that
// the tiles one XCD has in flight share their A and B panels in its L2.
#include "tgx_internal.h"

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128;
constexpr int TILE_GM = 8;  // tile rows per group of the GEMM tile order (tile_of)

enum { EPI_STORE = TGX_EPI_STORE_F32, EPI_GAP = 1, EPI_SILU = TGX_EPI_SILU, EPI_SCALE_AUX = TGX_EPI_SCALE_AUX,
       EPI_ACT2 = TGX_EPI_ACT2 };

struct EpiArgs {
  double pref;           // -zeta delta^2 (EPI_GAP) / output scale (EPI_STORE, EPI_SCALE_AUX)
  double zeta;
  int zeta_int;          // >= 0 when zeta is an integer
  const double* alpha;   // [N] (FP64 path)
  const float* alpha_f;  // [N] (BF16 path)
  void* out;             // Kd, Qss or the MLP output
  int ldo;
  double* epart;         // [M][n_tiles] (EPI_GAP)
  int n_tiles;
  const float* bias;     // [N] (EPI_SILU)
  float* aux;            // silu'(v) written by EPI_SILU, read by EPI_SCALE_AUX
  int ldaux;
  int64_t sA, sB, sOut, sAux, sBias;  // per-blockIdx.z element strides
  const float* zscale;   // per-z scale (EPI_SCALE_AUX), or null for pref
  const float* w3;       // [N] output weights (EPI_ACT2)
  int64_t sW3, sEpart;   // per-blockIdx.z strides of w3 and epart
};

// GM > 1: inside its range an XCD walks groups of GM tile rows column by
// column, so that the tiles it has in flight at once share their A and B
// panels GM ways each (instead of one row of tiles sharing A only).
template <int GM = 1>
__device__ __forceinline__ void tile_of(int& tm, int& tn) {
  const int ntn = gridDim.x, ntm = gridDim.y;
  const int total = ntn * ntm;
  const int lin = blockIdx.x + ntn * blockIdx.y;
  const int xcd = lin & 7, local = lin >> 3;
  // tiles with lin % 8 == x go to XCD x; give each XCD a contiguous range
  const int base = total / 8, rem = total % 8;
  const int start = xcd * base + (xcd < rem ? xcd : rem);
  const int tile = start + local;
  if (GM == 1) {
    tm = tile / ntn;
    tn = tile - tm * ntn;
  } else {
    const int group = tile / (GM * ntn), first = group * GM;
    const int gm = min(ntm - first, GM), in = tile - first * ntn;
    tm = first + in % gm;
    tn = in / gm;
  }
}

template <class T, bool ZINT>
__device__ __forceinline__ void kpow(T k, double zeta, int zi, T& kz, T& kzm1) {
  if (ZINT) {
    T p = 1;
    for (int i = 0; i < zi - 1; ++i) p *= k;
    kzm1 = p;
    kz = p * k;
  } else {
    kzm1 = (T)pow((double)k, zeta - 1.0);
    kz = (T)pow((double)k, zeta);
  }
}

// Shared epilogue. acc holds, for tile (ti, tj) and register r, the element
// (row0 + ti*16 + rr, col0 + tj*16 + lane%16) where the row within the tile is
//   rr = 4*r + lane/16   for v_mfma_f64_16x16x4f64   (register-major)
//   rr = 4*(lane/16) + r for the FP32-output MFMAs   (lane-group-major)
// (checked against a host GEMM on the device).
template <bool F64> __device__ __forceinline__ int tile_row(int lane, int r) {
  return F64 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
}

// TI x TJ MFMA tiles of 16 x 16 per wavefront (a WM x WN wave tile), the
// workgroup tile BMT rows tall with two wavefronts along N.
template <int EPI, class OutT, class AccT, class VecT, bool F64, bool ZINT, int TI = 4, int TJ = 4, int WM = 64,
          int WN = 64, int BMT = BM>
__device__ __forceinline__ void epilogue(const VecT (&acc)[TI][TJ], int m0, int n0, int tn, int wm, int wn, const EpiArgs& e,
                         double* sE) {
  const int lane = threadIdx.x & 63;
  const int rbase = m0 + wm * WM;
  const int cbase = n0 + wn * WN + (lane & 15);
  OutT* out = reinterpret_cast<OutT*>(e.out);
  if (EPI == EPI_SILU || EPI == EPI_SCALE_AUX) {
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) {
      const int col = cbase + tj * 16;
      const float bc = EPI == EPI_SILU ? e.bias[col] : 0.0f;
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const size_t row = (size_t)(rbase + ti * 16 + tile_row<F64>(lane, r));
          if (EPI == EPI_SILU) {
            const float v = (float)acc[ti][tj][r] + bc;
            out[row * e.ldo + col] = (OutT)tgx_silu(v);
            e.aux[row * e.ldaux + col] = tgx_silu1(v);
          } else {
            out[row * e.ldo + col] = (OutT)((float)acc[ti][tj][r] * e.aux[row * e.ldaux + col] * (float)e.pref);
          }
        }
    }
    return;
  }
  if (EPI == EPI_ACT2) {
    // per 16-row slab: a2 = w3 silu'(v) out, and the row sums of w3 silu(v)
    // over the tile's columns as partial energies (as the GAP epilogue's)
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
      double er[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj) {
        const int col = cbase + tj * 16;
        const float bc = e.bias[col], wc = e.w3[col];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const size_t row = (size_t)(rbase + ti * 16 + tile_row<F64>(lane, r));
          const float v = (float)acc[ti][tj][r] + bc;
          out[row * e.ldo + col] = (OutT)(wc * tgx_silu1(v));
          er[r] += (double)(wc * tgx_silu(v));
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double v = er[r];
        v += __shfl_xor(v, 1);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 8);
        if ((lane & 15) == 0) sE[wn * BMT + wm * WM + ti * 16 + tile_row<F64>(lane, r)] = v;
      }
    }
    __syncthreads();
    if (threadIdx.x < BMT) {
      const int row = m0 + threadIdx.x;
      e.epart[(size_t)row * e.n_tiles + tn] = sE[threadIdx.x] + sE[BMT + threadIdx.x];
    }
    return;
  }
  if (EPI == EPI_STORE) {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          out[(size_t)(rbase + ti * 16 + tile_row<F64>(lane, r)) * e.ldo + cbase + tj * 16] =
              (OutT)((AccT)e.pref * acc[ti][tj][r]);
    return;
  }
  // one 16-row slab at a time keeps only four row partials live
#pragma unroll
  for (int ti = 0; ti < TI; ++ti) {
    double er[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) {
      const int col = cbase + tj * 16;
      const AccT al = e.alpha_f ? (AccT)e.alpha_f[col] : (AccT)e.alpha[col];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        AccT kz, kzm1;
        kpow<AccT, ZINT>(acc[ti][tj][r], e.zeta, e.zeta_int, kz, kzm1);
        out[(size_t)(rbase + ti * 16 + tile_row<F64>(lane, r)) * e.ldo + col] = (OutT)((AccT)e.pref * kzm1 * al);
        er[r] += (double)(al * kz);
      }
    }
    // reduce over the 16 lanes that share rows
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double v = er[r];
      v += __shfl_xor(v, 1);
      v += __shfl_xor(v, 2);
      v += __shfl_xor(v, 4);
      v += __shfl_xor(v, 8);
      if ((lane & 15) == 0) sE[wn * BMT + wm * WM + ti * 16 + tile_row<F64>(lane, r)] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < BMT) {
    const int row = m0 + threadIdx.x;
    e.epart[(size_t)row * e.n_tiles + tn] = sE[threadIdx.x] + sE[BMT + threadIdx.x];
  }
}

// FP64 NT GEMM. M, N multiples of 128; K multiple of 16.
// LDS holds both operands k-major ([k][row], row stride 144 doubles: the four
// k rows an MFMA fragment read touches fall on disjoint bank halves), in two
// stages: the global loads for stage s+1 are in flight during the MFMAs of
// stage s and land in the other buffer, one barrier per stage. Threads 0-127
// stage one row of A each, 128-255 one row of B (16 doubles, 128 contiguous
// bytes), so a wavefront's LDS writes cover 64 consecutive rows.
template <int EPI, bool ZINT>
__global__ __launch_bounds__(256, 2) void k_gemm_f64(int K, const double* __restrict__ A, int lda,
                                                     const double* __restrict__ B, int ldb, EpiArgs e) {
  constexpr int BK = 16, LS = BM + 16;
  __shared__ __align__(16) double As[2][BK * LS];
  __shared__ __align__(16) double Bs[2][BK * LS];
  __shared__ double sE[2 * BM];
  int tm, tn;
  tile_of<TILE_GM>(tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int srow = tid & 127;
  const bool isB = tid >= 128;
  const double* G = isB ? B + (size_t)(n0 + srow) * ldb : A + (size_t)(m0 + srow) * lda;
  f64x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f64x4){0, 0, 0, 0};
  double2 r0, r1, r2, r3, r4, r5, r6, r7;
  auto gload = [&](int k) {
    const double2* g = reinterpret_cast<const double2*>(G + k);
    r0 = g[0]; r1 = g[1]; r2 = g[2]; r3 = g[3]; r4 = g[4]; r5 = g[5]; r6 = g[6]; r7 = g[7];
  };
  auto lstore = [&](int buf) {
    double* d = (isB ? Bs[buf] : As[buf]) + srow;
    d[0 * LS] = r0.x; d[1 * LS] = r0.y; d[2 * LS] = r1.x; d[3 * LS] = r1.y;
    d[4 * LS] = r2.x; d[5 * LS] = r2.y; d[6 * LS] = r3.x; d[7 * LS] = r3.y;
    d[8 * LS] = r4.x; d[9 * LS] = r4.y; d[10 * LS] = r5.x; d[11 * LS] = r5.y;
    d[12 * LS] = r6.x; d[13 * LS] = r6.y; d[14 * LS] = r7.x; d[15 * LS] = r7.y;
  };
  gload(0);
  lstore(0);
  __syncthreads();
  const int n_stages = K / BK;
  for (int st = 0; st < n_stages; ++st) {
    const int buf = st & 1;
    gload(min(st + 1, n_stages - 1) * BK);  // the last stage reloads itself: always in bounds
    __builtin_amdgcn_sched_barrier(0);
    const double* as = As[buf] + wm * 64 + (lane & 15);
    const double* bs = Bs[buf] + wn * 64 + (lane & 15);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kr = (ks * 4 + (lane >> 4)) * LS;
      double a[4], b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = as[kr + t * 16];
        b[t] = bs[kr + t * 16];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    lstore(buf ^ 1);
    __syncthreads();
  }
  epilogue<EPI, double, double, f64x4, true, ZINT>(acc, m0, n0, tn, wm, wn, e, sE);
}

// BF16 NT GEMM with FP32 accumulation. M, N multiples of 128; K multiple of 64.
// Two LDS stages: the global loads of stage k+1 are in flight while the MFMAs
// of stage k run, and one barrier per stage separates the LDS write of k+1
// from the reads of k. Each thread moves four 16-byte chunks per operand and
// stage: chunk t%8 of rows t/8 + 32 j, so a wavefront reads eight full 128-byte
// row segments per load. The next-stage offset is clamped to the last stage
// (a redundant, in-bounds reload) so the staging registers are written on
// every iteration and never demoted to private memory.
template <int EPI, class OutT, bool ZINT>
__global__ __launch_bounds__(256) void k_gemm_bf16(int K, const __bf16* __restrict__ A, int lda,
                                                   const __bf16* __restrict__ B, int ldb, EpiArgs e) {
  constexpr int BK = 64, LDS_STRIDE = BK + 8;  // 144-byte row pitch: conflict-free 16-byte reads
  __shared__ __align__(16) __bf16 As[2][BM * LDS_STRIDE];
  __shared__ __align__(16) __bf16 Bs[2][BN * LDS_STRIDE];
  __shared__ double sE[2 * BM];
  int tm, tn;
  tile_of<TILE_GM>(tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int lr = tid >> 3, lc = (tid & 7) * 8;
  {
    const int z = blockIdx.z;  // batch member: offset every operand
    A += z * e.sA;
    B += z * e.sB;
    e.out = reinterpret_cast<OutT*>(e.out) + z * e.sOut;
    if (e.aux) e.aux += z * e.sAux;
    if (e.bias) e.bias += z * e.sBias;
    if (e.w3) e.w3 += z * e.sW3;
    if (e.epart) e.epart += z * e.sEpart;
    if (e.zscale) e.pref = e.zscale[z];
  }
  const __bf16* Ag = A + (size_t)(m0 + lr) * lda + lc;
  const __bf16* Bg = B + (size_t)(n0 + lr) * ldb + lc;
  const size_t a32 = (size_t)32 * lda, b32 = (size_t)32 * ldb;
  // The scaled epilogue's factors are read before the main loop, so that their
  // latency overlaps it (a short K leaves the epilogue most of the time; the
  // 64 registers fit in the two waves per SIMD that the LDS allows)
  float auxv[4][4][4];
  if constexpr (EPI == EPI_SCALE_AUX) {
    const int rb = m0 + wm * 64, cb = n0 + wn * 64 + (lane & 15);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          auxv[ti][tj][r] = e.aux[(size_t)(rb + ti * 16 + tile_row<false>(lane, r)) * e.ldaux + cb + tj * 16];
  }
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0, 0, 0, 0};
  uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
  auto gload = [&](int k) {
    ra0 = *reinterpret_cast<const uint4*>(Ag + k);
    ra1 = *reinterpret_cast<const uint4*>(Ag + a32 + k);
    ra2 = *reinterpret_cast<const uint4*>(Ag + 2 * a32 + k);
    ra3 = *reinterpret_cast<const uint4*>(Ag + 3 * a32 + k);
    rb0 = *reinterpret_cast<const uint4*>(Bg + k);
    rb1 = *reinterpret_cast<const uint4*>(Bg + b32 + k);
    rb2 = *reinterpret_cast<const uint4*>(Bg + 2 * b32 + k);
    rb3 = *reinterpret_cast<const uint4*>(Bg + 3 * b32 + k);
  };
  auto lstore = [&](int buf) {
    __bf16* as = &As[buf][lr * LDS_STRIDE + lc];
    __bf16* bs = &Bs[buf][lr * LDS_STRIDE + lc];
    *reinterpret_cast<uint4*>(as) = ra0;
    *reinterpret_cast<uint4*>(as + 32 * LDS_STRIDE) = ra1;
    *reinterpret_cast<uint4*>(as + 64 * LDS_STRIDE) = ra2;
    *reinterpret_cast<uint4*>(as + 96 * LDS_STRIDE) = ra3;
    *reinterpret_cast<uint4*>(bs) = rb0;
    *reinterpret_cast<uint4*>(bs + 32 * LDS_STRIDE) = rb1;
    *reinterpret_cast<uint4*>(bs + 64 * LDS_STRIDE) = rb2;
    *reinterpret_cast<uint4*>(bs + 96 * LDS_STRIDE) = rb3;
  };
  gload(0);
  lstore(0);
  __syncthreads();
  const int n_stages = K / BK;
  for (int st = 0; st < n_stages; ++st) {
    const int buf = st & 1;
    const int kn = min(st + 1, n_stages - 1) * BK;
    gload(kn);
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the MFMAs they overlap
    const __bf16* as = &As[buf][(wm * 64 + (lane & 15)) * LDS_STRIDE + 8 * (lane >> 4)];
    const __bf16* bs = &Bs[buf][(wn * 64 + (lane & 15)) * LDS_STRIDE + 8 * (lane >> 4)];
#pragma unroll
    for (int ks = 0; ks < BK / 32; ++ks) {
      bf16x8 a[4], b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        a[t] = *reinterpret_cast<const bf16x8*>(as + t * 16 * LDS_STRIDE + ks * 32);
        b[t] = *reinterpret_cast<const bf16x8*>(bs + t * 16 * LDS_STRIDE + ks * 32);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    lstore(buf ^ 1);
    __syncthreads();
  }
  if constexpr (EPI == EPI_SCALE_AUX) {
    OutT* out = reinterpret_cast<OutT*>(e.out);
    const int rb = m0 + wm * 64, cb = n0 + wn * 64 + (lane & 15);
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          out[(size_t)(rb + ti * 16 + tile_row<false>(lane, r)) * e.ldo + cb + tj * 16] =
              (OutT)((float)acc[ti][tj][r] * auxv[ti][tj][r] * (float)e.pref);
  } else {
    epilogue<EPI, OutT, float, f32x4, false, ZINT>(acc, m0, n0, tn, wm, wn, e, sE);
  }
}

// BF16 NT GEMM, workgroup tile 256 x 128: four wavefronts of 128 x 64 (8 x 4
// MFMA tiles), BK = 32 in two LDS stages of (256 + 128) rows x 80 bytes
// (61 KB: two workgroups per CU). A wavefront reads 12 fragments per 32 MFMAs
// against 8 per 16 in the 128 x 128 kernel, a quarter less LDS traffic per
// flop. M a multiple of 256, N of 128, K of 32.
template <int EPI, class OutT, bool ZINT>
__global__ __launch_bounds__(256, 2) void k_gemm_bf16_w(int K, const __bf16* __restrict__ A, int lda,
                                                        const __bf16* __restrict__ B, int ldb, EpiArgs e) {
  constexpr int TM = 256, BK = 32, LS = BK + 8;  // 80-byte rows: 16-byte fragment reads hit distinct banks
  __shared__ __align__(16) __bf16 As[2][TM * LS];
  __shared__ __align__(16) __bf16 Bs[2][BN * LS];
  __shared__ double sE[2 * TM];
  int tm, tn;
  tile_of<TILE_GM>(tm, tn);
  const int m0 = tm * TM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  // rows lr + 64 j, 16-byte chunk lc; the eight lanes of a store group take rows r and r + 4 of an
  // 8-row run (at the 80-byte pitch rows r and r + 1 would share a bank at byte 128)
  const int jr = tid >> 2;
  const int lr = (jr & ~7) | ((jr & 1) << 2) | ((jr >> 1) & 3), lc = (tid & 3) * 8;
  {
    const int z = blockIdx.z;
    A += z * e.sA;
    B += z * e.sB;
    e.out = reinterpret_cast<OutT*>(e.out) + z * e.sOut;
    if (e.aux) e.aux += z * e.sAux;
    if (e.bias) e.bias += z * e.sBias;
    if (e.w3) e.w3 += z * e.sW3;
    if (e.epart) e.epart += z * e.sEpart;
    if (e.zscale) e.pref = e.zscale[z];
  }
  const __bf16* Ag = A + (size_t)(m0 + lr) * lda + lc;
  const __bf16* Bg = B + (size_t)(n0 + lr) * ldb + lc;
  const size_t a64 = (size_t)64 * lda, b64 = (size_t)64 * ldb;
  f32x4 acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0, 0, 0, 0};
  uint4 ra0, ra1, ra2, ra3, rb0, rb1;  // named: arrays of staging registers are demoted to private memory
  auto gload = [&](int k) {
    ra0 = *reinterpret_cast<const uint4*>(Ag + k);
    ra1 = *reinterpret_cast<const uint4*>(Ag + a64 + k);
    ra2 = *reinterpret_cast<const uint4*>(Ag + 2 * a64 + k);
    ra3 = *reinterpret_cast<const uint4*>(Ag + 3 * a64 + k);
    rb0 = *reinterpret_cast<const uint4*>(Bg + k);
    rb1 = *reinterpret_cast<const uint4*>(Bg + b64 + k);
  };
  auto lstore = [&](int buf) {
    __bf16* as = &As[buf][lr * LS + lc];
    __bf16* bs = &Bs[buf][lr * LS + lc];
    *reinterpret_cast<uint4*>(as) = ra0;
    *reinterpret_cast<uint4*>(as + 64 * LS) = ra1;
    *reinterpret_cast<uint4*>(as + 128 * LS) = ra2;
    *reinterpret_cast<uint4*>(as + 192 * LS) = ra3;
    *reinterpret_cast<uint4*>(bs) = rb0;
    *reinterpret_cast<uint4*>(bs + 64 * LS) = rb1;
  };
  gload(0);
  lstore(0);
  __syncthreads();
  const int n_stages = K / BK;
  for (int st = 0; st < n_stages; ++st) {
    const int buf = st & 1;
    gload(min(st + 1, n_stages - 1) * BK);
    __builtin_amdgcn_sched_barrier(0);
    const __bf16* as = &As[buf][(wm * 128 + (lane & 15)) * LS + 8 * (lane >> 4)];
    const __bf16* bs = &Bs[buf][(wn * 64 + (lane & 15)) * LS + 8 * (lane >> 4)];
    bf16x8 b[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) b[t] = *reinterpret_cast<const bf16x8*>(bs + t * 16 * LS);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(as + i * 16 * LS);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);
    lstore(buf ^ 1);
    __syncthreads();
  }
  epilogue<EPI, OutT, float, f32x4, false, ZINT, 8, 4, 128, 64, TM>(acc, m0, n0, tn, wm, wn, e, sE);
}

// BF16 NT GEMM fed by direct global-to-LDS loads (global_load_lds_dwordx4):
// workgroup tile TM x 128, four wavefronts of (TM / 2) x 64, BK = 32, an
// NS-slot ring of LDS stages. A stage is TM + 128 rows of 64 bytes (the A
// rows, then the B rows); the four 16-byte chunks of row r sit at chunk
// c ^ ((r >> 1) & 3), so that the eight rows a 16-byte fragment read serves
// per cycle cover all 32 banks. One load instruction fills 16 rows (lane l
// writes chunk l & 3 of row l >> 2 and fetches the chunk that belongs there);
// nothing is staged through registers. The fragments of each stage are held
// in registers, double-buffered: stage s + 1's fragments are read while stage
// s's MFMAs run, so a ring slot is free as soon as every wavefront has read it
// and stage s + NS is loaded into stage s's slot right after the barrier of
// iteration s (NS stages in flight). Iteration s: the MFMAs of the first half
// of the tile rows, then (vmcnt: this wavefront's part of stage s + 1 landed)
// the barrier, the loads of stage s + NS and the fragment reads of stage
// s + 1, then the second half of the MFMAs, which cover the reads' latency.
// M a multiple of TM, N of 128, K of 32.
constexpr int DK = 32;                      // k per stage
constexpr int DROW = DK * 2;                // bytes per staged row
__host__ __device__ constexpr int dma_stage_bytes(int tm) { return (tm + BN) * DROW; }
__host__ __device__ constexpr int dma_lds_bytes(int tm, int ns) { return ns * dma_stage_bytes(tm) + 2 * tm * 8; }

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 16, "vmcnt immediate");
  __builtin_amdgcn_s_waitcnt(0x3F70 | N);  // vmcnt(N); expcnt and lgkmcnt not waited on
}

// One stage's loads of a wavefront: its P blocks of 16 rows, block b of the
// stage into bytes [1024 b, 1024 b + 1024) of the stage buffer.
template <int P>
__device__ __forceinline__ void dma_issue(const __bf16* const (&src)[P], int k0, unsigned char* stage, int w) {
#pragma unroll
  for (int j = 0; j < P; ++j)
    __builtin_amdgcn_global_load_lds(const_cast<__bf16*>(src[j] + k0), stage + (w + 4 * j) * 1024, 16, 0, 0);
}

template <int TI>
__device__ __forceinline__ void pf_read(const unsigned char* sb, int aoff, int boff, bf16x8 (&a)[TI], bf16x8 (&b)[4]) {
#pragma unroll
  for (int t = 0; t < 4; ++t) b[t] = *reinterpret_cast<const bf16x8*>(sb + boff + t * 16 * DROW);
#pragma unroll
  for (int i = 0; i < TI; ++i) a[i] = *reinterpret_cast<const bf16x8*>(sb + aoff + i * 16 * DROW);
}

template <int TI, int I0, int I1>
__device__ __forceinline__ void pf_mfma(f32x4 (&acc)[TI][4], const bf16x8 (&a)[TI], const bf16x8 (&b)[4]) {
#pragma unroll
  for (int i = I0; i < I1; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
}

template <int TI, int P, int NS, int SB>
__device__ __forceinline__ void pf_stage(f32x4 (&acc)[TI][4], const bf16x8 (&a)[TI], const bf16x8 (&b)[4],
                                         bf16x8 (&an)[TI], bf16x8 (&bn)[4], const __bf16* const (&src)[P],
                                         unsigned char* dsm, int st, int n_stages, int& slot, int aoff, int boff,
                                         int w) {
  pf_mfma<TI, 0, TI / 2>(acc, a, b);
  __builtin_amdgcn_sched_barrier(0);
  wait_vmcnt<(NS - 2) * P>();
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_sched_barrier(0);
  // every wavefront has read this stage's slot: it takes stage st + NS (an in-bounds repeat past the end)
  dma_issue<P>(src, min(st + NS, n_stages - 1) * DK, dsm + slot * SB, w);
  slot = slot + 1 == NS ? 0 : slot + 1;
  pf_read<TI>(dsm + slot * SB, aoff, boff, an, bn);
  __builtin_amdgcn_sched_barrier(0);
  pf_mfma<TI, TI / 2, TI>(acc, a, b);
}

template <int EPI, class OutT, bool ZINT, int TM, int NS, int OCC>
__global__ __launch_bounds__(256, OCC) void k_gemm_bf16_pf(int K, const __bf16* __restrict__ A, int lda,
                                                          const __bf16* __restrict__ B, int ldb, EpiArgs e) {
  constexpr int TI = TM / 32;
  constexpr int NBA = TM / 16, NBB = BN / 16;
  constexpr int P = (NBA + NBB) / 4;
  static_assert((NBA + NBB) % 4 == 0, "load blocks per wavefront");
  static_assert(NS >= 3 && (NS - 1) * P < 16, "vmcnt range");
  constexpr int SB = dma_stage_bytes(TM);
  extern __shared__ __align__(16) unsigned char dsm[];
  double* sE = reinterpret_cast<double*>(dsm + NS * SB);
  int tm, tn;
  tile_of<TILE_GM>(tm, tn);
  const int m0 = tm * TM, n0 = tn * BN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  {
    const int z = blockIdx.z;
    A += z * e.sA;
    B += z * e.sB;
    e.out = reinterpret_cast<OutT*>(e.out) + z * e.sOut;
    if (e.aux) e.aux += z * e.sAux;
    if (e.bias) e.bias += z * e.sBias;
    if (e.w3) e.w3 += z * e.sW3;
    if (e.epart) e.epart += z * e.sEpart;
    if (e.zscale) e.pref = e.zscale[z];
  }
  const int lrow = lane >> 2, lchunk = (lane & 3) ^ ((lane >> 3) & 3);
  const __bf16* src[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int b = w + 4 * j;
    src[j] = b < NBA ? A + (size_t)(m0 + 16 * b + lrow) * lda + 8 * lchunk
                     : B + (size_t)(n0 + 16 * (b - NBA) + lrow) * ldb + 8 * lchunk;
  }
  const int li = lane & 15, lg = lane >> 4;
  const int foff = li * DROW + ((lg ^ ((li >> 1) & 3)) << 4);
  const int aoff = wm * (TM / 2) * DROW + foff, boff = TM * DROW + wn * 64 * DROW + foff;
  f32x4 acc[TI][4];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (f32x4){0, 0, 0, 0};
  const int n_stages = K / DK;
#pragma unroll
  for (int s = 0; s < NS; ++s) dma_issue<P>(src, min(s, n_stages - 1) * DK, dsm + s * SB, w);
  wait_vmcnt<(NS - 1) * P>();
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_sched_barrier(0);
  bf16x8 a0[TI], b0[4], a1[TI], b1[4];
  pf_read<TI>(dsm, aoff, boff, a0, b0);
  int slot = 0, st = 0;
  for (; st + 1 < n_stages; st += 2) {
    pf_stage<TI, P, NS, SB>(acc, a0, b0, a1, b1, src, dsm, st, n_stages, slot, aoff, boff, w);
    pf_stage<TI, P, NS, SB>(acc, a1, b1, a0, b0, src, dsm, st + 1, n_stages, slot, aoff, boff, w);
  }
  if (st < n_stages) pf_stage<TI, P, NS, SB>(acc, a0, b0, a1, b1, src, dsm, st, n_stages, slot, aoff, boff, w);
  // the loads issued past the last stage land before the workgroup's LDS is released
  wait_vmcnt<0>();
  __builtin_amdgcn_sched_barrier(0);
  epilogue<EPI, OutT, float, f32x4, false, ZINT, TI, 4, TM / 2, 64, TM>(acc, m0, n0, tn, wm, wn, e, sE);
}

// E_i = delta^2 sum_t epart[i][t] + e0, the partials summed in tile order.
__global__ void k_energy_finalize(int n, const double* __restrict__ epart, int n_tiles, double d2, double e0,
                                  double* __restrict__ e_site) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int t = 0; t < n_tiles; ++t) s += epart[(size_t)i * n_tiles + t];
  e_site[i] = d2 * s + e0;
}

__global__ void k_f32_to_bf16(int64_t n, const float* __restrict__ x, __bf16* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) y[i] = (__bf16)x[i];
}

}  // namespace

static EpiArgs gap_epi(const tgx_model_t* m) {
  EpiArgs e{};
  const double z = m->zeta;
  const bool zint = std::fabs(z - (double)(int)z) < 1.0e-5;
  e.zeta_int = zint ? (int)z : 0;
  e.zeta = zint ? (double)(int)z : z;
  e.pref = -e.zeta * m->delta * m->delta;
  return e;
}

// BF16 NT GEMM launch. Plain FP32 stores with at least 512 tiles of 256 x 128
// (one round of two workgroups per CU) go to the kernel fed by direct LDS
// loads (0.389 -> 0.345 ms on the 32,768 x 1,280 x 4,096 Qss product, with the
// grouped tile order); the GAP and second-layer epilogues to its 128 x 128
// form with three workgroups per CU (GAP kernel 0.495 -> 0.476 ms, second
// layer 0.052 -> 0.047 ms). The SiLU and scaled epilogues measured faster on
// the register-staged kernels.
// Otherwise 256 x 128 tiles when M is a multiple of 256 and there are at least
// 1,024 of them (two rounds of two workgroups per CU; with fewer, the last
// round's imbalance costs more than the wider tile saves, measured on the
// 32,768-row MLP shapes), else 128 x 128. The GAP epilogue writes one energy
// partial per row and 128-column tile either way.
template <int EPI, class OutT, bool ZINT>
static int launch_bf16(tgx_ctx* ctx, const char* name, int M, int N, int K, int batch, const __bf16* A, int lda,
                       const __bf16* B, int ldb, EpiArgs e) {
  TGX_REQUIRE(M % BM == 0 && N % BN == 0 && K % 64 == 0 && lda % 8 == 0 && ldb % 8 == 0,
              "gemm_bf16 %s: unpadded shape %d x %d x %d", name, M, N, K);
  if constexpr (EPI == EPI_STORE) {
    if (M % 256 == 0 && (long)(M / 256) * (N / BN) * batch >= 512) {
      TGX_LAUNCH(ctx, name, (k_gemm_bf16_pf<EPI, OutT, ZINT, 256, 3, 2>), dim3(N / BN, M / 256, batch), dim3(256),
                 dma_lds_bytes(256, 3), K, A, lda, B, ldb, e);
      return TGX_OK;
    }
  }
  if constexpr (EPI == EPI_GAP || EPI == EPI_ACT2) {
    if (K % DK == 0) {
      TGX_LAUNCH(ctx, name, (k_gemm_bf16_pf<EPI, OutT, ZINT, 128, 3, 3>), dim3(N / BN, M / BM, batch), dim3(256),
                 dma_lds_bytes(128, 3), K, A, lda, B, ldb, e);
      return TGX_OK;
    }
  }
  if (M % 256 == 0 && (long)(M / 256) * (N / BN) * batch >= 1024)
    TGX_LAUNCH(ctx, name, (k_gemm_bf16_w<EPI, OutT, ZINT>), dim3(N / BN, M / 256, batch), dim3(256), 0, K, A, lda, B,
               ldb, e);
  else
    TGX_LAUNCH(ctx, name, (k_gemm_bf16<EPI, OutT, ZINT>), dim3(N / BN, M / BM, batch), dim3(256), 0, K, A, lda, B, ldb,
               e);
  return TGX_OK;
}

static int launch_f64(tgx_ctx* ctx, int epi, int M, int N, int K, const double* A, int lda, const double* B, int ldb,
                      EpiArgs e, const char* name) {
  TGX_REQUIRE(M % BM == 0 && N % BN == 0 && K % 16 == 0, "gemm_f64: unpadded shape %d x %d x %d", M, N, K);
  dim3 grid(N / BN, M / BM);
  if (epi == EPI_GAP && e.zeta_int >= 1)
    TGX_LAUNCH(ctx, name, (k_gemm_f64<EPI_GAP, true>), grid, dim3(256), 0, K, A, lda, B, ldb, e);
  else if (epi == EPI_GAP)
    TGX_LAUNCH(ctx, name, (k_gemm_f64<EPI_GAP, false>), grid, dim3(256), 0, K, A, lda, B, ldb, e);
  else
    TGX_LAUNCH(ctx, name, (k_gemm_f64<EPI_STORE, true>), grid, dim3(256), 0, K, A, lda, B, ldb, e);
  return TGX_OK;
}

int tgx_gap_contract_f64(tgx_ctx* ctx, const tgx_model_t* m, int ns, const double* q, double* kd, double* epart,
                         double* e_site, double* qss) {
  if (ns == 0) return TGX_OK;
  const int Mp = (int)tgx_round_up(ns, BM);
  const int Sp = m->Spad, Dp = m->Dpad;
  const int n_tiles = Sp / BN;
  EpiArgs e = gap_epi(m);
  e.epart = epart;
  e.n_tiles = n_tiles;
  e.alpha = m->alpha.as<double>();
  e.out = kd;
  e.ldo = Sp;
  TGX_TRY(launch_f64(ctx, EPI_GAP, Mp, Sp, Dp, q, Dp, m->Qs.as<double>(), Dp, e, "gemm_kernel_f64"));
  TGX_LAUNCH(ctx, "energy_finalize", k_energy_finalize, dim3((ns + 255) / 256), dim3(256), 0, ns, epart, n_tiles,
             m->delta * m->delta, m->e0, e_site);
  if (!qss) return TGX_OK;
  EpiArgs s{};
  s.pref = 1.0;
  s.out = qss;
  s.ldo = Dp;
  return launch_f64(ctx, EPI_STORE, Mp, Dp, Sp, kd, Sp, m->QsT.as<double>(), Sp, s, "gemm_qss_f64");
}

int tgx_gap_contract(tgx_ctx* ctx, tgx_model_t* m, int path, bool do_forces) {
  const int ns = m->n_sites_b;
  if (ns == 0) return TGX_OK;
  const int Mp = (int)tgx_round_up(ns, BM);
  const int Sp = m->Spad, Dp = m->Dpad;
  const int n_tiles = Sp / BN;
  TGX_TRY(m->epart.ensure(sizeof(double) * (size_t)Mp * n_tiles));
  TGX_TRY(m->e_site.ensure(sizeof(double) * (size_t)Mp));
  if (path == TGX_PATH_FP64_PLAIN || path == TGX_PATH_FP64_ADJOINT) {
    TGX_TRY(m->kd64.ensure(sizeof(double) * (size_t)Mp * Sp));
    if (do_forces) TGX_TRY(m->qss64.ensure(sizeof(double) * (size_t)Mp * Dp));
    return tgx_gap_contract_f64(ctx, m, ns, m->q64.as<double>(), m->kd64.as<double>(), m->epart.as<double>(),
                                m->e_site.as<double>(), do_forces ? m->qss64.as<double>() : nullptr);
  }
  // mixed path: qbf (BF16 descriptors) comes from the FP32 descriptor pass
  EpiArgs e = gap_epi(m);
  e.epart = m->epart.as<double>();
  e.n_tiles = n_tiles;
  TGX_REQUIRE(m->qbf.bytes >= sizeof(__bf16) * (size_t)Mp * Dp, "gemm_bf16: BF16 descriptors missing");
  TGX_TRY(m->kdbf.ensure(sizeof(__bf16) * (size_t)Mp * Sp));
  e.alpha_f = m->alpha_f.as<float>();
  e.out = m->kdbf.p;
  e.ldo = Sp;
  if (e.zeta_int >= 1)
    TGX_TRY((launch_bf16<EPI_GAP, __bf16, true>(ctx, "gemm_kernel_bf16", Mp, Sp, Dp, 1, m->qbf.as<__bf16>(), Dp,
                                                 m->Qs_bf.as<__bf16>(), Dp, e)));
  else
    TGX_TRY((launch_bf16<EPI_GAP, __bf16, false>(ctx, "gemm_kernel_bf16", Mp, Sp, Dp, 1, m->qbf.as<__bf16>(), Dp,
                                                  m->Qs_bf.as<__bf16>(), Dp, e)));
  TGX_LAUNCH(ctx, "energy_finalize", k_energy_finalize, dim3((ns + 255) / 256), dim3(256), 0, ns,
             m->epart.as<double>(), n_tiles, m->delta * m->delta, m->e0, m->e_site.as<double>());
  if (!do_forces) return TGX_OK;
  EpiArgs s{};
  s.pref = 1.0;
  TGX_TRY(m->qss32.ensure(sizeof(float) * (size_t)Mp * Dp));
  s.out = m->qss32.p;
  s.ldo = Dp;
  return launch_bf16<EPI_STORE, float, true>(ctx, "gemm_qss_bf16", Mp, Dp, Sp, 1, m->kdbf.as<__bf16>(), Sp,
                                             m->QsT_bf.as<__bf16>(), Sp, s);
}

// C[M][N] (FP32) = scale * A[M][K] . B[N][K]^T in BF16 with FP32 accumulation.
// M, N multiples of 128, K a multiple of 64, leading dimensions in elements.
int tgx_gemm_bf16_nt(tgx_ctx* ctx, const char* name, int M, int N, int K, const __bf16* A, int lda, const __bf16* B,
                     int ldb, float* C, int ldc, double scale) {
  EpiArgs s{};
  s.pref = scale;
  s.out = C;
  s.ldo = ldc;
  return launch_bf16<EPI_STORE, float, true>(ctx, name, M, N, K, 1, A, lda, B, ldb, s);
}

int tgx_gemm_bf16(tgx_ctx* ctx, const char* name, const tgx_gemm_desc& d) {
  TGX_REQUIRE(d.M % BM == 0 && d.N % BN == 0 && d.K % 64 == 0 && d.lda % 8 == 0 && d.ldb % 8 == 0 && d.sA % 8 == 0 &&
                  d.sB % 8 == 0 && d.batch >= 1 && d.batch <= 65535,
              "gemm_bf16 %s: unpadded shape %d x %d x %d (batch %d)", name, d.M, d.N, d.K, d.batch);
  TGX_REQUIRE(d.epi != TGX_EPI_SILU || (d.bias && d.aux), "gemm_bf16 %s: SiLU epilogue needs bias and aux", name);
  TGX_REQUIRE(d.epi != TGX_EPI_SCALE_AUX || d.aux, "gemm_bf16 %s: scaled epilogue needs aux", name);
  TGX_REQUIRE(d.epi != TGX_EPI_ACT2 || (d.bias && d.w3 && d.epart), "gemm_bf16 %s: ACT2 needs bias, w3, epart",
              name);
  EpiArgs e{};
  e.pref = d.scale;
  e.out = d.C;
  e.ldo = d.ldc;
  e.bias = d.bias;
  e.aux = d.aux;
  e.ldaux = d.ldaux;
  e.sA = d.sA;
  e.sB = d.sB;
  e.sOut = d.sC;
  e.sAux = d.sAux;
  e.sBias = d.sBias;
  e.zscale = d.zscale;
  e.w3 = d.w3;
  e.sW3 = d.sW3;
  e.epart = d.epart;
  e.sEpart = d.sEpart;
  e.n_tiles = d.N / BN;
  switch (d.epi) {
    case TGX_EPI_STORE_F32:
      return launch_bf16<EPI_STORE, float, true>(ctx, name, d.M, d.N, d.K, d.batch, d.A, d.lda, d.B, d.ldb, e);
    case TGX_EPI_SILU:
      return launch_bf16<EPI_SILU, __bf16, true>(ctx, name, d.M, d.N, d.K, d.batch, d.A, d.lda, d.B, d.ldb, e);
    case TGX_EPI_SCALE_AUX:
      return launch_bf16<EPI_SCALE_AUX, __bf16, true>(ctx, name, d.M, d.N, d.K, d.batch, d.A, d.lda, d.B, d.ldb, e);
    case TGX_EPI_ACT2:
      return launch_bf16<EPI_ACT2, __bf16, true>(ctx, name, d.M, d.N, d.K, d.batch, d.A, d.lda, d.B, d.ldb, e);
    default:
      TGX_REQUIRE(false, "gemm_bf16 %s: unknown epilogue %d", name, d.epi);
  }
}

int tgx_f32_to_bf16(tgx_ctx* ctx, const float* x, __bf16* y, int64_t n) {
  TGX_LAUNCH(ctx, "to_bf16", k_f32_to_bf16, dim3(1024), dim3(256), 0, n, x, y);
  return TGX_OK;
}
