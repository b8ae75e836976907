// Device utilities: zero fill, deterministic exclusive scan, deterministic sum.
#include "tgx_util.h"

namespace {

constexpr int SCAN_BLOCK = 1024;
constexpr int64_t SCAN_SINGLE_MAX = 16 * 1024;  // one workgroup scans up to this many values

__global__ void k_fill_zero(uint32_t* __restrict__ p, size_t n_words) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n_words; i += stride) p[i] = 0u;
}

__global__ void k_fill_zero_bytes(uint8_t* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Block-local exclusive scan of 1024 values (one value per thread).
__device__ int64_t block_exclusive_scan(int64_t v, int64_t* sh, int64_t* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < SCAN_BLOCK; o <<= 1) {
    int64_t add = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  int64_t incl = sh[t];
  *total = sh[SCAN_BLOCK - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_local(const int* __restrict__ in, int64_t* __restrict__ out,
                                                          int64_t n, int64_t* __restrict__ block_sums) {
  __shared__ int64_t sh[SCAN_BLOCK];
  const int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  int64_t v = i < n ? in[i] : 0;
  int64_t tot;
  int64_t ex = block_exclusive_scan(v, sh, &tot);
  if (i < n) out[i] = ex;
  if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// Serial-in-blocks scan of the block sums by one workgroup.
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_sums(int64_t* __restrict__ sums, int64_t nb) {
  __shared__ int64_t sh[SCAN_BLOCK];
  int64_t carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += SCAN_BLOCK) {
    const int64_t i = b0 + threadIdx.x;
    int64_t v = i < nb ? sums[i] : 0;
    int64_t tot;
    int64_t ex = block_exclusive_scan(v, sh, &tot);
    if (i < nb) sums[i] = ex + carry;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ void k_scan_add(int64_t* __restrict__ out, int64_t n, const int64_t* __restrict__ sums) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += sums[i / SCAN_BLOCK];
  if (i == 0) out[n] = sums[(n + SCAN_BLOCK - 1) / SCAN_BLOCK];
}

// Whole exclusive scan in one workgroup (n up to SCAN_SINGLE_MAX = 16 per
// thread): thread t owns the contiguous chunk [t c, (t+1) c), loaded at once into
// registers; the 1024 chunk sums are scanned in LDS and each thread writes its
// chunk's prefixes; out[n] is the total.
constexpr int SCAN_CMAX = 16;
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_single(const int* __restrict__ in, int64_t* __restrict__ out,
                                                           int64_t n) {
  __shared__ int64_t sh[SCAN_BLOCK];
  const int c = (int)((n + SCAN_BLOCK - 1) / SCAN_BLOCK);
  const int64_t b = (int64_t)threadIdx.x * c;
  int v[SCAN_CMAX];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_CMAX; ++k) {
    v[k] = (k < c && b + k < n) ? in[b + k] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t run = block_exclusive_scan(s, sh, &tot);
#pragma unroll
  for (int k = 0; k < SCAN_CMAX; ++k) {
    if (k < c && b + k < n) out[b + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) out[n] = tot;
}

// Exclusive scan of one value per thread over a 1024-thread workgroup: wave
// scans by shuffles, the 16 wave totals scanned by the first wave, two barriers.
__device__ int64_t block_exclusive_scan_waves(int64_t v, int64_t* sh, int64_t* total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  constexpr int NW = SCAN_BLOCK / 64;
  int64_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  if (w == 0) {
    int64_t ws = lane < NW ? sh[lane] : 0;
#pragma unroll
    for (int o = 1; o < NW; o <<= 1) {
      const int64_t y = __shfl_up(ws, o);
      if (lane >= o) ws += y;
    }
    if (lane < NW) sh[NW + lane] = ws;
  }
  __syncthreads();
  *total = sh[2 * NW - 1];
  return (w > 0 ? sh[NW + w - 1] : 0) + x - v;
}

// Two-launch exclusive scan of up to SCAN_2L_MAX values, one or two arrays
// (blockIdx.y picks in0/out0 or in1/out1): k_scan_local2 scans each block of
// 1024 values and stores the block totals; k_scan_add2 adds to every block the
// sum of the totals before it (at most 63 of them, summed by its first
// wavefront), and the last block writes out[n].
constexpr int64_t SCAN_2L_MAX = 64 * (int64_t)SCAN_BLOCK;
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_local2(const int* __restrict__ in0, int64_t* __restrict__ out0,
                                                           const int* __restrict__ in1, int64_t* __restrict__ out1,
                                                           int64_t n, int64_t* __restrict__ sums) {
  __shared__ int64_t sh[2 * SCAN_BLOCK / 64];
  const int* in = blockIdx.y ? in1 : in0;
  int64_t* out = blockIdx.y ? out1 : out0;
  const int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  const int64_t v = i < n ? in[i] : 0;
  int64_t tot;
  const int64_t ex = block_exclusive_scan_waves(v, sh, &tot);
  if (i < n) out[i] = ex;
  if (threadIdx.x == 0) sums[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add2(int64_t* __restrict__ out0, int64_t* __restrict__ out1,
                                                         int64_t n, const int64_t* __restrict__ sums) {
  __shared__ int64_t off;
  int64_t* out = blockIdx.y ? out1 : out0;
  const int64_t* bs = sums + (size_t)blockIdx.y * gridDim.x;
  if (threadIdx.x < 64) {
    int64_t v = 0;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += 64) v += bs[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (threadIdx.x == 0) off = v;
  }
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * SCAN_BLOCK + threadIdx.x;
  if (i < n) out[i] += off;
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = off + bs[blockIdx.x];
}

// Streaming copy, 16 bytes per thread and iteration (the HBM reference).
__global__ __launch_bounds__(256) void k_hbm_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

// Fixed-order sum: each block reduces a contiguous slice in a fixed tree.
__global__ __launch_bounds__(256) void k_sum_partial(const double* __restrict__ x, int64_t n, double* __restrict__ part) {
  __shared__ double sh[256];
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t b = (int64_t)blockIdx.x * per, e = b + per < n ? b + per : n;
  double acc = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += 256) acc += x[i];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace

int tgx_zero(tgx_ctx* ctx, void* p, size_t bytes) {
  if (bytes == 0) return TGX_OK;
  if ((bytes & 3) == 0 && ((uintptr_t)p & 3) == 0) {
    size_t w = bytes / 4;
    unsigned g = (unsigned)((w + 255) / 256);
    if (g > 65536) g = 65536;
    TGX_LAUNCH(ctx, "fill_zero", k_fill_zero, dim3(g), dim3(256), 0, (uint32_t*)p, w);
  } else {
    TGX_LAUNCH(ctx, "fill_zero_bytes", k_fill_zero_bytes, dim3((unsigned)((bytes + 255) / 256)), dim3(256), 0,
               (uint8_t*)p, bytes);
  }
  return TGX_OK;
}

int tgx_scan_i32(tgx_ctx* ctx, const int* in, int64_t* out, int64_t n, int64_t* total, dbuf* tmp) {
  const int64_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
  dbuf& sb = tmp ? *tmp : ctx->scan_tmp;
  TGX_TRY(sb.ensure(sizeof(int64_t) * (size_t)(nb + 2)));
  int64_t* sums = sb.as<int64_t>();
  if (n > 0 && n <= SCAN_SINGLE_MAX) {
    TGX_LAUNCH(ctx, "scan_single", k_scan_single, dim3(1), dim3(SCAN_BLOCK), 0, in, out, n);
  } else if (n > 0 && n <= SCAN_2L_MAX) {
    TGX_LAUNCH(ctx, "scan_local", k_scan_local2, dim3((unsigned)nb, 1), dim3(SCAN_BLOCK), 0, in, out,
               (const int*)nullptr, (int64_t*)nullptr, n, sums);
    TGX_LAUNCH(ctx, "scan_add", k_scan_add2, dim3((unsigned)nb, 1), dim3(SCAN_BLOCK), 0, out, (int64_t*)nullptr, n,
               (const int64_t*)sums);
  } else if (n > 0) {
    TGX_LAUNCH(ctx, "scan_local", k_scan_local, dim3((unsigned)nb), dim3(SCAN_BLOCK), 0, in, out, n, sums);
    TGX_LAUNCH(ctx, "scan_sums", k_scan_sums, dim3(1), dim3(SCAN_BLOCK), 0, sums, nb);
    TGX_LAUNCH(ctx, "scan_add", k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, out, n, sums);
  } else {
    TGX_TRY(tgx_zero(ctx, out, sizeof(int64_t)));
  }
  if (total) {
    TGX_CHECK_HIP(hipMemcpyAsync(total, out + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  return TGX_OK;
}

int tgx_scan2_i32(tgx_ctx* ctx, const int* in0, int64_t* out0, const int* in1, int64_t* out1, int64_t n,
                  dbuf& tmp) {
  if (n > SCAN_SINGLE_MAX && n <= SCAN_2L_MAX) {
    const int64_t nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    TGX_TRY(tmp.ensure(sizeof(int64_t) * 2 * (size_t)nb));
    int64_t* sums = tmp.as<int64_t>();
    TGX_LAUNCH(ctx, "scan_local", k_scan_local2, dim3((unsigned)nb, 2), dim3(SCAN_BLOCK), 0, in0, out0, in1, out1, n,
               sums);
    TGX_LAUNCH(ctx, "scan_add", k_scan_add2, dim3((unsigned)nb, 2), dim3(SCAN_BLOCK), 0, out0, out1, n,
               (const int64_t*)sums);
    return TGX_OK;
  }
  // one after the other on the caller's stream, both through the caller's tmp
  TGX_TRY(tgx_scan_i32(ctx, in0, out0, n, nullptr, &tmp));
  return tgx_scan_i32(ctx, in1, out1, n, nullptr, &tmp);
}

int tgx_device_sum_launch(tgx_ctx* ctx, const double* d_x, int64_t n, double* d_part) {
  TGX_LAUNCH(ctx, "sum_partial", k_sum_partial, dim3(TGX_SUM_PARTS), dim3(256), 0, d_x, n, d_part);
  return TGX_OK;
}

int tgx_device_sum(tgx_ctx* ctx, const double* d_x, int64_t n, double* out_host) {
  const int nbk = TGX_SUM_PARTS;
  TGX_TRY(ctx->red.ensure(sizeof(double) * nbk));
  TGX_LAUNCH(ctx, "sum_partial", k_sum_partial, dim3(nbk), dim3(256), 0, d_x, n, ctx->red.as<double>());
  double part[nbk];
  TGX_CHECK_HIP(hipMemcpyAsync(part, ctx->red.p, sizeof(part), hipMemcpyDeviceToHost, ctx->stream));
  TGX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  double s = 0.0;
  for (int i = 0; i < nbk; ++i) s += part[i];
  *out_host = s;
  return TGX_OK;
}

int tgx_hbm_copy(tgx_ctx* ctx, const void* src, void* dst, size_t bytes) {
  const size_t n16 = bytes / 16;
  TGX_LAUNCH(ctx, "hbm_copy", k_hbm_copy, dim3(4096), dim3(256), 0, reinterpret_cast<const uint4*>(src),
             reinterpret_cast<uint4*>(dst), n16);
  return TGX_OK;
}
