// Device utilities shared by the libtgx translation units.
#pragma once
#include "tgx_internal.h"

// Exclusive prefix sum of n int32 values into int64 offsets; out has n + 1
// entries (out[n] is the total). When total is not NULL the total is copied to
// the host (this synchronises the stream). The block totals go to tmp, or to
// the context's own scratch when tmp is NULL; scans that may run on different
// streams at once must each bring their own.
int tgx_scan_i32(tgx_ctx* ctx, const int* in, int64_t* out, int64_t n, int64_t* total, dbuf* tmp = nullptr);
// The same for two arrays of n values (two launches for both when 16,384 < n <=
// 65,536, four or six otherwise); block totals in tmp, which belongs to the
// caller so that scans on different streams do not share it.
int tgx_scan2_i32(tgx_ctx* ctx, const int* in0, int64_t* out0, const int* in1, int64_t* out1, int64_t n,
                  dbuf& tmp);

// Partial sums of n doubles into TGX_SUM_PARTS device values, each over a fixed
// contiguous slice in a fixed tree (summed on the host in order).
constexpr int TGX_SUM_PARTS = 256;
int tgx_device_sum_launch(tgx_ctx* ctx, const double* d_x, int64_t n, double* d_part);

// Streaming device-to-device copy of `bytes` (a multiple of 16) by k_hbm_copy.
int tgx_hbm_copy(tgx_ctx* ctx, const void* src, void* dst, size_t bytes);
