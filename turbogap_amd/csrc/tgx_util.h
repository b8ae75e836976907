// Device utilities shared by the libtgx translation units.
#pragma once
#include "tgx_internal.h"

// Exclusive prefix sum of n int32 values into int64 offsets; out has n + 1
// entries (out[n] is the total). When total is not NULL the total is copied to
// the host (this synchronises the stream).
int tgx_scan_i32(tgx_ctx* ctx, const int* in, int64_t* out, int64_t n, int64_t* total);

// Streaming device-to-device copy of `bytes` (a multiple of 16) by k_hbm_copy.
int tgx_hbm_copy(tgx_ctx* ctx, const void* src, void* dst, size_t bytes);
