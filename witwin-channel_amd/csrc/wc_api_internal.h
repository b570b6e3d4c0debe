// Error reporting shared by every C-ABI entry point: a non-zero status plus a
// thread-local message readable through wc_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstddef>

// printf-style message; returns 1
int wc_fail(const char* format, ...) __attribute__((format(printf, 1, 2)));
int wc_fail_hip(const char* where, hipError_t err);

namespace wc {
int radix_sort_pairs(unsigned int* keys, int* values, unsigned int* keys_tmp, int* values_tmp, int* hist, int n,
                     int key_bits, hipStream_t stream);
size_t radix_sort_hist_ints(int n);
// exclusive scan of n ints into out[0..n] (out[n] = total), *total (device
// int64) = the sum; workspace of scan_exclusive_workspace_bytes(n) bytes
size_t scan_exclusive_workspace_bytes(long long n);
int scan_exclusive(const int* values, int n, int* out, void* workspace, long long* total, hipStream_t stream);
}  // namespace wc
