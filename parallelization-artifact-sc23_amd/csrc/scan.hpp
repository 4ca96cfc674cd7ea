// Device-wide exclusive prefix sum (int32).
#pragma once

#include <cstddef>

#include "runtime.hpp"

namespace indigo {

// Bytes of device scratch `temp` needed for n values.
size_t scan_temp_bytes(int n);
// out[i] = sum(in[0..i)), out[n] = total.  Three launches on `stream`.
int exclusive_scan(hipStream_t stream, const int* d_in, int* d_out, int n, int* d_temp);

}  // namespace indigo
