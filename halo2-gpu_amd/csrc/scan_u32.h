// Exclusive prefix sum of uint32 counts on the device (used by the MSM bucket
// sort and its load-balancing levels). out has n + 1 entries; out[n] = total.
#pragma once

#include "common.h"

namespace h2 {

// scratch bytes needed by scan_u32_exclusive for n inputs
uint64_t scan_u32_scratch(uint64_t n);

RustError scan_u32_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, void* scratch, hipStream_t stream);

} // namespace h2
