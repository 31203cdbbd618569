// Three-phase exclusive scan: per-block scans (1024 threads x 4 elements),
// a recursive scan of the block totals, then a uniform add per block.
#include "scan_u32.h"

namespace h2 {

namespace {

constexpr unsigned kThreads = 1024;
constexpr unsigned kPer = 4;
constexpr unsigned kTile = kThreads * kPer;

__global__ void __launch_bounds__(kThreads) scan_tile_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
    uint64_t n, uint32_t* __restrict__ block_sums)
{
    __shared__ uint32_t s[kThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (unsigned k = 0; k < kPer; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        sum += v[k];
    }
    s[threadIdx.x] = sum;
    for (unsigned d = 1; d < kThreads; d <<= 1) {
        __syncthreads();
        uint32_t t = threadIdx.x >= d ? s[threadIdx.x - d] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
    }
    __syncthreads();
    uint32_t run = threadIdx.x ? s[threadIdx.x - 1] : 0u;
#pragma unroll
    for (unsigned k = 0; k < kPer; ++k) {
        if (base + k < n)
            out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kThreads - 1)
        block_sums[blockIdx.x] = s[kThreads - 1];
}

__global__ void add_block_kernel(uint32_t* __restrict__ out, uint64_t n, const uint32_t* __restrict__ block_pre)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] += block_pre[i / kTile];
}

__global__ void write_total_kernel(uint32_t* out, uint64_t n, const uint32_t* in)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        out[n] = (n ? out[n - 1] + in[n - 1] : 0u);
}

} // namespace

uint64_t scan_u32_scratch(uint64_t n)
{
    uint64_t bytes = 0;
    while (n > 1) {
        uint64_t nb = (n + kTile - 1) / kTile;
        bytes += align_up(4 * (nb + 1), 256) * 2;
        n = nb;
    }
    return bytes + 256;
}

static RustError scan_rec(const uint32_t* in, uint32_t* out, uint64_t n, uint8_t* scratch, hipStream_t stream)
{
    const uint64_t nb = (n + kTile - 1) / kTile;
    uint32_t* sums = (uint32_t*)scratch;
    uint32_t* sums_pre = (uint32_t*)(scratch + align_up(4 * (nb + 1), 256));
    uint8_t* next = scratch + 2 * align_up(4 * (nb + 1), 256);
    scan_tile_kernel<<<(unsigned)nb, kThreads, 0, stream>>>(in, out, n, sums);
    H2_LAUNCH_CHECK();
    if (nb > 1) {
        RustError e = scan_rec(sums, sums_pre, nb, next, stream);
        if (e.code)
            return e;
        add_block_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(out, n, sums_pre);
        H2_LAUNCH_CHECK();
    }
    return ok();
}

RustError scan_u32_exclusive(const uint32_t* in, uint32_t* out, uint64_t n, void* scratch, hipStream_t stream)
{
    if (n == 0) {
        H2_TRY(hipMemsetAsync(out, 0, 4, stream));
        return ok();
    }
    RustError e = scan_rec(in, out, n, (uint8_t*)scratch, stream);
    if (e.code)
        return e;
    write_total_kernel<<<1, 64, 0, stream>>>(out, n, in);
    H2_LAUNCH_CHECK();
    return ok();
}

} // namespace h2
