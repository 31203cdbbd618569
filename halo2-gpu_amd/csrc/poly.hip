// Element-wise polynomial helpers of the quotient path and the omega-powers
// table, mirroring cuda/src/polynomial/polynomial.cu:247-318 and
// cuda/src/polynomial/omega.cu of the reference. All are HBM-streaming
// kernels: one 32-byte read and write per element plus a tiny table that stays
// in L1/L2.
#include "common.h"
#include "field.h"

using namespace h2;

namespace {

constexpr unsigned kBlock = 256;

unsigned stream_grid(uint64_t n)
{
    // enough work-groups to cover all 256 CUs several times; grid-stride beyond
    return grid_for(n, kBlock, 256u * 16u);
}

__global__ void extended_from_lagrange_vec_kernel(Fr* __restrict__ out, const Fr* const* __restrict__ parts,
    uint32_t num_parts, uint64_t n)
{
    const uint64_t total = n * num_parts;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += stride) {
        const uint64_t i = j / num_parts;
        const uint32_t p = (uint32_t)(j - i * num_parts);
        Fr::load(parts[p] + i).store(out + j);
    }
}

__global__ void divide_by_vanishing_kernel(Fr* __restrict__ a, const Fr* __restrict__ t, uint32_t t_len, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        (Fr::load(a + i) * Fr::load(t + (i % t_len))).store(a + i);
}

__global__ void distribute_powers_zeta_kernel(Fr* __restrict__ a, const Fr* __restrict__ cp, uint32_t len, uint64_t n)
{
    const uint32_t mod = len + 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t k = (uint32_t)(i % mod);
        if (k)
            (Fr::load(a + i) * Fr::load(cp + (k - 1))).store(a + i);
    }
}

// out[i] = omega^i: thread i builds omega^(i mod 2^8) from a 256-entry LDS table
// times omega^(256 * (i >> 8)) computed by square-and-multiply once per block.
__global__ void omega_powers_kernel(Fr* __restrict__ out, const Fr* __restrict__ d_omega, uint64_t n)
{
    __shared__ Fr lo[256];
    __shared__ Fr blk;
    const Fr w = Fr::load(d_omega);
    const uint64_t base = (uint64_t)blockIdx.x * 256;
    if (threadIdx.x == 0) {
        Fr w256 = w;
        for (int s = 0; s < 8; ++s)
            w256 = Fr::sqr(w256);
        blk = Fr::pow_u64(w256, blockIdx.x);
    }
    lo[threadIdx.x] = Fr::pow_u64(w, threadIdx.x);
    __syncthreads();
    const uint64_t i = base + threadIdx.x;
    if (i < n)
        (blk * lo[threadIdx.x]).store(out + i);
}

} // namespace

extern "C" {

RustError _halo2_extended_from_lagrange_vec_device(void* d_out, const void* d_parts, uint32_t num_parts, uint64_t n,
    void* stream)
{
    if (n == 0 || num_parts == 0)
        return ok();
    extended_from_lagrange_vec_kernel<<<stream_grid(n * num_parts), kBlock, 0, (hipStream_t)stream>>>(
        (Fr*)d_out, (const Fr* const*)d_parts, num_parts, n);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_divide_by_vanishing_poly(void* d_poly, const void* d_t_evals, uint32_t t_len, uint64_t n,
    void* stream)
{
    if (n == 0)
        return ok();
    H2_CHECK_ARG(t_len > 0, "_halo2_divide_by_vanishing_poly: t_len must be > 0");
    divide_by_vanishing_kernel<<<stream_grid(n), kBlock, 0, (hipStream_t)stream>>>((Fr*)d_poly, (const Fr*)d_t_evals,
        t_len, n);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_distribute_powers_zeta(void* d_a, const void* d_coset_powers, uint32_t coset_powers_len, uint64_t n,
    void* stream)
{
    if (n == 0 || coset_powers_len == 0)
        return ok();
    distribute_powers_zeta_kernel<<<stream_grid(n), kBlock, 0, (hipStream_t)stream>>>((Fr*)d_a,
        (const Fr*)d_coset_powers, coset_powers_len, n);
    H2_LAUNCH_CHECK();
    return ok();
}

uint64_t _halo2_generate_omega_powers_workspace_size(uint32_t log_n)
{
    (void)log_n;
    return 0;
}

RustError _halo2_generate_omega_powers(void* d_omega_powers, const void* d_omega, uint32_t log_n, void* scratch,
    uint64_t scratch_bytes, void* stream)
{
    (void)scratch;
    (void)scratch_bytes;
    H2_CHECK_ARG(log_n <= 28, "_halo2_generate_omega_powers: log_n %u > 28", log_n);
    const uint64_t n = 1ull << log_n;
    omega_powers_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>((Fr*)d_omega_powers,
        (const Fr*)d_omega, n);
    H2_LAUNCH_CHECK();
    return ok();
}

} // extern "C"
