// Lookup and permutation quotient identities on one extended-domain part,
// folded into the running quotient accumulator with Horner in y. Mirrors
// cuda/src/quotient_lookups.cu + kernel/quotient.h:320-380 and
// cuda/src/permutation_quotient.cu + kernel/permutation_quotient.h:56-148.
// Both are row-parallel streaming kernels (one thread per row, grid-stride).
#include "common.h"
#include "field.h"

using namespace h2;

namespace {

constexpr unsigned kBlock = 256;

__device__ __forceinline__ uint64_t rot(uint64_t idx, int64_t r, int64_t size)
{
    int64_t v = ((int64_t)idx + r) % size;
    if (v < 0)
        v += size;
    return (uint64_t)v;
}

// Lookup identities, in the reference's order (quotient.h:320-380):
//   l0 (1 - Z), l_last (Z^2 - Z), l_active (Z(wX)(A'+b)(S'+g) - Z * table),
//   l0 (A' - S'), l_active (A' - S')(A' - A'(w^-1 X)); rotations wrap mod length.
__global__ void quotient_lookups_kernel(Fr* __restrict__ values, const Fr* __restrict__ table,
    const Fr* __restrict__ z, const Fr* __restrict__ ai, const Fr* __restrict__ st, const Fr* __restrict__ l0,
    const Fr* __restrict__ l_last, const Fr* __restrict__ l_active, const Fr* __restrict__ d_beta,
    const Fr* __restrict__ d_gamma, const Fr* __restrict__ d_y, uint64_t n)
{
    const Fr beta = Fr::load(d_beta), gamma = Fr::load(d_gamma), y = Fr::load(d_y), one = Fr::one();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t rn = rot(i, 1, (int64_t)n), rp = rot(i, -1, (int64_t)n);
        const Fr zi = Fr::load(z + i);
        const Fr a = Fr::load(ai + i);
        const Fr s = Fr::load(st + i);
        const Fr L0 = Fr::load(l0 + i), LA = Fr::load(l_active + i);
        const Fr ams = a - s;
        Fr v = Fr::load(values + i);
        v = v * y + (one - zi) * L0;
        v = v * y + (zi * zi - zi) * Fr::load(l_last + i);
        v = v * y + (Fr::load(z + rn) * (a + beta) * (s + gamma) - zi * Fr::load(table + i)) * LA;
        v = v * y + ams * L0;
        v = v * y + ams * (a - Fr::load(ai + rp)) * LA;
        v.store(values + i);
    }
}

// Permutation identities (permutation_quotient.h:56-148, the CPU restatement is
// src/cpu/evaluator.rs:63-147). `delta_start * ext_omega * omega^row` seeds the
// per-row delta power that advances by delta per column across all sets.
// omega^row for the first row of each thread: omega^(block base) from one
// exponentiation per work-group times omega^threadIdx.x from an LDS table
// built by doubling (8 dependent products), instead of an exponentiation per
// thread; the per-thread stride power is shared the same way.
__global__ void __launch_bounds__(kBlock) quotient_permutation_kernel(Fr* __restrict__ values,
    const Fr* const* __restrict__ zs, const Fr* const* __restrict__ perms, const Fr* const* __restrict__ cols,
    const Fr* __restrict__ l0, const Fr* __restrict__ l_last, const Fr* __restrict__ l_active, Fr beta, Fr gamma,
    Fr y, Fr delta, Fr delta_start, Fr cur_ext_omega, Fr omega, uint64_t n_sets, uint64_t chunk_len,
    uint64_t n_cols, int32_t last_rotation, int32_t rot_scale, int32_t isize, uint64_t n)
{
    __shared__ Fr pw[kBlock + 1];
    __shared__ Fr s_base, s_step;
    const Fr one = Fr::one();
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (threadIdx.x == 0) {
        pw[0] = one;
        pw[1] = omega;
        s_base = cur_ext_omega * Fr::pow_u64(omega, (uint64_t)blockIdx.x * blockDim.x);
        s_step = Fr::pow_u64(omega, stride);
    }
    __syncthreads();
    for (unsigned h = 1; h < kBlock; h <<= 1) { // pw[h + 1 .. 2h] from pw[1 .. h] and pw[h]
        if (threadIdx.x < h)
            pw[h + 1 + threadIdx.x] = pw[1 + threadIdx.x] * pw[h];
        __syncthreads();
    }
    Fr beta_term = s_base * pw[threadIdx.x];
    const Fr wstep = s_step;
    for (uint64_t i = tid; i < n; i += stride) {
        const uint64_t rn = rot(i, (int64_t)rot_scale, isize);
        const uint64_t rl = rot(i, (int64_t)last_rotation * rot_scale, isize);
        const Fr L0 = Fr::load(l0 + i);
        Fr v = Fr::load(values + i);
        const Fr z0 = Fr::load(zs[0] + i);
        v = v * y + (one - z0) * L0;
        const Fr zl = Fr::load(zs[n_sets - 1] + i);
        v = v * y + (zl * zl - zl) * Fr::load(l_last + i);
        for (uint64_t s = 1; s < n_sets; ++s)
            v = v * y + (Fr::load(zs[s] + i) - Fr::load(zs[s - 1] + rl)) * L0;
        const Fr LA = Fr::load(l_active + i);
        Fr cd = delta_start * beta_term;
        for (uint64_t s = 0; s < n_sets; ++s) {
            Fr left = Fr::load(zs[s] + rn);
            Fr right = Fr::load(zs[s] + i);
            const uint64_t c0 = s * chunk_len;
            const uint64_t cnt = (c0 + chunk_len <= n_cols) ? chunk_len : n_cols - c0;
            for (uint64_t j = 0; j < cnt; ++j) {
                const Fr sc = Fr::load(cols[c0 + j] + i) + gamma;
                left = left * (sc + beta * Fr::load(perms[c0 + j] + i));
                right = right * (sc + cd);
                cd = cd * delta;
            }
            v = v * y + (left - right) * LA;
        }
        v.store(values + i);
        beta_term = beta_term * wstep;
    }
}

Fr host_fr(const void* p)
{
    Fr r;
    memcpy(r.v, p, 32);
    return r;
}

} // namespace

extern "C" {

RustError _halo2_quotient_lookups(void* values, const void* table_values, const void* product_coset,
    const void* permuted_input_coset, const void* permuted_table_coset, const void* l0, const void* l_last,
    const void* l_active_row, const void* beta, const void* gamma, const void* y, size_t poly_length, void* stream)
{
    if (poly_length == 0)
        return ok();
    quotient_lookups_kernel<<<grid_for(poly_length, kBlock, 256u * 8u), kBlock, 0, (hipStream_t)stream>>>(
        (Fr*)values, (const Fr*)table_values, (const Fr*)product_coset, (const Fr*)permuted_input_coset,
        (const Fr*)permuted_table_coset, (const Fr*)l0, (const Fr*)l_last, (const Fr*)l_active_row, (const Fr*)beta,
        (const Fr*)gamma, (const Fr*)y, poly_length);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_quotient_permutation(void* values, const void* l0, const void* l_last, const void* l_active_row,
    const void* perm_prod_cosets, const void* perm_cosets, const void* column_values, const void* beta,
    const void* gamma, const void* y, const void* delta, const void* delta_start, const void* current_extended_omega,
    const void* omega, uint64_t n_sets, uint64_t chunk_len, uint64_t n_perm_cols, int32_t last_rotation,
    int32_t rot_scale, int32_t isize_, uint64_t poly_length, void* stream)
{
    if (n_sets == 0 || poly_length == 0)
        return ok();
    H2_CHECK_ARG(chunk_len > 0 && n_perm_cols <= n_sets * chunk_len && isize_ > 0,
        "_halo2_quotient_permutation: inconsistent set layout (sets=%lu chunk=%lu cols=%lu isize=%d)",
        (unsigned long)n_sets, (unsigned long)chunk_len, (unsigned long)n_perm_cols, isize_);
    quotient_permutation_kernel<<<grid_for(poly_length, kBlock, 256u * 8u), kBlock, 0, (hipStream_t)stream>>>(
        (Fr*)values, (const Fr* const*)perm_prod_cosets, (const Fr* const*)perm_cosets,
        (const Fr* const*)column_values, (const Fr*)l0, (const Fr*)l_last, (const Fr*)l_active_row, host_fr(beta),
        host_fr(gamma), host_fr(y), host_fr(delta), host_fr(delta_start), host_fr(current_extended_omega),
        host_fr(omega), n_sets, chunk_len, n_perm_cols, last_rotation, rot_scale, isize_, poly_length);
    H2_LAUNCH_CHECK();
    return ok();
}

} // extern "C"
