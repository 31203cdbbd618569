// Lookup and permutation quotient identities on one extended-domain part,
// folded into the running quotient accumulator with Horner in y. Mirrors
// cuda/src/quotient_lookups.cu + kernel/quotient.h:320-380 and
// cuda/src/permutation_quotient.cu + kernel/permutation_quotient.h:56-148.
// Both are row-parallel streaming kernels (one thread per row, grid-stride) on
// the 29-bit limb form of field29.h: stored words (x * 2^256) stay in that
// scale, the Horner factor y and the other per-call constants that multiply a
// stored word are moved to the 2^261 scale once (shl5), and a product of two
// stored words is mul(shl5(a), b). The bounds in the comments are multiples of
// m = r, using 32 m < 0.19 * 2^261.
#include "common.h"
#include "field.h"
#include "field29.h"

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

using R29 = Fr29;

__device__ __forceinline__ R29 ld29(const Fr* p) { return R29::load32(p); }
// a per-call constant, the same in every lane, held in scalar registers: the products
// take it as their scalar operand, and the vector registers it would occupy go to
// more resident waves
__device__ __forceinline__ R29 uniform(R29 x)
{
#pragma unroll
    for (int w = 0; w < 9; ++w)
        x.l[w] = __builtin_amdgcn_readfirstlane(x.l[w]);
    return x;
}
__device__ __forceinline__ R29 c261(const R29& stored) { return R29::shl5(stored); } // < 32m, normalised
// v * y + t (v < 4.1m normalised, t < 2.3m normalised): < 4.1m (v -> 3.3m + 0.19 v), normalised
__device__ __forceinline__ R29 horner(const R29& v, const R29& y32, const R29& t)
{
    return R29::normalize(R29::add(R29::mul(v, y32), t));
}
// a * b for a < 7m normalised (limbs < 2^29) or with limbs < 2^31, b a stored word: < 2.4m
__device__ __forceinline__ R29 times_stored(const R29& a, const Fr* b) { return R29::mul(a, c261(ld29(b))); }

// Lookup identities, in the reference's order (quotient.h:320-380):
//   l0 (1 - Z), l_last (Z^2 - Z), l_active (Z(wX)(A'+b)(S'+g) - Z * table),
//   l0 (A' - S'), l_active (A' - S')(A' - A'(w^-1 X)); rotations wrap mod length.
__global__ void __launch_bounds__(kBlock) quotient_lookups_kernel(Fr* __restrict__ values,
    const Fr* __restrict__ table, const Fr* __restrict__ z, const Fr* __restrict__ ai, const Fr* __restrict__ st,
    const Fr* __restrict__ l0, const Fr* __restrict__ l_last, const Fr* __restrict__ l_active,
    const Fr* __restrict__ d_beta, const Fr* __restrict__ d_gamma, const Fr* __restrict__ d_y, uint64_t n)
{
    const R29 beta = uniform(ld29(d_beta)), gamma = uniform(ld29(d_gamma)), y32 = uniform(c261(ld29(d_y)));
    const R29 one = R29::from_fp(Fr::one());
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t rn = rot(i, 1, (int64_t)n), rp = rot(i, -1, (int64_t)n);
        const R29 zi = ld29(z + i), a = ld29(ai + i), s = ld29(st + i);
        const R29 ams = R29::sub2(a, s); // < 3m
        R29 v = ld29(values + i);
        v = horner(v, y32, times_stored(R29::sub2(one, zi), l0 + i));
        const R29 zz = R29::mul(c261(zi), zi); // < 1.2m
        v = horner(v, y32, times_stored(R29::sub2(zz, zi), l_last + i));
        // Z(wX) (A' + b)(S' + g) - Z table: (A' + b) < 2m -> 64m in the 2^261 scale
        const R29 ab = R29::mul(c261(R29::normalize(R29::add(a, beta))), R29::add(s, gamma)); // < 1.8m
        const R29 lhs = R29::mul(c261(ld29(z + rn)), ab); // < 1.4m
        const R29 rhs = R29::mul(c261(zi), ld29(table + i)); // < 1.2m
        v = horner(v, y32, times_stored(R29::sub2(lhs, rhs), l_active + i));
        v = horner(v, y32, times_stored(ams, l0 + i));
        const R29 q = R29::mul(c261(R29::normalize(ams)), R29::sub2(a, ld29(ai + rp))); // 96m * 3m: < 2.8m
        v = horner(v, y32, times_stored(q, l_active + i));
        R29::reduce_full(v).store32(values + i);
    }
}

// Permutation identities (permutation_quotient.h:56-148, the CPU restatement is
// src/cpu/evaluator.rs:63-147). `delta_start * ext_omega * omega^row` seeds the
// per-row delta power that advances by delta per column across all sets.
// omega^row for the first row of each thread: omega^(block base) from one
// exponentiation per work-group times omega^threadIdx.x from an LDS table
// built by doubling (8 dependent products), instead of an exponentiation per
// thread; the per-thread stride power is shared the same way.
// Running products: z * prod (c + beta s + gamma) with factors < 3.4m stays
// below 2.6m (x -> m + 0.61 x), so its 2^261-scale image stays below 2^261.
__global__ void __launch_bounds__(kBlock) quotient_permutation_kernel(Fr* __restrict__ values,
    const Fr* const* __restrict__ zs, const Fr* const* __restrict__ perms, const Fr* const* __restrict__ cols,
    const Fr* __restrict__ l0, const Fr* __restrict__ l_last, const Fr* __restrict__ l_active, Fr beta, Fr gamma,
    Fr y, Fr delta, Fr delta_start, Fr cur_ext_omega, Fr omega, uint64_t n_sets, uint64_t chunk_len,
    uint64_t n_cols, int32_t last_rotation, int32_t rot_scale, int32_t isize, uint64_t n)
{
    __shared__ Fr pw[kBlock + 1];
    __shared__ Fr s_base, s_step;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (threadIdx.x == 0) {
        pw[0] = Fr::one();
        pw[1] = omega;
        s_base = cur_ext_omega * Fr::pow_u64(omega, (uint64_t)blockIdx.x * blockDim.x);
    } else if (threadIdx.x == 64) { // the other exponentiation on the second wave, beside it
        s_step = Fr::pow_u64(omega, stride);
    }
    __syncthreads();
    for (unsigned h = 1; h < kBlock; h <<= 1) { // pw[h + 1 .. 2h] from pw[1 .. h] and pw[h]
        if (threadIdx.x < h)
            pw[h + 1 + threadIdx.x] = pw[1 + threadIdx.x] * pw[h];
        __syncthreads();
    }
    R29 beta_term = R29::from_fp(s_base * pw[threadIdx.x]); // ext_omega omega^row, < 1.2m
    const R29 wstep32 = uniform(c261(R29::from_fp(s_step))), y32 = uniform(c261(R29::from_fp(y)));
    const R29 beta32 = uniform(c261(R29::from_fp(beta))), gamma29 = uniform(R29::from_fp(gamma));
    const R29 delta32 = uniform(c261(R29::from_fp(delta))), dstart32 = uniform(c261(R29::from_fp(delta_start)));
    const R29 one = R29::from_fp(Fr::one());
    for (uint64_t i = tid; i < n; i += stride) {
        const uint64_t rn = rot(i, (int64_t)rot_scale, isize);
        const uint64_t rl = rot(i, (int64_t)last_rotation * rot_scale, isize);
        R29 v = ld29(values + i);
        v = horner(v, y32, times_stored(R29::sub2(one, ld29(zs[0] + i)), l0 + i));
        const R29 zl = ld29(zs[n_sets - 1] + i);
        v = horner(v, y32, times_stored(R29::sub2(R29::mul(c261(zl), zl), zl), l_last + i));
        for (uint64_t s = 1; s < n_sets; ++s)
            v = horner(v, y32, times_stored(R29::sub2(ld29(zs[s] + i), ld29(zs[s - 1] + rl)), l0 + i));
        R29 cd = R29::mul(dstart32, beta_term); // < 1.4m
        for (uint64_t s = 0; s < n_sets; ++s) {
            R29 left = ld29(zs[s] + rn);
            R29 right = ld29(zs[s] + i);
            const uint64_t c0 = s * chunk_len;
            const uint64_t cnt = (c0 + chunk_len <= n_cols) ? chunk_len : n_cols - c0;
            for (uint64_t j = 0; j < cnt; ++j) {
                const R29 sc = R29::add(ld29(cols[c0 + j] + i), gamma29); // < 2m, limbs < 2^30
                const R29 fl = R29::add(sc, R29::mul(beta32, ld29(perms[c0 + j] + i))); // < 3.2m, limbs < 2^31
                const R29 fr = R29::add(sc, cd); // < 3.4m, limbs < 2^31
                left = R29::mul(c261(left), fl);
                right = R29::mul(c261(right), fr);
                cd = R29::mul(cd, delta32);
            }
            const R29 diff = R29::normalize(R29::sub4(left, right)); // < 6.9m
            v = horner(v, y32, times_stored(diff, l_active + i));
        }
        R29::reduce_full(v).store32(values + i);
        beta_term = R29::mul(beta_term, wstep32);
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
