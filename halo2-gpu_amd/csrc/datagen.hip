// Seeded device-side input generation for benchmarks and large tests: uniform
// Fr elements from a ChaCha20 keystream, produced directly in HBM.
//
// Element i is ChaCha20 block (start + i) (RFC 8439 block function, 64-bit
// block counter, zero stream id) under the key that rand_core's
// `seed_from_u64(seed)` derives with PCG32; its 64 bytes are read as a 512-bit
// little-endian integer and reduced mod r into Montgomery form exactly like
// halo2curves' `Fr::from_uniform_bytes` (d0 * R^2 + d1 * R^3, Montgomery
// products). That is the byte stream `Fr::random(ChaCha20Rng::seed_from_u64)`
// consumes, one block per element.
#include "common.h"
#include "field.h"

using namespace h2;

namespace {

struct ChachaKey {
    uint32_t k[8];
};

__device__ __forceinline__ uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }

#define QR(a, b, c, d)          \
    a += b;                     \
    d = rotl(d ^ a, 16);        \
    c += d;                     \
    b = rotl(b ^ c, 12);        \
    a += b;                     \
    d = rotl(d ^ a, 8);         \
    c += d;                     \
    b = rotl(b ^ c, 7);

__global__ void random_fr_kernel(Fr* __restrict__ out, uint64_t n, ChachaKey key, uint64_t start)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t ctr = start + i;
    uint32_t s[16] = { 0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key.k[0], key.k[1], key.k[2], key.k[3],
        key.k[4], key.k[5], key.k[6], key.k[7], (uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u };
    uint32_t x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
        x[j] = s[j];
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    Fr d0, d1;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        d0.v[j] = x[j] + s[j];
        d1.v[j] = x[8 + j] + s[8 + j];
    }
    Fr r3;
    const uint32_t R3[8] = { 0xb4bf0040u, 0x5e94d8e1u, 0x1cfbb6b8u, 0x2a489cbeu, 0xa19fcfedu, 0x893cc664u,
        0x7fcc657cu, 0x0cf8594bu }; // 2^768 mod r
#pragma unroll
    for (int j = 0; j < 8; ++j)
        r3.v[j] = R3[j];
    (Fr::mul(d0, Fr::r2()) + Fr::mul(d1, r3)).store(out + i);
}

ChachaKey seed_key(uint64_t seed)
{
    // rand_core 0.6 SeedableRng::seed_from_u64: PCG32 (XSH RR) words fill the 32-byte seed
    const uint64_t MUL = 6364136223846793005ull, INC = 11634580027462260723ull;
    ChachaKey k;
    uint64_t st = seed;
    for (int i = 0; i < 8; ++i) {
        st = st * MUL + INC;
        uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27);
        uint32_t r = (uint32_t)(st >> 59);
        k.k[i] = (xs >> r) | (xs << ((32 - r) & 31));
    }
    return k;
}

} // namespace

extern "C" RustError h2amd_random_fr(void* d_out, uint64_t n, uint64_t seed, uint64_t start, void* stream)
{
    if (n == 0)
        return ok();
    H2_CHECK_ARG(n < (1ull << 40), "h2amd_random_fr: n too large");
    random_fr_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>((Fr*)d_out, n, seed_key(seed),
        start);
    H2_LAUNCH_CHECK();
    return ok();
}
