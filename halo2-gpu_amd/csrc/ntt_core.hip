// Fr NTT engine: table generation and the LDS-staged Stockham pass kernel.
// See ntt_core.h for the decomposition.
//
// Arithmetic inside the pass is the 9 x 29-bit lazy-reduction form of
// field29.h: elements stay in the boundary's 2^256 Montgomery scale and every
// constant the pass multiplies by (twiddles, coset powers, zeta, the iFFT
// scale) is held in the 2^261 scale.
#include <atomic>
#include <mutex>
#include <vector>

#include "field29.h"
#include "ntt_core.h"
#include "profile.h"

namespace h2 {

// zeta (a primitive cube root of unity, halo2's g_coset) and zeta^2 times
// 2^261 mod r, in 29-bit limbs. The reference hard-codes zeta and zeta^2 in
// 2^256 Montgomery form at cuda/include/kernel/zeta.h:19-30; these are the
// same field elements rescaled.
__device__ __forceinline__ Fr29 zeta261(uint32_t m)
{
    constexpr uint32_t z1[9] = { 0x0f9aca5du, 0x06c39377u, 0x03fb9328u, 0x0dd12516u, 0x19b77045u, 0x1b53b497u,
        0x086e5ab0u, 0x0b729bf7u, 0x00226320u };
    constexpr uint32_t z2[9] = { 0x1065364du, 0x19a50e73u, 0x05348a9cu, 0x02af1c94u, 0x11265ae2u, 0x0a454b5au,
        0x06633a88u, 0x1dbec294u, 0x000038f7u };
    Fr29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i)
        r.l[i] = m == 1 ? z1[i] : z2[i];
    return r;
}

// x * 2^256 -> x * 2^261 (mod r) on the host: one Montgomery product with
// 2^261 mod r (CIOS over 32-bit words).
static void host_to261(const Fr& x, uint32_t (&out)[9])
{
    const uint32_t* a = x.v;
    const uint32_t* b = Fr29Params::TO261;
    const uint32_t* m = FrParams::M;
    uint32_t t[10] = { 0 };
    for (int i = 0; i < 8; ++i) {
        uint64_t c = 0;
        for (int j = 0; j < 8; ++j) {
            c += (uint64_t)t[j] + (uint64_t)a[j] * b[i];
            t[j] = (uint32_t)c;
            c >>= 32;
        }
        c += t[8];
        t[8] = (uint32_t)c;
        t[9] = (uint32_t)(c >> 32);
        const uint32_t q = t[0] * FrParams::INV;
        c = ((uint64_t)t[0] + (uint64_t)q * m[0]) >> 32;
        for (int j = 1; j < 8; ++j) {
            c += (uint64_t)t[j] + (uint64_t)q * m[j];
            t[j - 1] = (uint32_t)c;
            c >>= 32;
        }
        c += t[8];
        t[7] = (uint32_t)c;
        t[8] = t[9] + (uint32_t)(c >> 32);
    }
    // t < 2m: reduce once
    uint32_t d[8];
    int64_t br = 0;
    for (int j = 0; j < 8; ++j) {
        const int64_t v = (int64_t)t[j] - m[j] + br;
        d[j] = (uint32_t)v;
        br = v >> 32;
    }
    const bool ge = t[8] != 0 || br == 0;
    uint32_t w[8];
    for (int j = 0; j < 8; ++j)
        w[j] = ge ? d[j] : t[j];
    for (int i = 0; i < 8; ++i) {
        const int wi = (29 * i) >> 5, s = (29 * i) & 31;
        const uint64_t two = ((uint64_t)w[wi + 1] << 32) | w[wi];
        out[i] = (uint32_t)(two >> s) & 0x1fffffffu;
    }
    out[8] = w[7] >> 8;
}

// Tuned work-group tile sizes: log2 of the elements per work-group for
// transforms of 2^log_n (0 = the default rule). Set with h2amd_ntt_set_tiles.
// Atomic entries: a setter on one host thread may race planners on others.
static std::atomic<uint8_t> g_tile_table[NTT_MAX_LOG + 1] = {};

NttPlan ntt_plan(uint32_t log_n)
{
    NttPlan p {};
    if (log_n <= NTT_MAX_LOG_R) {
        p.P = 1;
        p.log_r[0] = log_n;
    } else if (log_n <= 2 * NTT_MAX_LOG_R) {
        p.P = 2;
        p.log_r[0] = (log_n + 1) / 2;
        p.log_r[1] = log_n - p.log_r[0];
    } else {
        p.P = 3;
        p.log_r[0] = (log_n + 2) / 3;
        p.log_r[1] = (log_n - p.log_r[0] + 1) / 2;
        p.log_r[2] = log_n - p.log_r[0] - p.log_r[1];
    }
    for (int i = 0; i < p.P; ++i) {
        // aim for >= 256 work-groups per transform, <= 4096 elements per work-group
        int g = (int)log_n - (int)p.log_r[i] - 8;
        int gmax = (int)NTT_MAX_LOG_R - (int)p.log_r[i];
        if (g > gmax)
            g = gmax;
        if (g < 0)
            g = 0;
        const uint8_t tuned = log_n <= NTT_MAX_LOG ? g_tile_table[log_n].load(std::memory_order_relaxed) : 0;
        if (tuned) {
            g = (int)tuned - (int)p.log_r[i];
            if (g > gmax)
                g = gmax;
            if (g > (int)log_n - (int)p.log_r[i])
                g = (int)log_n - (int)p.log_r[i];
            if (g < 0)
                g = 0;
        }
        if (p.P == 1)
            g = 0;
        p.log_g[i] = (uint32_t)g;
    }
    return p;
}

bool ntt_needs_scratch(uint32_t log_n, bool aliased)
{
    NttPlan p = ntt_plan(log_n);
    return p.P == 3 || (p.P == 2 && aliased);
}

static inline uint32_t split_h(uint32_t log_n) { return (log_n + 1) / 2; }

void ntt_tables_size(WorkspaceSizer& ws, uint32_t log_n, bool powers_twist)
{
    NttPlan p = ntt_plan(log_n);
    uint32_t h = split_h(log_n);
    ws.take(32ull << h);
    ws.take(32ull << (log_n - h));
    for (int i = 0; i < p.P; ++i)
        ws.take(32ull * ((1ull << p.log_r[i]) / 2 + 1));
    if (powers_twist) {
        ws.take(32ull << h);
        ws.take(32ull << (log_n - h));
    }
}

struct TableJob {
    Fr* out;
    uint32_t shift;
    uint32_t count;
    uint32_t use_c;
};
struct TableJobs {
    TableJob j[7];
};

// out[i] = (base^(2^shift))^i * 2^261 / 2^256, base = omega (device) or c (by
// value): the tables hold the 2^261-scale representatives the pass multiplies by.
__global__ void ntt_tables_kernel(TableJobs tj, const Fr* d_omega, Fr c)
{
    const TableJob job = tj.j[blockIdx.y];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= job.count)
        return;
    Fr b = job.use_c ? c : Fr::load(d_omega);
    for (uint32_t s = 0; s < job.shift; ++s)
        b = Fr::sqr(b);
    Fr to;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        to.v[k] = Fr29Params::TO261[k];
    (Fr::pow_u64(b, i) * to).store(job.out + i);
}

RustError ntt_tables_build(Workspace& ws, NttTables& t, uint32_t log_n, const Fr* d_omega, bool powers_twist,
    const Fr& c, hipStream_t stream)
{
    NttPlan p = ntt_plan(log_n);
    uint32_t h = split_h(log_n);
    TableJobs tj {};
    int nj = 0;
    uint32_t maxc = 0;
    auto add = [&](uint64_t count, uint32_t shift, uint32_t use_c) -> Fr* {
        Fr* out = (Fr*)ws.take(32ull * count);
        tj.j[nj++] = TableJob { out, shift, (uint32_t)count, use_c };
        if (count > maxc)
            maxc = (uint32_t)count;
        return out;
    };
    t.h = h;
    t.tw_lo = add(1ull << h, 0, 0);
    t.tw_hi = add(1ull << (log_n - h), h, 0);
    for (int i = 0; i < p.P; ++i)
        t.twr[i] = add((1ull << p.log_r[i]) / 2 + 1, log_n - p.log_r[i], 0);
    if (powers_twist) {
        t.c_lo = add(1ull << h, 0, 1);
        t.c_hi = add(1ull << (log_n - h), h, 1);
    }
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "NTT workspace too small for the twiddle tables (log_n=%u)", log_n);
    dim3 grid((maxc + 127) / 128, nj);
    ntt_tables_kernel<<<grid, 128, 0, stream>>>(tj, d_omega, c);
    H2_LAUNCH_CHECK();
    return ok();
}

struct PassBatch {
    const Fr* in[NTT_MAX_BATCH];
    Fr* out[NTT_MAX_BATCH];
};

struct PassParams {
    uint32_t log_n, log_r, log_g, log_l;
    uint64_t in_len;
    const Fr* tw_lo;
    const Fr* tw_hi;
    const Fr* twr;
    const Fr* c_lo;
    const Fr* c_hi;
    uint32_t h;
    int twist;
    int post;
    int last;
    uint32_t xcd_remap;
    uint32_t scale[9]; // iFFT scale in the 2^261 scale
};

// Slot i of the tile lives at word lds_slot(i) of each plane: the low 6 bits of i
// XOR a function of the higher bits, a permutation inside every 64-slot block (so a
// range of whole blocks keeps its words). In slot order the load's bit-reversed
// writes put all 64 lanes of a wave on one bank (64-way conflicts), the first stage
// pairs' strided accesses 4-8 lanes per bank, and the multi-column store up to 16;
// through this permutation every one of these patterns, for every tile shape the
// plans use, spreads a wave over distinct banks (64 banks; 2 lanes per bank with 32),
// as enumerating the patterns shows.
__device__ __forceinline__ uint32_t lds_slot(uint32_t i)
{
    const uint32_t x = i >> 6;
    return i ^ (((x ^ (x >> 6)) & 63u) ^ ((x & 3u) * 20u));
}

// LDS holds the tile as 9 limb planes: limb q of (physical) word i at lds[q * RG + i]
__device__ __forceinline__ Fr29 lds_get(const uint32_t* lds, uint32_t RG, uint32_t i)
{
    Fr29 r;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        r.l[q] = lds[q * RG + i];
    return r;
}
__device__ __forceinline__ void lds_put(uint32_t* lds, uint32_t RG, uint32_t i, const Fr29& x)
{
#pragma unroll
    for (int q = 0; q < 9; ++q)
        lds[q * RG + i] = x.l[q];
}

// Value bounds (m = r): the load writes normalised limbs of values < 1.25m
// (inputs are < 2^256 and a twisted input is a product, < 1.25m). Stage 0
// forms a + c and a + 2m - c; every later stage normalises a, multiplies c by
// its twiddle (< 1.25m, normalised) and forms the same pair. So after stage s
// every value is < (1.25 + 2s) m < 25.25m < 2^259 with limbs < 2^29 + 2^30,
// within mul()'s operand limits; the write multiplies by the inter-pass
// twiddle (or the scale), or reduces fully in the last pass.
// Work-group `b` (of nwg) of one pass of one transform.
__device__ __forceinline__ void ntt_pass_tile(
    const Fr* __restrict__ in, Fr* __restrict__ out, const PassParams& pp, uint32_t b, uint32_t* lds)
{
    const uint32_t log_r = pp.log_r, log_g = pp.log_g, log_l = pp.log_l;
    const uint32_t R = 1u << log_r;
    const uint32_t G = 1u << log_g;
    const uint32_t RG = R << log_g;
    const uint32_t T = blockDim.x;
    const uint32_t nwg = (uint32_t)((1ull << pp.log_n) >> (log_r + log_g));
    if (pp.xcd_remap) // blocks are dealt to the 8 XCDs round robin: keep neighbouring columns on one L2
        b = (b & 7u) * (nwg >> 3) + (b >> 3);
    const uint32_t j0 = b << log_g;
    const uint64_t stride = (1ull << pp.log_n) >> log_r;

    for (uint32_t f = threadIdx.x; f < RG; f += T) {
        const uint32_t g = f & (G - 1);
        const uint32_t n1 = f >> log_g;
        const uint64_t gi = (uint64_t)(j0 + g) + (uint64_t)n1 * stride;
        Fr29 x;
        if (gi < pp.in_len) {
            x = Fr29::load32(in + gi);
            if (pp.twist == TWIST_ZETA) {
                const uint32_t m = (uint32_t)(gi % 3u);
                if (m)
                    x = Fr29::mul(x, zeta261(m));
            } else if (pp.twist == TWIST_POWERS) {
                const uint32_t hm = (1u << pp.h) - 1;
                const uint32_t lo = (uint32_t)gi & hm, hi = (uint32_t)(gi >> pp.h);
                Fr29 w = Fr29::load32(pp.c_lo + lo);
                if (hi)
                    w = Fr29::mul(w, Fr29::load32(pp.c_hi + hi));
                x = Fr29::mul(x, w);
            }
        } else {
            x = Fr29::zero();
        }
        const uint32_t slot = (g << log_r) + (__brev(n1) >> (32 - log_r));
        lds_put(lds, RG, lds_slot(slot), x);
    }

    // Stages go in pairs (radix 4: four elements per thread through LDS once
    // for two stages); an odd stage count starts with one radix-2 stage 0,
    // which has only unit twiddles. Per element the operations are the same
    // as radix-2, so the bounds above hold.
    uint32_t s = 0;
    if (log_r & 1u) {
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < (RG >> 1); q += T) {
            const uint32_t i0 = q << 1;
            const uint32_t p0 = lds_slot(i0), p1 = lds_slot(i0 + 1);
            const Fr29 a = lds_get(lds, RG, p0);
            const Fr29 c = lds_get(lds, RG, p1);
            lds_put(lds, RG, p0, Fr29::add(a, c));
            lds_put(lds, RG, p1, Fr29::sub2(a, c));
        }
        s = 1;
    }
    // A stage pair of half-width h touches, per wave, the 256 slots
    // [256 w', 256 w' + 256) that the same wave wrote in the previous pair whenever
    // its blocks (4h slots) fit in 256 slots or a column has at most 256 slots
    // (q -> slot ranges are a fixed function of q for every such pair): those
    // pairs need only wave-level ordering of the LDS accesses (LDS serves one
    // wave's instructions in order), not a work-group barrier (lds_slot keeps a
    // 256-slot range in the same words). The first pair after the load or the
    // radix-2 stage reads other waves' slots.
    const uint32_t quarterR = R >> 2;
    const uint32_t s_first = s;
    for (; s + 1 < log_r; s += 2) {
        const uint32_t h = 1u << s;
        if (s == s_first || ((4u << s) > 256u && R > 256u)) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        for (uint32_t q = threadIdx.x; q < (RG >> 2); q += T) {
            const uint32_t g = q >> (log_r - 2);
            const uint32_t qq = q & (quarterR - 1);
            const uint32_t pos = qq & (h - 1);
            const uint32_t blk = qq >> s;
            const uint32_t i0 = (g << log_r) + (blk << (s + 2)) + pos;
            const uint32_t p0 = lds_slot(i0), p1 = lds_slot(i0 + h), p2 = lds_slot(i0 + 2 * h),
                           p3 = lds_slot(i0 + 3 * h);
            Fr29 x0 = lds_get(lds, RG, p0);
            Fr29 x1 = lds_get(lds, RG, p1);
            Fr29 x2 = lds_get(lds, RG, p2);
            Fr29 x3 = lds_get(lds, RG, p3);
            if (s) {
                const Fr29 w = Fr29::load32(pp.twr + (pos << (log_r - 1 - s)));
                x0 = Fr29::normalize(x0);
                x2 = Fr29::normalize(x2);
                x1 = Fr29::mul(x1, w);
                x3 = Fr29::mul(x3, w);
            }
            Fr29 y0 = Fr29::add(x0, x1), y1 = Fr29::sub2(x0, x1);
            Fr29 y2 = Fr29::add(x2, x3), y3 = Fr29::sub2(x2, x3);
            y0 = Fr29::normalize(y0);
            y1 = Fr29::normalize(y1);
            y2 = Fr29::mul(y2, Fr29::load32(pp.twr + (pos << (log_r - 2 - s))));
            y3 = Fr29::mul(y3, Fr29::load32(pp.twr + ((pos + h) << (log_r - 2 - s))));
            lds_put(lds, RG, p0, Fr29::add(y0, y2));
            lds_put(lds, RG, p2, Fr29::sub2(y0, y2));
            lds_put(lds, RG, p1, Fr29::add(y1, y3));
            lds_put(lds, RG, p3, Fr29::sub2(y1, y3));
        }
    }
    __syncthreads();

    Fr29 scale;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        scale.l[q] = pp.scale[q];
    const bool mapA = (G > 1) && (log_l > 0);
    const uint32_t lmask = (1u << log_l) - 1;
    for (uint32_t f = threadIdx.x; f < RG; f += T) {
        uint32_t g, k1;
        if (mapA) {
            g = f & (G - 1);
            k1 = f >> log_g;
        } else {
            k1 = f & (R - 1);
            g = f >> log_r;
        }
        Fr29 x = lds_get(lds, RG, lds_slot((g << log_r) + k1));
        const uint32_t j = j0 + g;
        const uint32_t K = j & lmask;
        const uint32_t n2 = j >> log_l;
        const uint64_t o = ((uint64_t)n2 << (log_l + log_r)) + K + ((uint64_t)k1 << log_l);
        if (!pp.last) {
            const uint32_t e = (n2 * k1) << log_l;
            const uint32_t hm = (1u << pp.h) - 1;
            Fr29 w = Fr29::load32(pp.tw_lo + (e & hm)); // entry 0 is one: the multiply also bounds x
            const uint32_t hi = e >> pp.h;
            if (hi)
                w = Fr29::mul(w, Fr29::load32(pp.tw_hi + hi));
            x = Fr29::mul(x, w);
        } else {
            if (pp.post != POST_NONE) {
                x = Fr29::mul(x, scale);
                if (pp.post == POST_SCALE_ZETA_INV) {
                    const uint32_t m = (uint32_t)(o % 3u);
                    if (m)
                        x = Fr29::mul(x, zeta261(3u - m));
                }
            } else {
                x = Fr29::normalize(x);
            }
            x = Fr29::reduce_full(x);
        }
        x.store32(out + o);
    }
}

__global__ void __launch_bounds__(1024) ntt_pass_kernel(PassBatch pb, PassParams pp)
{
    extern __shared__ uint32_t lds[];
    ntt_pass_tile(pb.in[blockIdx.y], pb.out[blockIdx.y], pp, blockIdx.x, lds);
}

// One pass of many transforms of different sizes: work-groups
// [block0, block0 + nwg) of the launch belong to the transform.
struct BatchPass {
    const Fr* in;
    Fr* out;
    uint32_t block0, nwg;
    PassParams pp;
};

__global__ void __launch_bounds__(1024) ntt_pass_batch_kernel(const BatchPass* __restrict__ jobs, uint32_t njobs)
{
    extern __shared__ uint32_t lds[];
    uint32_t lo = 0, hi = njobs; // last job with block0 <= blockIdx.x (uniform)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].block0 <= blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    const BatchPass& jb = jobs[lo];
    const PassParams pp = jb.pp;
    ntt_pass_tile(jb.in, jb.out, pp, blockIdx.x - jb.block0, lds);
}

// The dynamic-LDS limit of the pass kernels, raised once per process (std::call_once:
// concurrent first calls from several host threads are safe).
static hipError_t raise_lds_limits()
{
    static std::once_flag once;
    static hipError_t status = hipSuccess;
    std::call_once(once, [] {
        status = hipFuncSetAttribute((const void*)ntt_pass_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
            36 << NTT_MAX_LOG_R);
        if (status == hipSuccess)
            status = hipFuncSetAttribute((const void*)ntt_pass_batch_kernel,
                hipFuncAttributeMaxDynamicSharedMemorySize, 36 << NTT_MAX_LOG_R);
    });
    return status;
}

RustError ntt_run(const NttTables& t, uint32_t log_n, const NttJob* jobs, int count, uint64_t in_len,
    Fr* scratch, uint64_t scratch_slots, NttTwist twist, NttPost post, const Fr& scale, hipStream_t stream)
{
    if (count <= 0)
        return ok();
    H2_CHECK_ARG(log_n >= 1 && log_n <= NTT_MAX_LOG, "NTT log_n=%u out of range [1, %u]", log_n, NTT_MAX_LOG);
    H2_TRY(raise_lds_limits());
    const NttPlan p = ntt_plan(log_n);
    const uint64_t N = 1ull << log_n;
    bool aliased = false;
    for (int i = 0; i < count; ++i)
        aliased |= (const void*)jobs[i].in == (const void*)jobs[i].out;
    const bool need_scratch = p.P == 3 || (p.P == 2 && aliased);
    H2_CHECK_ARG(!need_scratch || (scratch != nullptr && scratch_slots > 0), "NTT of 2^%u needs a scratch buffer",
        log_n);
    int max_batch = NTT_MAX_BATCH;
    if (need_scratch && scratch_slots < (uint64_t)max_batch)
        max_batch = (int)scratch_slots;
    uint32_t scale261[9] = { 0 };
    if (post != POST_NONE)
        host_to261(scale, scale261);

    for (int base = 0; base < count; base += max_batch) {
        const int nb = count - base < max_batch ? count - base : max_batch;
        for (int pass = 0; pass < p.P; ++pass) {
            PassBatch pb {};
            for (int i = 0; i < nb; ++i) {
                const NttJob& jb = jobs[base + i];
                Fr* s = scratch ? scratch + (uint64_t)i * N : nullptr;
                const Fr* src;
                Fr* dst;
                if (p.P == 1) {
                    src = jb.in;
                    dst = jb.out;
                } else if (p.P == 2) {
                    Fr* mid = need_scratch ? s : jb.out;
                    src = pass == 0 ? jb.in : mid;
                    dst = pass == 0 ? mid : jb.out;
                } else {
                    src = pass == 0 ? jb.in : (pass == 1 ? s : jb.out);
                    dst = pass == 0 ? s : jb.out;
                }
                pb.in[i] = src;
                pb.out[i] = dst;
            }
            PassParams pp {};
            pp.log_n = log_n;
            pp.log_r = p.log_r[pass];
            pp.log_g = p.log_g[pass];
            uint32_t log_l = 0;
            for (int q = 0; q < pass; ++q)
                log_l += p.log_r[q];
            pp.log_l = log_l;
            pp.in_len = pass == 0 ? in_len : N;
            pp.tw_lo = t.tw_lo;
            pp.tw_hi = t.tw_hi;
            pp.h = t.h;
            pp.twr = t.twr[pass];
            pp.c_lo = t.c_lo;
            pp.c_hi = t.c_hi;
            pp.twist = pass == 0 ? (int)twist : TWIST_NONE;
            pp.last = pass == p.P - 1;
            pp.post = pp.last ? (int)post : POST_NONE;
            for (int q = 0; q < 9; ++q)
                pp.scale[q] = scale261[q];
            const uint32_t RG = 1u << (pp.log_r + pp.log_g);
            const uint32_t nwg = (uint32_t)(N >> (pp.log_r + pp.log_g));
            pp.xcd_remap = (nwg % 8 == 0 && nwg >= 64) ? 1u : 0u;
            uint32_t T = RG / 2;
            if (T > 1024)
                T = 1024;
            if (T < 64)
                T = 64;
            dim3 grid(nwg, nb);
            {
                ProfScope prof("ntt_pass", stream);
                ntt_pass_kernel<<<grid, T, 36u * RG, stream>>>(pb, pp);
            }
            H2_LAUNCH_CHECK();
        }
    }
    return ok();
}

// ---------------------------------------------------------- uneven batches

uint64_t ntt_batch_pass_bytes(int count) { return 3ull * (uint64_t)count * sizeof(BatchPass); }

RustError ntt_run_batch(const NttBatchJob* jobs, int count, NttTwist twist, NttPost post, void* d_pass,
    hipStream_t stream)
{
    if (count <= 0)
        return ok();
    H2_TRY(raise_lds_limits());
    std::vector<BatchPass> host(3ull * count);
    uint32_t nlaunch[3] = { 0, 0, 0 }, grid[3] = { 0, 0, 0 }, rgmax[3] = { 0, 0, 0 };
    for (int i = 0; i < count; ++i) {
        const NttBatchJob& jb = jobs[i];
        H2_CHECK_ARG(jb.log_n >= 1 && jb.log_n <= NTT_MAX_LOG, "NTT batch: log_n=%u out of range [1, %u]", jb.log_n,
            NTT_MAX_LOG);
        NttPlan p = ntt_plan(jb.log_n);
        if (p.P > 1) // whole 4096-element tiles: a batch fills the chip with many transforms, not wide ones
            for (int q = 0; q < p.P; ++q) {
                const uint32_t g = NTT_MAX_LOG_R - p.log_r[q];
                p.log_g[q] = g < jb.log_n - p.log_r[q] ? g : jb.log_n - p.log_r[q];
            }
        H2_CHECK_ARG(p.P < 3 || jb.scratch, "NTT batch: 2^%u needs a scratch buffer", jb.log_n);
        H2_CHECK_ARG(p.P != 2 || (const void*)jb.in != (const void*)jb.out, "NTT batch: in-place 2^%u transform",
            jb.log_n);
        uint32_t scale261[9] = { 0 };
        if (post != POST_NONE)
            host_to261(jb.scale, scale261);
        for (int pass = 0; pass < p.P; ++pass) {
            BatchPass bp {};
            if (p.P == 1) {
                bp.in = jb.in;
                bp.out = jb.out;
            } else if (p.P == 2) {
                bp.in = pass == 0 ? jb.in : jb.out;
                bp.out = jb.out;
            } else {
                bp.in = pass == 0 ? jb.in : (pass == 1 ? jb.scratch : jb.out);
                bp.out = pass == 0 ? jb.scratch : jb.out;
            }
            PassParams& pp = bp.pp;
            pp.log_n = jb.log_n;
            pp.log_r = p.log_r[pass];
            pp.log_g = p.log_g[pass];
            uint32_t log_l = 0;
            for (int q = 0; q < pass; ++q)
                log_l += p.log_r[q];
            pp.log_l = log_l;
            pp.in_len = pass == 0 ? jb.in_len : (1ull << jb.log_n);
            pp.tw_lo = jb.tables->tw_lo;
            pp.tw_hi = jb.tables->tw_hi;
            pp.h = jb.tables->h;
            pp.twr = jb.tables->twr[pass];
            pp.c_lo = jb.tables->c_lo;
            pp.c_hi = jb.tables->c_hi;
            pp.twist = pass == 0 ? (int)twist : TWIST_NONE;
            pp.last = pass == p.P - 1;
            pp.post = pp.last ? (int)post : POST_NONE;
            for (int q = 0; q < 9; ++q)
                pp.scale[q] = scale261[q];
            pp.xcd_remap = 0;
            const uint32_t rg = 1u << (pp.log_r + pp.log_g);
            bp.nwg = (uint32_t)((1ull << jb.log_n) >> (pp.log_r + pp.log_g));
            bp.block0 = grid[pass];
            grid[pass] += bp.nwg;
            if (rg > rgmax[pass])
                rgmax[pass] = rg;
            host[(uint64_t)pass * count + nlaunch[pass]++] = bp;
        }
    }
    // pinned staging: the host table may go away on return without draining the stream
    {
        RustError e = stage_upload(d_pass, host.data(), sizeof(BatchPass) * host.size(), nullptr, nullptr, 0, stream);
        if (e.code)
            return e;
    }
    const BatchPass* dp = (const BatchPass*)d_pass;
    for (int pass = 0; pass < 3; ++pass) {
        if (!nlaunch[pass])
            continue;
        uint32_t T = rgmax[pass] / 2;
        if (T > 1024)
            T = 1024;
        if (T < 64)
            T = 64;
        ProfScope prof("ntt_pass_batch", stream);
        ntt_pass_batch_kernel<<<grid[pass], T, 36u * rgmax[pass], stream>>>(dp + (uint64_t)pass * count,
            nlaunch[pass]);
        H2_LAUNCH_CHECK();
    }
    return ok();
}

RustError ntt_set_tiles(const uint8_t* tiles, uint32_t count)
{
    H2_CHECK_ARG(count <= NTT_MAX_LOG + 1, "h2amd_ntt_set_tiles: %u entries exceed %u", count, NTT_MAX_LOG + 1);
    for (uint32_t i = 0; i < count; ++i)
        H2_CHECK_ARG(tiles[i] == 0 || (tiles[i] >= 6 && tiles[i] <= NTT_MAX_LOG_R),
            "h2amd_ntt_set_tiles: tile 2^%u for 2^%u outside [2^6, 2^%u]", tiles[i], i, NTT_MAX_LOG_R);
    for (uint32_t i = 0; i <= NTT_MAX_LOG; ++i)
        g_tile_table[i].store(i < count ? tiles[i] : 0, std::memory_order_relaxed);
    return ok();
}

} // namespace h2
