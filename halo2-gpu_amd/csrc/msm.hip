// BN254 G1 multi-scalar multiplication (Pippenger) for gfx950, behind the
// reference's `_halo2_multiexp*` entry points (cuda/src/multiexp.cu:166-525).
//
// Pipeline (all on one stream, caller-owned workspace):
//  1. scalars: Montgomery -> canonical (one Fr multiplication by 1).
//  2. signed c-bit digits (carry-propagated, |d| <= 2^(c-1)), W = ceil(255/c)
//     windows, B = 2^(c-1) buckets per window; a zero digit adds nothing.
//  3. counting sort of (point index, sign) by (window, bucket): histogram with
//     device atomics, exclusive scan, scatter.
//  4. bucket sums with load balancing: every bucket is cut into pieces of at
//     most S0 entries, one thread per piece accumulates XYZZ mixed additions;
//     the partial sums are then merged level by level (pieces of S1 partials)
//     until every bucket has one value. A dense bucket (all scalars equal) is
//     therefore spread over many threads instead of one.
//  5. per window, sum_b b * B_b: each thread takes a run of L buckets, computes
//     the running-sum pair (sum j*B, sum B) and adds lo * (sum B); segment
//     results are summed in groups of 16.
//  6. Horner over windows (c doublings per window), conversion to affine.
// The result is written as (x, y, 1) or the identity (0, 1, 0) in Montgomery
// form, the layout of the reference's `cukernel_to_affine` output.
#include <memory>

#include "common.h"
#include "curve.h"
#include "profile.h"
#include "scan_u32.h"

using namespace h2;

namespace {

constexpr uint32_t SCALAR_BITS = 254;
constexpr uint32_t MSM_S1 = 16; // partials merged per thread at levels >= 1
constexpr uint32_t RED_GROUP = 16; // segment results summed per thread

struct MsmPlan {
    uint64_t n;
    uint32_t c, W, B;
    uint64_t nb; // W * B buckets
    uint32_t S0;
    uint32_t levels;
    uint64_t ub[16]; // upper bound on pieces per level
    uint32_t L; // buckets per reduction segment
    uint64_t M; // segments per window
};

uint32_t log2_floor(uint64_t n)
{
    uint32_t l = 0;
    while ((2ull << l) <= n)
        ++l;
    return l;
}

MsmPlan make_plan(uint64_t n)
{
    MsmPlan p {};
    p.n = n;
    int c = (int)log2_floor(n < 1 ? 1 : n) - 4;
    if (c < 4)
        c = 4;
    if (c > 20)
        c = 20;
    p.c = (uint32_t)c;
    p.W = (255 + p.c - 1) / p.c;
    p.B = 1u << (p.c - 1);
    p.nb = (uint64_t)p.W * p.B;
    uint64_t work = n * p.W;
    uint64_t s0 = (work + (1ull << 21) - 1) >> 21;
    if (s0 < 4)
        s0 = 4;
    p.S0 = (uint32_t)s0;
    // levels: level 0 leaves <= ceil(n / S0) pieces in the worst bucket, each later level divides by S1
    uint64_t worst = (n + p.S0 - 1) / p.S0;
    p.levels = 1;
    while (worst > 1) {
        worst = (worst + MSM_S1 - 1) / MSM_S1;
        ++p.levels;
    }
    p.ub[0] = (work + p.S0 - 1) / p.S0 + p.nb;
    for (uint32_t l = 1; l < p.levels; ++l)
        p.ub[l] = (p.ub[l - 1] + MSM_S1 - 1) / MSM_S1 + p.nb;
    p.L = p.B < 8 ? p.B : 8;
    p.M = p.B / p.L;
    return p;
}

struct MsmBuffers {
    uint32_t* canon; // n x 8 limbs
    uint32_t* counts; // nb
    uint32_t* off0; // nb + 1
    uint32_t* cursor; // nb
    uint32_t* entries; // n * W
    uint32_t* pc; // nb
    uint32_t* poff[2]; // nb + 1 each
    G1Xyzz* part[2]; // ub[0], max(ub[1..])
    G1Xyzz* bsum; // nb
    G1Xyzz* red[2]; // W * M, W * ceil(M / 16)
    void* scan_scratch;
    uint64_t* out; // 12 limbs
};

template <class WS>
void carve(WS& ws, const MsmPlan& p, MsmBuffers* b)
{
    auto take = [&](uint64_t bytes) -> void* { return ws.take(bytes); };
    uint64_t ub1 = 1;
    for (uint32_t l = 1; l < p.levels; ++l)
        ub1 = p.ub[l] > ub1 ? p.ub[l] : ub1;
    void* v[16];
    v[0] = take(32 * p.n);
    v[1] = take(4 * p.nb);
    v[2] = take(4 * (p.nb + 1));
    v[3] = take(4 * p.nb);
    v[4] = take(4 * p.n * p.W);
    v[5] = take(4 * p.nb);
    v[6] = take(4 * (p.nb + 1));
    v[7] = take(4 * (p.nb + 1));
    v[8] = take(128 * p.ub[0]);
    v[9] = take(128 * ub1);
    v[10] = take(128 * p.nb);
    v[11] = take(128 * p.W * p.M);
    v[12] = take(128 * p.W * ((p.M + RED_GROUP - 1) / RED_GROUP));
    v[13] = take(scan_u32_scratch(p.nb + 1));
    v[14] = take(96);
    if (b) {
        b->canon = (uint32_t*)v[0];
        b->counts = (uint32_t*)v[1];
        b->off0 = (uint32_t*)v[2];
        b->cursor = (uint32_t*)v[3];
        b->entries = (uint32_t*)v[4];
        b->pc = (uint32_t*)v[5];
        b->poff[0] = (uint32_t*)v[6];
        b->poff[1] = (uint32_t*)v[7];
        b->part[0] = (G1Xyzz*)v[8];
        b->part[1] = (G1Xyzz*)v[9];
        b->bsum = (G1Xyzz*)v[10];
        b->red[0] = (G1Xyzz*)v[11];
        b->red[1] = (G1Xyzz*)v[12];
        b->scan_scratch = v[13];
        b->out = (uint64_t*)v[14];
    }
}

struct SizerAdapter {
    WorkspaceSizer s;
    void* take(uint64_t n)
    {
        s.take(n);
        return nullptr;
    }
};

uint64_t msm_workspace_bytes(uint64_t n)
{
    MsmPlan p = make_plan(n);
    SizerAdapter sa;
    carve(sa, p, nullptr);
    return sa.s.bytes + 4096;
}

// ------------------------------------------------------------- kernels

__global__ void to_canonical_kernel(const Fr* __restrict__ scal, uint32_t* __restrict__ canon, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Fr v = Fr::from_mont(Fr::load(scal + i));
    v.store(canon + 8 * i);
}

// signed digit of window w; `carry` flows from window w-1 to w
__device__ __forceinline__ int32_t next_digit(const uint32_t* s, uint32_t w, uint32_t c, uint32_t& carry)
{
    const uint32_t bit = w * c;
    uint32_t raw = 0;
    if (bit < SCALAR_BITS) {
        const uint32_t limb = bit >> 5, sh = bit & 31;
        uint64_t two = s[limb];
        if (limb + 1 < 8)
            two |= (uint64_t)s[limb + 1] << 32;
        raw = (uint32_t)(two >> sh) & ((1u << c) - 1u);
    }
    raw += carry;
    const uint32_t half = 1u << (c - 1);
    if (raw > half) {
        carry = 1;
        return (int32_t)raw - (int32_t)(1u << c);
    }
    carry = 0;
    return (int32_t)raw;
}

__global__ void histogram_kernel(const uint32_t* __restrict__ canon, uint64_t n, uint32_t c, uint32_t W, uint32_t B,
    uint32_t* __restrict__ counts)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        s[k] = canon[8 * i + k];
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const int32_t d = next_digit(s, w, c, carry);
        if (d != 0) {
            const uint32_t mag = (uint32_t)(d < 0 ? -d : d);
            atomicAdd(&counts[(uint64_t)w * B + mag - 1], 1u);
        }
    }
}

__global__ void scatter_kernel(const uint32_t* __restrict__ canon, uint64_t n, uint32_t c, uint32_t W, uint32_t B,
    uint32_t* __restrict__ cursor, uint32_t* __restrict__ entries)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t s[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        s[k] = canon[8 * i + k];
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; ++w) {
        const int32_t d = next_digit(s, w, c, carry);
        if (d != 0) {
            const uint32_t mag = (uint32_t)(d < 0 ? -d : d);
            const uint32_t pos = atomicAdd(&cursor[(uint64_t)w * B + mag - 1], 1u);
            entries[pos] = (uint32_t)i | (d < 0 ? 0x80000000u : 0u);
        }
    }
}

__global__ void piece_count_kernel(const uint32_t* __restrict__ off, uint64_t nb, uint32_t S, uint32_t* __restrict__ pc)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb)
        return;
    const uint32_t cnt = off[b + 1] - off[b];
    pc[b] = (cnt + S - 1) / S;
}

// largest b with poff[b] <= p (poff nondecreasing, nb + 1 entries, poff[0] = 0)
__device__ __forceinline__ uint64_t find_bucket(const uint32_t* __restrict__ poff, uint64_t nb, uint32_t p)
{
    uint64_t lo = 0, hi = nb; // invariant: poff[lo] <= p < poff[hi]
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (poff[mid] <= p)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// level 0: one thread per piece of <= S entries of one bucket; affine inputs
__global__ void __launch_bounds__(256) sum_level0_kernel(const uint32_t* __restrict__ entries, const uint8_t* __restrict__ points,
    const uint32_t* __restrict__ off, const uint32_t* __restrict__ poff, uint64_t nb, uint32_t S,
    G1Xyzz* __restrict__ out)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= poff[nb])
        return;
    const uint64_t b = find_bucket(poff, nb, (uint32_t)p);
    const uint64_t start = off[b] + (p - poff[b]) * (uint64_t)S;
    uint64_t end = start + S;
    if (end > off[b + 1])
        end = off[b + 1];
    G1Xyzz acc = G1Xyzz::inf();
    for (uint64_t e = start; e < end; ++e) {
        const uint32_t ent = entries[e];
        G1Affine q = G1Affine::load(points + 64ull * (ent & 0x7fffffffu));
        if (ent >> 31)
            q = q.neg();
        acc = G1Xyzz::add_affine(acc, q);
    }
    acc.store(out + p);
}

// level >= 1: one thread per piece of <= S partials of one bucket
__global__ void __launch_bounds__(256) sum_level_kernel(const G1Xyzz* __restrict__ in, const uint32_t* __restrict__ off,
    const uint32_t* __restrict__ poff, uint64_t nb, uint32_t S, G1Xyzz* __restrict__ out)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= poff[nb])
        return;
    const uint64_t b = find_bucket(poff, nb, (uint32_t)p);
    const uint64_t start = off[b] + (p - poff[b]) * (uint64_t)S;
    uint64_t end = start + S;
    if (end > off[b + 1])
        end = off[b + 1];
    G1Xyzz acc = G1Xyzz::inf();
    for (uint64_t e = start; e < end; ++e)
        acc = G1Xyzz::add(acc, G1Xyzz::load(in + e));
    acc.store(out + p);
}

__global__ void gather_buckets_kernel(const G1Xyzz* __restrict__ part, const uint32_t* __restrict__ poff, uint64_t nb,
    G1Xyzz* __restrict__ bsum)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb)
        return;
    if (poff[b + 1] > poff[b])
        G1Xyzz::load(part + poff[b]).store(bsum + b);
    else
        G1Xyzz::inf().store(bsum + b);
}

// segment m of window w covers buckets (1-based) lo+1 .. lo+L, lo = m*L:
// result = sum_j j*B_{lo+j} + lo * sum_j B_{lo+j}
__global__ void __launch_bounds__(256) bucket_segment_kernel(const G1Xyzz* __restrict__ bsum, uint32_t W, uint32_t B, uint32_t L,
    uint64_t M, G1Xyzz* __restrict__ out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)W * M)
        return;
    const uint64_t w = t / M, m = t % M;
    const G1Xyzz* bw = bsum + w * B;
    const uint32_t lo = (uint32_t)(m * L);
    G1Xyzz run = G1Xyzz::inf(), acc = G1Xyzz::inf();
    for (int j = (int)L - 1; j >= 0; --j) {
        run = G1Xyzz::add(run, G1Xyzz::load(bw + lo + j));
        acc = G1Xyzz::add(acc, run);
    }
    if (lo)
        acc = G1Xyzz::add(acc, G1Xyzz::mul_small(run, lo));
    acc.store(out + t);
}

// out[w * Mo + g] = sum of in[w * Mi + g*G .. + G)
__global__ void __launch_bounds__(256) group_sum_kernel(const G1Xyzz* __restrict__ in, uint64_t Mi, uint32_t W, uint32_t G,
    G1Xyzz* __restrict__ out, uint64_t Mo)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)W * Mo)
        return;
    const uint64_t w = t / Mo, g = t % Mo;
    G1Xyzz acc = G1Xyzz::inf();
    for (uint64_t k = g * G; k < (g + 1) * G && k < Mi; ++k)
        acc = G1Xyzz::add(acc, G1Xyzz::load(in + w * Mi + k));
    acc.store(out + t);
}

__global__ void __launch_bounds__(64) finalize_kernel(const G1Xyzz* __restrict__ win, uint32_t W, uint32_t c, uint64_t* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    G1Xyzz acc = G1Xyzz::load(win + (W - 1));
    for (int w = (int)W - 2; w >= 0; --w) {
        for (uint32_t k = 0; k < c; ++k)
            acc = G1Xyzz::dbl(acc);
        acc = G1Xyzz::add(acc, G1Xyzz::load(win + w));
    }
    Fq one = Fq::one();
    if (acc.is_inf()) {
        Fq::zero().store(out);
        one.store(out + 4);
        Fq::zero().store(out + 8);
    } else {
        G1Affine a = acc.to_affine();
        a.x.store(out);
        a.y.store(out + 4);
        one.store(out + 8);
    }
}

unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Full device pipeline; the 96-byte result lands in b.out (device).
RustError msm_device(const MsmPlan& p, const MsmBuffers& b, const Fr* d_scalars, const uint8_t* d_points,
    hipStream_t stream)
{
    const unsigned T = 256;
    ProfScope prof_all("msm_total", stream);
    std::unique_ptr<ProfScope> prof_sort(new ProfScope("msm_sort", stream));
    to_canonical_kernel<<<blocks_for(p.n, T), T, 0, stream>>>(d_scalars, b.canon, p.n);
    H2_LAUNCH_CHECK();
    H2_TRY(hipMemsetAsync(b.counts, 0, 4 * p.nb, stream));
    histogram_kernel<<<blocks_for(p.n, T), T, 0, stream>>>(b.canon, p.n, p.c, p.W, p.B, b.counts);
    H2_LAUNCH_CHECK();
    RustError e = scan_u32_exclusive(b.counts, b.off0, p.nb, b.scan_scratch, stream);
    if (e.code)
        return e;
    H2_TRY(hipMemcpyAsync(b.cursor, b.off0, 4 * p.nb, hipMemcpyDeviceToDevice, stream));
    scatter_kernel<<<blocks_for(p.n, T), T, 0, stream>>>(b.canon, p.n, p.c, p.W, p.B, b.cursor, b.entries);
    H2_LAUNCH_CHECK();
    prof_sort.reset();

    // level 0
    piece_count_kernel<<<blocks_for(p.nb, T), T, 0, stream>>>(b.off0, p.nb, p.S0, b.pc);
    H2_LAUNCH_CHECK();
    e = scan_u32_exclusive(b.pc, b.poff[0], p.nb, b.scan_scratch, stream);
    if (e.code)
        return e;
    {
        ProfScope prof("msm_bucket_level0", stream);
        sum_level0_kernel<<<blocks_for(p.ub[0], T), T, 0, stream>>>(b.entries, d_points, b.off0, b.poff[0], p.nb,
            p.S0, b.part[0]);
    }
    H2_LAUNCH_CHECK();
    int cur = 0;
    for (uint32_t l = 1; l < p.levels; ++l) {
        piece_count_kernel<<<blocks_for(p.nb, T), T, 0, stream>>>(b.poff[cur], p.nb, MSM_S1, b.pc);
        H2_LAUNCH_CHECK();
        e = scan_u32_exclusive(b.pc, b.poff[cur ^ 1], p.nb, b.scan_scratch, stream);
        if (e.code)
            return e;
        sum_level_kernel<<<blocks_for(p.ub[l], T), T, 0, stream>>>(b.part[cur], b.poff[cur], b.poff[cur ^ 1], p.nb,
            MSM_S1, b.part[cur ^ 1]);
        H2_LAUNCH_CHECK();
        cur ^= 1;
    }
    gather_buckets_kernel<<<blocks_for(p.nb, T), T, 0, stream>>>(b.part[cur], b.poff[cur], p.nb, b.bsum);
    H2_LAUNCH_CHECK();

    // per-window weighted bucket sums
    ProfScope prof_red("msm_reduce", stream);
    bucket_segment_kernel<<<blocks_for((uint64_t)p.W * p.M, T), T, 0, stream>>>(b.bsum, p.W, p.B, p.L, p.M, b.red[0]);
    H2_LAUNCH_CHECK();
    uint64_t Mi = p.M;
    int r = 0;
    while (Mi > 1) {
        const uint64_t Mo = (Mi + RED_GROUP - 1) / RED_GROUP;
        group_sum_kernel<<<blocks_for((uint64_t)p.W * Mo, T), T, 0, stream>>>(b.red[r], Mi, p.W, RED_GROUP,
            b.red[r ^ 1], Mo);
        H2_LAUNCH_CHECK();
        r ^= 1;
        Mi = Mo;
    }
    finalize_kernel<<<1, 64, 0, stream>>>(b.red[r], p.W, p.c, b.out);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError msm_entry(const char* who, const void* scalars, bool scalars_on_device, const void* points,
    bool points_on_device, void* h_out, uint64_t n, void* scratch, uint64_t scratch_bytes, hipStream_t stream)
{
    if (n == 0) {
        // identity (0, 1, 0)
        uint64_t* o = (uint64_t*)h_out;
        memset(o, 0, 96);
        Fq one = Fq::one();
        memcpy(o + 4, one.v, 32);
        return ok();
    }
    H2_CHECK_ARG(n < (1ull << 31), "%s: length %lu exceeds 2^31", who, (unsigned long)n);
    const MsmPlan p = make_plan(n);
    H2_CHECK_ARG((uint64_t)p.n * p.W < (1ull << 32), "%s: length %lu too large for 32-bit bucket offsets", who,
        (unsigned long)n);
    Workspace ws(scratch, scratch_bytes);
    const Fr* d_scal = (const Fr*)scalars;
    const uint8_t* d_pts = (const uint8_t*)points;
    if (!scalars_on_device)
        d_scal = (const Fr*)ws.take(32 * n);
    if (!points_on_device)
        d_pts = (const uint8_t*)ws.take(64 * n);
    MsmBuffers b {};
    carve(ws, p, &b);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "%s: workspace of %lu bytes is too small for length %lu", who,
            (unsigned long)scratch_bytes, (unsigned long)n);
    if (!scalars_on_device)
        H2_TRY(hipMemcpyAsync((void*)d_scal, scalars, 32 * n, hipMemcpyHostToDevice, stream));
    if (!points_on_device)
        H2_TRY(hipMemcpyAsync((void*)d_pts, points, 64 * n, hipMemcpyHostToDevice, stream));
    RustError e = msm_device(p, b, d_scal, d_pts, stream);
    if (e.code)
        return e;
    H2_TRY(hipMemcpyAsync(h_out, b.out, 96, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

} // namespace

extern "C" {

size_t _halo2_msm_max_length(uint64_t free_bytes)
{
    for (uint32_t log_n = 28; log_n > 0; --log_n) {
        uint64_t n = 1ull << log_n;
        if (msm_workspace_bytes(n) + 96 * n < free_bytes)
            return (size_t)n;
    }
    return 0;
}

uint64_t _halo2_multiexp_workspace_size(uint64_t length) { return msm_workspace_bytes(length) + 96 * length + 512; }

uint64_t _halo2_multiexp_device_bases_workspace_size(uint64_t length)
{
    return msm_workspace_bytes(length) + 32 * length + 256;
}

uint64_t _halo2_multiexp_device_scalars_device_bases_workspace_size(uint64_t length)
{
    return msm_workspace_bytes(length);
}

RustError _halo2_multiexp(const FFITraitObject* scalar, const FFITraitObject* point, const FFITraitObject* output,
    size_t length, void* scratch, uint64_t scratch_bytes, void* stream)
{
    return msm_entry("_halo2_multiexp", scalar->ptr, false, point->ptr, false, output->ptr, length, scratch,
        scratch_bytes, (hipStream_t)stream);
}

RustError _halo2_multiexp_device_bases(const FFITraitObject* h_scalar, const void* d_bases,
    const FFITraitObject* h_output, size_t length, void* scratch, uint64_t scratch_bytes, void* stream)
{
    return msm_entry("_halo2_multiexp_device_bases", h_scalar->ptr, false, d_bases, true, h_output->ptr, length,
        scratch, scratch_bytes, (hipStream_t)stream);
}

RustError _halo2_multiexp_device_scalars_device_bases(const void* d_scalar, const void* d_bases,
    const FFITraitObject* h_output, size_t length, void* scratch, uint64_t scratch_bytes, void* stream)
{
    return msm_entry("_halo2_multiexp_device_scalars_device_bases", d_scalar, true, d_bases, true, h_output->ptr,
        length, scratch, scratch_bytes, (hipStream_t)stream);
}

} // extern "C"
