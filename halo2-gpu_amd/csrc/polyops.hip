// Polynomial helpers of the opening phase (SHPLONK) and of key generation,
// behind the reference's entry points in cuda/src/polynomial/polynomial.cu,
// kate_division.cu, decode_assigned.cu, omega.cu and cuda/src/multiopen.cu:
//   evaluation at a point (single, batched), the multiopen fold, the FMA /
//   elementwise / fill / short-subtract helpers, division by (X - u),
//   Assigned<F> decoding, omega power tables and omega * delta mappings.
//
// Evaluation: each work-group folds a span of coefficients (Horner in
// u^256 per thread, coalesced reads), scales by the span's power of u and
// sums; one small kernel adds the group sums. Division by (X - u):
// q_j = a_{j+1} + u q_{j+1}, a linear recurrence solved as a suffix scan: 8
// outputs per thread, a work-group scan over 2048-output tiles, one work-group
// scan over the tiles, then every tile again with its carry.
#include "common.h"
#include "field.h"
#include "field29.h"

using namespace h2;

namespace {

constexpr unsigned kT = 256;

unsigned blocks_of(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

struct Sizer {
    WorkspaceSizer s;
    void* take(uint64_t n)
    {
        s.take(n);
        return nullptr;
    }
};

// ------------------------------------------------------------ evaluation

// A work-group covers kEvalSpan = kEvalT * kEvalRows consecutive coefficients
// from base = blockIdx.x * kEvalSpan: thread t folds a_{base + t + m kEvalT}
// over m by Horner in U = u^kEvalT (coalesced: for each m the group reads
// kEvalT consecutive words) and scales by u^t; the group sum is
// P_b = sum_i a_{base + i} u^i, without the factor u^base. The combine kernel
// then adds sum_b P_b V^b, V = u^kEvalSpan, by the same scheme over blocks.
// Powers u^0..u^kEvalT are built in LDS by doubling (8 dependent products),
// so a thread's chain of dependent products is about 8 + kEvalRows + 1 per
// kernel: evaluation is latency-bound at these sizes, not bandwidth-bound.
constexpr unsigned kEvalT = 256;
constexpr unsigned kEvalRows = 16;
constexpr unsigned kEvalLogRows = 4;
constexpr uint64_t kEvalSpan = (uint64_t)kEvalT * kEvalRows;
constexpr unsigned kEvalMaxJobs = 32; // polynomials per launch (the table is a kernel argument)

__host__ __device__ inline uint64_t eval_blocks(uint64_t n) { return (n + kEvalSpan - 1) / kEvalSpan; }

// One polynomial: coefficients a[0, n), the point (from du when set, else u),
// the result slot.
struct EvalJob {
    const Fr* a;
    uint64_t n;
    const Fr* du;
    Fr u;
    Fr* out;
};

struct EvalJobs {
    EvalJob j[kEvalMaxJobs];
};

// bytes: point, per-group sums, result
uint64_t eval_scratch(uint64_t n) { return 3 * 256 + align_up(32 * (eval_blocks(n) + 1), 256); }

struct EvalBufs {
    Fr* point;
    Fr* partial;
    Fr* result;
};

template <class WS>
EvalBufs eval_carve(WS& ws, uint64_t n)
{
    EvalBufs b;
    b.point = (Fr*)ws.take(32);
    b.partial = (Fr*)ws.take(32 * (eval_blocks(n) + 1));
    b.result = (Fr*)ws.take(32);
    return b;
}

__device__ __forceinline__ Fr block_sum(Fr v, Fr* sh)
{
    sh[threadIdx.x] = v;
    for (unsigned d = blockDim.x / 2; d > 0; d >>= 1) {
        __syncthreads();
        if (threadIdx.x < d)
            sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + d];
    }
    __syncthreads();
    return sh[0];
}

// pw[t] = x^t for t <= kEvalT (whole work-group): with pw[0..h] known, threads
// i < h set pw[h + 1 + i] = pw[1 + i] * pw[h].
__device__ __forceinline__ void lds_powers(const Fr& x, Fr* pw)
{
    if (threadIdx.x == 0) {
        pw[0] = Fr::one();
        pw[1] = x;
    }
    __syncthreads();
    for (unsigned h = 1; h < kEvalT; h <<= 1) {
        if (threadIdx.x < h)
            pw[h + 1 + threadIdx.x] = pw[1 + threadIdx.x] * pw[h];
        __syncthreads();
    }
}

__device__ __forceinline__ Fr job_point(const EvalJob& jb) { return jb.du ? Fr::load(jb.du) : jb.u; }

// grid (max blocks, jobs): partial[y * pstride + x] = P_x of job y
__global__ void __launch_bounds__(kEvalT) eval_span_kernel(const EvalJobs jobs, uint64_t pstride,
    Fr* __restrict__ partial)
{
    __shared__ Fr pw[kEvalT + 1];
    __shared__ Fr sh[kEvalT];
    const EvalJob& jb = jobs.j[blockIdx.y];
    const uint64_t base = (uint64_t)blockIdx.x * kEvalSpan;
    if (base >= jb.n)
        return; // uniform over the group
    lds_powers(job_point(jb), pw);
    const Fr big_u = pw[kEvalT];
    const uint64_t rem = jb.n - base;
    const unsigned rows = rem >= kEvalSpan ? kEvalRows : (unsigned)((rem + kEvalT - 1) / kEvalT);
    const Fr* a = jb.a + base;
    Fr acc = Fr::zero();
    for (int m = (int)rows - 1; m >= 0; --m) {
        const uint64_t k = (uint64_t)m * kEvalT + threadIdx.x;
        acc = acc * big_u;
        if (k < rem)
            acc = acc + Fr::load(a + k);
    }
    acc = acc * pw[threadIdx.x];
    const Fr s = block_sum(acc, sh);
    if (threadIdx.x == 0)
        s.store(partial + blockIdx.y * pstride + blockIdx.x);
}

// one work-group per job: out = sum_b P_b V^b with V = u^kEvalSpan; thread t
// takes blocks [t q, t q + q) by Horner in V and scales by (V^q)^t
__global__ void __launch_bounds__(kEvalT) eval_combine_kernel(const EvalJobs jobs, uint64_t pstride,
    const Fr* __restrict__ partial)
{
    __shared__ Fr pw[kEvalT + 1];
    __shared__ Fr sh[kEvalT];
    const EvalJob& jb = jobs.j[blockIdx.x];
    const uint64_t nb = eval_blocks(jb.n);
    lds_powers(job_point(jb), pw);
    Fr v = pw[kEvalT];
#pragma unroll
    for (unsigned i = 0; i < kEvalLogRows; ++i)
        v = Fr::sqr(v);
    const uint64_t q = (nb + kEvalT - 1) / kEvalT;
    const Fr w = Fr::pow_u64(v, q ? q : 1);
    __syncthreads(); // every thread has read pw[kEvalT]
    lds_powers(w, pw);
    const Fr* p = partial + blockIdx.x * pstride;
    Fr acc = Fr::zero();
    const uint64_t b0 = threadIdx.x * q;
    for (uint64_t r = q; r-- > 0;) {
        acc = acc * v;
        if (b0 + r < nb)
            acc = acc + Fr::load(p + b0 + r);
    }
    acc = acc * pw[threadIdx.x];
    const Fr s = block_sum(acc, sh);
    if (threadIdx.x == 0)
        s.store(jb.out);
}

// jobs[0, count) evaluated into their out slots (count <= kEvalMaxJobs); partial
// holds count * pstride elements, pstride >= eval_blocks of every job
RustError eval_jobs(const EvalJobs& jobs, uint32_t count, uint64_t pstride, Fr* partial, hipStream_t stream)
{
    if (count == 0)
        return ok();
    uint64_t maxb = 0;
    for (uint32_t i = 0; i < count; ++i)
        if (eval_blocks(jobs.j[i].n) > maxb)
            maxb = eval_blocks(jobs.j[i].n);
    if (maxb) {
        eval_span_kernel<<<dim3((unsigned)maxb, count), kEvalT, 0, stream>>>(jobs, pstride, partial);
        H2_LAUNCH_CHECK();
    }
    eval_combine_kernel<<<count, kEvalT, 0, stream>>>(jobs, pstride, partial);
    H2_LAUNCH_CHECK();
    return ok();
}

// p(u) into *out (device); u already on the device
RustError eval_device(const Fr* a, uint64_t n, const Fr* d_u, const EvalBufs& b, Fr* out, hipStream_t stream)
{
    EvalJobs jobs;
    jobs.j[0] = EvalJob { a, n, d_u, Fr::zero(), out };
    return eval_jobs(jobs, 1, eval_blocks(n), b.partial, stream);
}

// ---------------------------------------------------------- elementwise

__global__ void multiply_add_kernel(Fr* acc, const Fr* in, const Fr* s,
    uint64_t n)
{
    const Fr k = Fr::load(s);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        (Fr::load(acc + i) + Fr::load(in + i) * k).store(acc + i);
}

// acc[i] += sum_j s_j p_j[i] for up to kFoldMax polynomials in one pass (p_j zero at and
// beyond len[j]): each element of acc is read and written once instead of once per
// polynomial. On the 29-bit form: the scalars go to the 2^261 scale once (shl5 of their
// stored words, held in scalar registers) and two products share one reduction
// (mul_sum: stored words < m and scalars < 32m, so each pair adds < 1.4m); the sum of
// acc and at most kFoldMax / 2 pairs stays below 13m < 2^261 with limbs normalised
// after every pair.
constexpr unsigned kFoldMax = 16;
struct FoldJobs {
    const Fr* p[kFoldMax];
    uint64_t len[kFoldMax];
};

__device__ __forceinline__ Fr29 fold_scalar(const Fr* s)
{
    Fr29 x = Fr29::shl5(Fr29::load32(s));
#pragma unroll
    for (int w = 0; w < 9; ++w)
        x.l[w] = __builtin_amdgcn_readfirstlane(x.l[w]);
    return x;
}

__global__ void __launch_bounds__(kT) multiply_add_many_kernel(Fr* __restrict__ acc, const FoldJobs jobs,
    const Fr* __restrict__ scalars, unsigned count, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        Fr29 a = Fr29::load32(acc + i);
        unsigned j = 0;
        for (; j + 1 < count; j += 2) {
            const Fr29 x = i < jobs.len[j] ? Fr29::load32(jobs.p[j] + i) : Fr29::zero();
            const Fr29 y = i < jobs.len[j + 1] ? Fr29::load32(jobs.p[j + 1] + i) : Fr29::zero();
            a = Fr29::normalize(Fr29::add(a, Fr29::mul_sum(x, fold_scalar(scalars + j), y, fold_scalar(scalars + j + 1))));
        }
        if (j < count && i < jobs.len[j])
            a = Fr29::normalize(Fr29::add(a, Fr29::mul(Fr29::load32(jobs.p[j] + i), fold_scalar(scalars + j))));
        Fr29::reduce_full(a).store32(acc + i);
    }
}

__global__ void elementwise_mul_kernel(Fr* out, const Fr* a, const Fr* b, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        (Fr::load(a + i) * Fr::load(b + i)).store(out + i);
}

__global__ void fill_one_kernel(Fr* __restrict__ out, uint64_t n)
{
    const Fr one = Fr::one();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        one.store(out + i);
}

__global__ void sub_short_kernel(Fr* out, const Fr* lng, const Fr* shrt, uint64_t short_len, uint64_t long_len)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < long_len;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const Fr a = Fr::load(lng + i);
        (i < short_len ? a - Fr::load(shrt + i) : a).store(out + i);
    }
}

__global__ void sub_scalar_at_zero_kernel(Fr* buf, const Fr* s)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        (Fr::load(buf) - Fr::load(s)).store(buf);
}

unsigned ew_blocks(uint64_t n)
{
    uint64_t b = (n + kT - 1) / kT;
    if (b > 4096)
        b = 4096;
    return b ? (unsigned)b : 1u;
}

// --------------------------------------------------------- kate division
//
// q_j = b_j + u q_{j+1} for j < len, q_len = 0, b_j = src[j + 1], as a suffix scan
// on three levels, every level on the 29-bit form (u and its powers in the 2^261
// scale, data in the stored 2^256 scale):
//  * outputs: tiles of kKateTile = 2048, one 256-thread work-group each; thread k
//    owns the 8 consecutive outputs [8k, 8k + 8) of its tile, read through LDS (the
//    tile is loaded with coalesced reads and transposed there);
//  * threads (scan pass): a thread solves its outputs from its top with a zero
//    boundary (7 dependent products) and a Kogge-Stone suffix scan over the
//    work-group in LDS (8 steps, multiplier u^(8d) at distance d) gives S'_k, the value
//    at thread k's bottom with a zero tile top; S'_k is kept for the output pass and
//    S'_0 = A_t is the tile's;
//  * tiles: the A_t are combined the same way by one work-group (kate_carries_kernel)
//    into the true values C_t = q_(tT);
//  * output pass: thread k's carry c_k = S'_(k+1) + u^(8 (255 - k)) C_(t+1) (one
//    product, the power from a table), then its 8 outputs by the recurrence itself.
// 7 + 8 + 1 + 8 products per thread of 8 outputs, at most 15 of them dependent in one
// pass. Launches: powers of u (one work-group), the scan pass (plus one work-group
// squaring out the carries' multipliers beside the tiles), the carries (more than one
// tile only), the outputs.

constexpr unsigned kKateT = 256; // threads per tile work-group
constexpr unsigned kKateE = 8; // outputs per thread
constexpr uint64_t kKateTile = (uint64_t)kKateT * kKateE;
constexpr unsigned kKateScanSteps = 8; // log2 kKateT
constexpr unsigned kKateP8 = kKateT + 1; // u^(8 j), j = 0..kKateT
constexpr unsigned kKateP8Stride = 264;

// powers of u in the 2^261 scale, normalised 29-bit limbs below 2m
struct KatePw {
    uint32_t u[16];
    uint32_t carry[kKateScanSteps][16]; // u^(kKateTile g 2^s), g = 2^log_g tiles per carries thread
    uint32_t p8[9][kKateP8Stride]; // limb w of u^(8 j) at p8[w][j]
};

// the thread's scanned values of one tile, limb planes: S'_k limb w at [w * kKateT + k]
constexpr uint64_t kKateScanWords = 9ull * kKateT;

__device__ __forceinline__ Fr29 pw_u(const KatePw* pw)
{
    Fr29 r;
#pragma unroll
    for (int w = 0; w < 9; ++w)
        r.l[w] = pw->u[w];
    return r;
}

__device__ __forceinline__ Fr29 pw_p8(const KatePw* pw, unsigned j)
{
    Fr29 r;
#pragma unroll
    for (int w = 0; w < 9; ++w)
        r.l[w] = pw->p8[w][j];
    return r;
}

// the same value held per lane: products of uniform operands would otherwise go to the
// scalar unit, which runs a 29-bit product about three times slower than one wave
__device__ __forceinline__ Fr29 per_lane(Fr29 x)
{
#pragma unroll
    for (int w = 0; w < 9; ++w)
        asm volatile("" : "+v"(x.l[w]));
    return x;
}

// LDS slot of tile element e: a permutation inside each 64-slot block, so that both
// the coalesced pattern (64 consecutive e per wave) and the per-thread pattern
// (e = 8 lane + i) put the 64 lanes of a wave on distinct banks
__device__ __forceinline__ unsigned kate_slot(unsigned e) { return e ^ ((e >> 6) & 7u); }

// x + y mult, brought below 2m: x, y, mult normalised below 2m (or x with limbs
// < 2^30 below 2.1m)
__device__ __forceinline__ Fr29 kate_step(const Fr29& x, const Fr29& y, const Fr29& mult)
{
    return Fr29::reduce_2m(Fr29::normalize(Fr29::add(x, Fr29::mul(y, mult))));
}

// Kogge-Stone suffix scan over the work-group's kKateT threads with a zero top: with
// X_k = x of thread k, returns S_k = sum_(k <= i < kKateT) X_i U^(i - k),
// U^(2^s) = mult(s). x normalised below 2m; returns normalised below 2m.
// lds: 2 * 9 * kKateT words.
template <class Mult>
__device__ __forceinline__ Fr29 kate_block_scan(Fr29 x, uint32_t* lds, Mult mult)
{
    const unsigned k = threadIdx.x;
#pragma unroll 1
    for (unsigned s = 0; s < kKateScanSteps; ++s) {
        const unsigned d = 1u << s;
        // double buffer: the step after next rewrites this buffer only after the next
        // barrier, which every thread reaches after its read below
        uint32_t* buf = lds + (s & 1) * 9 * kKateT;
#pragma unroll
        for (int w = 0; w < 9; ++w)
            buf[w * kKateT + k] = x.l[w];
        __syncthreads();
        if (k + d < kKateT) {
            Fr29 y;
#pragma unroll
            for (int w = 0; w < 9; ++w)
                y.l[w] = buf[w * kKateT + k + d];
            x = kate_step(x, y, mult(s));
        }
    }
    return x;
}

// One work-group of kKateT threads: u, and u^(8 j) for j <= kKateT by doubling in LDS
// (u^8 by 3 squarings, then 8 rounds in which threads i < h set
// p[h + 1 + i] = p[1 + i] p[h]): 11 dependent products.
__global__ void __launch_bounds__(kKateT) kate_table_kernel(const Fr* __restrict__ d_u, KatePw* __restrict__ pw)
{
    __shared__ uint32_t p[9][kKateP8Stride];
    const unsigned t = threadIdx.x;
    // u in the 2^261 scale: its stored word shifted by 5 bits (< 32m), reduced below 2m
    const Fr29 u = per_lane(Fr29::reduce_2m(Fr29::shl5(Fr29::load32(d_u))));
    if (t == 0) {
        const Fr29 u8 = Fr29::sqr(Fr29::sqr(Fr29::sqr(u)));
        const Fr29 one = Fr29::one();
#pragma unroll
        for (int w = 0; w < 9; ++w) {
            p[w][0] = one.l[w];
            p[w][1] = u8.l[w];
            pw->u[w] = u.l[w];
        }
    }
    __syncthreads();
#pragma unroll 1
    for (unsigned h = 1; h < kKateT; h <<= 1) {
        if (t < h) {
            Fr29 a, b;
#pragma unroll
            for (int w = 0; w < 9; ++w) {
                a.l[w] = p[w][1 + t];
                b.l[w] = p[w][h];
            }
            const Fr29 r = Fr29::mul(a, b);
#pragma unroll
            for (int w = 0; w < 9; ++w)
                p[w][h + 1 + t] = r.l[w];
        }
        __syncthreads();
    }
    for (unsigned j = t; j < kKateP8; j += kKateT)
#pragma unroll
        for (int w = 0; w < 9; ++w)
            pw->p8[w][j] = p[w][j];
}

// Tile passes. SCAN: S'_k of every thread into scan[t * kKateScanWords ...] and A_t into
// agg[t]; block `tiles` (one past the last tile) squares out the carries' multipliers
// u^(kKateTile 2^(log_g + s)) instead. !SCAN: the outputs q_j, j < min(cap, len), of
// tile t with the top carry C_(t+1) = carry[t + 1] (zero when carry is null: a single
// tile).
template <bool SCAN>
__global__ void __launch_bounds__(kKateT) kate_tile_kernel(const Fr* __restrict__ src, Fr* __restrict__ q,
    uint64_t len, uint64_t cap, KatePw* __restrict__ pw, uint32_t* __restrict__ scan, Fr* __restrict__ agg,
    const Fr* __restrict__ carry, uint64_t tiles, unsigned log_g)
{
    // the tile's words: 8 planes of kKateTile slots (the scan pass then takes the
    // first 2 * 9 * kKateT words for the scan)
    __shared__ uint32_t lds[8 * kKateTile];
    if (SCAN && blockIdx.x == tiles) {
        if (threadIdx.x < 64) {
            Fr29 x = per_lane(pw_p8(pw, kKateT)); // u^kKateTile
#pragma unroll 1
            for (unsigned i = 0; i < log_g; ++i)
                x = Fr29::sqr(x);
#pragma unroll 1
            for (unsigned s = 0; s < kKateScanSteps; ++s) {
                if (threadIdx.x == 0)
#pragma unroll
                    for (int w = 0; w < 9; ++w)
                        pw->carry[s][w] = x.l[w];
                if (s + 1 < kKateScanSteps)
                    x = Fr29::sqr(x);
            }
        }
        return; // the whole work-group
    }
    const unsigned k = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * kKateTile;
#pragma unroll
    for (unsigned r = 0; r < kKateE; ++r) {
        const unsigned e = k + r * kKateT;
        const uint64_t j = base + e;
        uint4 a = make_uint4(0, 0, 0, 0), b = make_uint4(0, 0, 0, 0);
        if (j < len) {
            const uint4* p = (const uint4*)(src + j + 1);
            a = p[0];
            b = p[1];
        }
        const unsigned sl = kate_slot(e);
        lds[0 * kKateTile + sl] = a.x;
        lds[1 * kKateTile + sl] = a.y;
        lds[2 * kKateTile + sl] = a.z;
        lds[3 * kKateTile + sl] = a.w;
        lds[4 * kKateTile + sl] = b.x;
        lds[5 * kKateTile + sl] = b.y;
        lds[6 * kKateTile + sl] = b.z;
        lds[7 * kKateTile + sl] = b.w;
    }
    const Fr29 u = pw_u(pw);
    auto tile_word = [&](unsigned i) {
        const unsigned sl = kate_slot(kKateE * k + i);
        uint32_t v[8];
#pragma unroll
        for (int w = 0; w < 8; ++w)
            v[w] = lds[w * kKateTile + sl];
        return Fr29::from32(v);
    };
    uint32_t* my_scan = scan + blockIdx.x * kKateScanWords;
    if (SCAN) {
        __syncthreads();
        // the thread's bottom value with a zero boundary at its top; limbs < 2^30,
        // values < 2.02m (a canonical b plus a product below 1.02m)
        Fr29 x = tile_word(kKateE - 1);
#pragma unroll
        for (int i = kKateE - 2; i >= 0; --i)
            x = Fr29::add(tile_word(i), Fr29::mul(x, u));
        __syncthreads(); // every thread has its words: the LDS is free
        const Fr29 s = kate_block_scan(Fr29::reduce_2m(Fr29::normalize(x)), lds,
            [&](unsigned st) { return pw_p8(pw, 1u << st); });
#pragma unroll
        for (int w = 0; w < 9; ++w)
            my_scan[w * kKateT + k] = s.l[w];
        if (k == 0)
            Fr29::reduce_full(s).store32(agg + blockIdx.x);
        return;
    }
    // c_k = S'_(k+1) + u^(8 (kKateT - 1 - k)) C_(t+1): the value just above thread k's
    // outputs (S'_kKateT = 0)
    Fr29 c = Fr29::zero();
    if (k + 1 < kKateT) {
#pragma unroll
        for (int w = 0; w < 9; ++w)
            c.l[w] = my_scan[w * kKateT + k + 1];
    }
    if (carry)
        c = kate_step(c, Fr29::load32(carry + blockIdx.x + 1), pw_p8(pw, kKateT - 1 - k));
    __syncthreads();
    // the outputs from the thread's top; each slot is read and rewritten by its thread
#pragma unroll
    for (int i = kKateE - 1; i >= 0; --i) {
        c = Fr29::add(tile_word(i), Fr29::mul(c, u));
        uint32_t v[8];
        Fr29::reduce_full(Fr29::normalize(c)).to32(v);
        const unsigned sl = kate_slot(kKateE * k + i);
#pragma unroll
        for (int w = 0; w < 8; ++w)
            lds[w * kKateTile + sl] = v[w];
    }
    __syncthreads();
    const uint64_t lim = cap < len ? cap : len;
#pragma unroll
    for (unsigned r = 0; r < kKateE; ++r) {
        const unsigned e = k + r * kKateT;
        const uint64_t j = base + e;
        if (j < lim) {
            const unsigned sl = kate_slot(e);
            uint4* p = (uint4*)(q + j);
            p[0] = make_uint4(lds[0 * kKateTile + sl], lds[1 * kKateTile + sl], lds[2 * kKateTile + sl],
                lds[3 * kKateTile + sl]);
            p[1] = make_uint4(lds[4 * kKateTile + sl], lds[5 * kKateTile + sl], lds[6 * kKateTile + sl],
                lds[7 * kKateTile + sl]);
        }
    }
}

// One work-group: thread k takes tiles [k g, k g + g) (g = 2^log_g, kKateT g >= tiles),
// solves them from the top with a zero boundary (C_t = A_t + u^kKateTile C_(t+1)), scans
// the threads' bottom values (multiplier u^(kKateTile g d) at distance d) and solves its
// tiles again from its true top carry: carry[t] = C_t for t < tiles, carry[tiles] = 0.
__global__ void __launch_bounds__(kKateT) kate_carries_kernel(const Fr* __restrict__ agg, Fr* __restrict__ carry,
    uint64_t tiles, unsigned log_g, const KatePw* __restrict__ pw)
{
    __shared__ uint32_t lds[3 * 9 * kKateT];
    const unsigned k = threadIdx.x;
    const uint64_t g = 1ull << log_g;
    const uint64_t lo = (uint64_t)k * g;
    const Fr29 ut = pw_p8(pw, kKateT);
    Fr29 x = Fr29::zero();
#pragma unroll 1
    for (uint64_t i = g; i-- > 0;)
        if (lo + i < tiles)
            x = kate_step(Fr29::load32(agg + lo + i), x, ut);
    const Fr29 s = kate_block_scan(x, lds, [&](unsigned st) {
        Fr29 r;
#pragma unroll
        for (int w = 0; w < 9; ++w)
            r.l[w] = pw->carry[st][w];
        return r;
    });
    uint32_t* xc = lds + 2 * 9 * kKateT;
#pragma unroll
    for (int w = 0; w < 9; ++w)
        xc[w * kKateT + k] = s.l[w];
    __syncthreads();
    Fr29 c = Fr29::zero();
    if (k + 1 < kKateT) {
#pragma unroll
        for (int w = 0; w < 9; ++w)
            c.l[w] = xc[w * kKateT + k + 1];
    }
#pragma unroll 1
    for (uint64_t i = g; i-- > 0;)
        if (lo + i < tiles) {
            c = kate_step(Fr29::load32(agg + lo + i), c, ut);
            Fr29::reduce_full(c).store32(carry + lo + i);
        }
    if (k == 0)
        Fr29::zero().store32(carry + tiles);
}

uint64_t kate_tiles(uint64_t len) { return (len + kKateTile - 1) / kKateTile; }

struct KateBufs {
    KatePw* pw;
    uint32_t* scan;
    Fr* agg;
    Fr* carry;
};

template <class WS>
KateBufs kate_carve(WS& ws, uint64_t len)
{
    const uint64_t m = kate_tiles(len);
    KateBufs b;
    b.pw = (KatePw*)ws.take(sizeof(KatePw));
    b.scan = (uint32_t*)ws.take(4 * kKateScanWords * m);
    b.agg = (Fr*)ws.take(32 * m);
    b.carry = m > 1 ? (Fr*)ws.take(32 * (m + 1)) : nullptr;
    return b;
}

uint64_t kate_scratch(uint64_t n)
{
    if (n < 2)
        return 0;
    Sizer z;
    kate_carve(z, n - 1);
    return z.s.bytes + 256;
}

// q_j = src(j + 1) + u q_{j + 1} for j < len with q_len = 0, written for j < cap
// (src and q must not overlap).
RustError kate_solve(const Fr* src, Fr* q, uint64_t len, uint64_t cap, const Fr* d_u, Workspace& ws,
    hipStream_t stream)
{
    const uint64_t m = kate_tiles(len);
    const KateBufs b = kate_carve(ws, len);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "kate division: workspace too small for %lu outputs",
            (unsigned long)len);
    if (m + 1 > 0x7fffffffull)
        return fail((int)hipErrorInvalidValue, "kate division: %lu outputs exceed the grid", (unsigned long)len);
    unsigned log_g = 0;
    while (((uint64_t)kKateT << log_g) < m)
        ++log_g;
    const uint64_t out_tiles = kate_tiles(cap < len ? cap : len);
    if (out_tiles == 0)
        return ok();
    kate_table_kernel<<<1, kKateT, 0, stream>>>(d_u, b.pw);
    H2_LAUNCH_CHECK();
    kate_tile_kernel<true><<<(unsigned)(m + (m > 1 ? 1 : 0)), kKateT, 0, stream>>>(src, nullptr, len, cap, b.pw,
        b.scan, b.agg, nullptr, m, log_g);
    H2_LAUNCH_CHECK();
    if (m > 1) {
        kate_carries_kernel<<<1, kKateT, 0, stream>>>(b.agg, b.carry, m, log_g, b.pw);
        H2_LAUNCH_CHECK();
    }
    kate_tile_kernel<false><<<(unsigned)out_tiles, kKateT, 0, stream>>>(src, q, len, cap, b.pw, b.scan, nullptr,
        b.carry, m, log_g);
    H2_LAUNCH_CHECK();
    return ok();
}

__global__ void zero_fill_kernel(Fr* __restrict__ out, uint64_t lo, uint64_t hi)
{
    const Fr z = Fr::zero();
    for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi;
         i += (uint64_t)gridDim.x * blockDim.x)
        z.store(out + i);
}

// ------------------------------------------------------ Assigned decode

__global__ void decode_assigned_kernel(Fr* __restrict__ nums, Fr* __restrict__ denoms, const uint8_t* __restrict__ raw,
    uint64_t n, AssignedLayout layout)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint8_t* base = raw + i * (uint64_t)layout.stride_bytes;
        const uint8_t tag = base[0];
        Fr num = Fr::zero(), den = Fr::one();
        if (tag != 0) {
            const uint32_t* p = (const uint32_t*)(base + layout.num_offset);
#pragma unroll
            for (int k = 0; k < 8; ++k)
                num.v[k] = p[k];
            if (tag != 1) {
                const uint32_t* d = (const uint32_t*)(base + layout.denom_offset);
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    den.v[k] = d[k];
            }
        }
        num.store(nums + i);
        den.store(denoms + i);
    }
}

// ----------------------------------------------------------- omega tables

constexpr uint32_t kDense = 10; // low-degree table: omega^i, i < 2^10

// lut[0] = 1, lut[i] = omega^(2^(i-1)), i = 1..log_n
__global__ void omega_log_lut_kernel(const Fr* __restrict__ d_omega, Fr* __restrict__ lut, uint32_t log_n)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    Fr w = Fr::load(d_omega);
    Fr::one().store(lut);
    for (uint32_t i = 1; i <= log_n; ++i) {
        w.store(lut + i);
        w = Fr::sqr(w);
    }
}

__global__ void power_from_lut_kernel(const Fr* __restrict__ lut, uint32_t pow, Fr* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    Fr r = Fr::one();
    for (uint32_t b = 0; b < 32; ++b)
        if ((pow >> b) & 1u)
            r = r * Fr::load(lut + b + 1);
    r.store(out);
}

// out[i] = (base^(2^shift))^i for i < count
__global__ void powers_kernel(const Fr* __restrict__ d_base, uint32_t shift, Fr* __restrict__ out, uint64_t count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    Fr b = Fr::load(d_base);
    for (uint32_t s = 0; s < shift; ++s)
        b = Fr::sqr(b);
    Fr::pow_u64(b, i).store(out + i);
}

// out[i - start] = omega^row * delta^col, (col, row) = mapping[i], i in [start, end]
__global__ void omegadelta_kernel(Fr* __restrict__ out, const uint32_t* __restrict__ mapping,
    const Fr* __restrict__ lo, const Fr* __restrict__ hi, const Fr* __restrict__ dl, uint64_t start, uint64_t end)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t i = start + t;
    if (i > end)
        return;
    const uint32_t col = mapping[2 * i], row = mapping[2 * i + 1];
    (Fr::load(lo + (row & ((1u << kDense) - 1))) * Fr::load(hi + (row >> kDense)) * Fr::load(dl + col)).store(out + t);
}

uint32_t bit_width(uint32_t x) { return x ? 32 - __builtin_clz(x) : 0; }

uint32_t ceil_log2(uint64_t x)
{
    uint32_t l = 0;
    while ((1ull << l) < x)
        ++l;
    return l;
}

// ------------------------------------------------------- host batch helpers

struct BatchBufs {
    Fr* poly[2];
    Fr* acc;
    Fr* points;
    Fr* results;
    Fr* challenges;
    EvalBufs ev;
};

template <class WS>
BatchBufs batch_carve(WS& ws, uint64_t len, uint64_t batch, bool with_acc)
{
    BatchBufs b {};
    b.poly[0] = (Fr*)ws.take(32 * len);
    b.poly[1] = (Fr*)ws.take(32 * len);
    if (with_acc) {
        b.acc = (Fr*)ws.take(32 * len);
        b.challenges = (Fr*)ws.take(32 * batch);
    }
    b.points = (Fr*)ws.take(32 * batch);
    b.results = (Fr*)ws.take(32 * batch);
    b.ev = eval_carve(ws, len);
    return b;
}

uint64_t batch_bytes(uint64_t len, uint64_t batch, bool with_acc)
{
    Sizer z;
    batch_carve(z, len, batch, with_acc);
    return z.s.bytes + 256;
}

// Evaluates host polys at host points (results to host) and, with an
// accumulator, folds acc += challenge_i * poly_i (acc to host).
RustError host_batch(const FFITraitObject* polys, const FFITraitObject* points, const FFITraitObject* results,
    const FFITraitObject* challenges, const FFITraitObject* acc_out, uint64_t offset, uint64_t len, uint64_t batch,
    void* scratch, uint64_t scratch_bytes, hipStream_t stream, const char* who)
{
    Workspace ws(scratch, scratch_bytes);
    const bool with_acc = acc_out != nullptr;
    BatchBufs b = batch_carve(ws, len, batch, with_acc);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "%s: workspace of %lu bytes is too small", who,
            (unsigned long)scratch_bytes);
    if (batch == 0)
        return ok();
    H2_TRY(hipMemcpyAsync(b.points, points->ptr, 32 * batch, hipMemcpyHostToDevice, stream));
    if (with_acc) {
        H2_TRY(hipMemcpyAsync(b.challenges, challenges->ptr, 32 * batch, hipMemcpyHostToDevice, stream));
        zero_fill_kernel<<<ew_blocks(len), kT, 0, stream>>>(b.acc, 0, len);
        H2_LAUNCH_CHECK();
    }
    for (uint64_t i = 0; i < batch; ++i) {
        Fr* cur = b.poly[i & 1];
        H2_TRY(hipMemcpyAsync(cur, polys[i].ptr + 4 * offset, 32 * len, hipMemcpyHostToDevice, stream));
        RustError e = eval_device(cur, len, b.points + i, b.ev, b.results + i, stream);
        if (e.code)
            return e;
        if (with_acc) {
            multiply_add_kernel<<<ew_blocks(len), kT, 0, stream>>>(b.acc, cur, b.challenges + i, len);
            H2_LAUNCH_CHECK();
        }
    }
    if (with_acc)
        H2_TRY(hipMemcpyAsync(acc_out->ptr + 4 * offset, b.acc, 32 * len, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipMemcpyAsync(results->ptr, b.results, 32 * batch, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

} // namespace

extern "C" {

// ------------------------------------------------------------ evaluation

uint64_t _halo2_eval_polynomial_workspace_size(uint64_t length) { return eval_scratch(length); }

RustError _halo2_eval_polynomial(const void* d_poly, const FFITraitObject* point, void** d_result_out,
    uint64_t length, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    Workspace ws(scratch, scratch_bytes);
    EvalBufs b = eval_carve(ws, length);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_eval_polynomial: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    H2_TRY(hipMemcpyAsync(b.point, point->ptr, 32, hipMemcpyHostToDevice, stream));
    RustError e = eval_device((const Fr*)d_poly, length, b.point, b, b.result, stream);
    if (e.code)
        return e;
    *d_result_out = b.result;
    return ok();
}

uint64_t h2amd_eval_polynomial_batch_device_workspace_size(const uint64_t* lengths, uint32_t count)
{
    uint64_t maxb = 1;
    for (uint32_t i = 0; i < count; ++i)
        if (eval_blocks(lengths[i]) > maxb)
            maxb = eval_blocks(lengths[i]);
    const uint64_t per = count < kEvalMaxJobs ? (count ? count : 1) : kEvalMaxJobs;
    return align_up(32 * maxb * per, 256) + 256;
}

RustError h2amd_eval_polynomial_batch_device(const void* const* d_polys, const uint64_t* lengths, const void* h_points,
    uint32_t count, void* d_results, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (count == 0)
        return ok();
    uint64_t maxb = 1;
    for (uint32_t i = 0; i < count; ++i)
        if (eval_blocks(lengths[i]) > maxb)
            maxb = eval_blocks(lengths[i]);
    const uint64_t per = count < kEvalMaxJobs ? count : kEvalMaxJobs;
    Workspace ws(scratch, scratch_bytes);
    Fr* partial = (Fr*)ws.take(32 * maxb * per);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "h2amd_eval_polynomial_batch_device: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    for (uint32_t c0 = 0; c0 < count; c0 += kEvalMaxJobs) {
        const uint32_t m = count - c0 < kEvalMaxJobs ? count - c0 : kEvalMaxJobs;
        EvalJobs jobs;
        for (uint32_t i = 0; i < m; ++i) {
            Fr u;
            memcpy(u.v, (const uint8_t*)h_points + 32ull * (c0 + i), 32);
            jobs.j[i] = EvalJob { (const Fr*)d_polys[c0 + i], lengths[c0 + i], nullptr, u, (Fr*)d_results + c0 + i };
        }
        RustError e = eval_jobs(jobs, m, maxb, partial, stream);
        if (e.code)
            return e;
    }
    return ok();
}

size_t _halo2_eval_poly_batch_max_len(size_t poly_length, size_t batch_size, uint64_t free_bytes)
{
    for (uint64_t l = poly_length; l > 0; l >>= 1)
        if (batch_bytes(l, batch_size, false) < free_bytes)
            return (size_t)l;
    return 0;
}

uint64_t _halo2_eval_polynomial_batch_workspace_size(uint64_t poly_length, uint64_t batch_size)
{
    return batch_bytes(poly_length, batch_size, false);
}

RustError _halo2_eval_polynomial_batch(const FFITraitObject* poly_in_many, const FFITraitObject* eval_points,
    const FFITraitObject* eval_result, size_t poly_offset, size_t poly_length, size_t batch_size, void* scratch,
    uint64_t scratch_bytes, void* stream)
{
    return host_batch(poly_in_many, eval_points, eval_result, nullptr, nullptr, poly_offset, poly_length, batch_size,
        scratch, scratch_bytes, (hipStream_t)stream, "_halo2_eval_polynomial_batch");
}

size_t _halo2_multiopen_poly_max_len(size_t poly_length, size_t batch_size, uint64_t free_bytes)
{
    for (uint64_t l = poly_length; l > 0; l >>= 1)
        if (batch_bytes(l, batch_size, true) < free_bytes)
            return (size_t)l;
    return 0;
}

uint64_t _halo2_multiopen_poly_calculation_workspace_size(uint64_t poly_length, uint64_t batch_size)
{
    return batch_bytes(poly_length, batch_size, true);
}

RustError _halo2_multiopen_poly_calculation(const FFITraitObject* poly_in_many, const FFITraitObject* poly_acc,
    size_t poly_offset, size_t poly_length, size_t batch_size, const FFITraitObject* challenge_point,
    const FFITraitObject* evaluate_point, const FFITraitObject* evaluate_result, void* scratch, uint64_t scratch_bytes,
    void* stream)
{
    return host_batch(poly_in_many, evaluate_point, evaluate_result, challenge_point, poly_acc, poly_offset,
        poly_length, batch_size, scratch, scratch_bytes, (hipStream_t)stream, "_halo2_multiopen_poly_calculation");
}

// ---------------------------------------------------------- elementwise

RustError _halo2_poly_multiply_add(void* d_poly_acc, const void* d_poly_in, const void* d_scalar, size_t poly_length,
    void* stream)
{
    if (poly_length == 0)
        return ok();
    multiply_add_kernel<<<ew_blocks(poly_length), kT, 0, (hipStream_t)stream>>>((Fr*)d_poly_acc, (const Fr*)d_poly_in,
        (const Fr*)d_scalar, poly_length);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError h2amd_poly_multiply_add_many(void* d_acc, const void* const* d_polys, const uint64_t* lengths,
    const void* d_scalars, uint32_t count, uint64_t n, void* stream)
{
    for (uint32_t j = 0; j < count; ++j)
        H2_CHECK_ARG(lengths[j] <= n, "h2amd_poly_multiply_add_many: polynomial %u has %lu coefficients, more than %lu",
            j, (unsigned long)lengths[j], (unsigned long)n);
    if (n == 0)
        return ok();
    for (uint32_t c0 = 0; c0 < count; c0 += kFoldMax) {
        const uint32_t m = count - c0 < kFoldMax ? count - c0 : kFoldMax;
        FoldJobs jobs;
        for (uint32_t j = 0; j < kFoldMax; ++j) {
            jobs.p[j] = j < m ? (const Fr*)d_polys[c0 + j] : nullptr;
            jobs.len[j] = j < m ? lengths[c0 + j] : 0;
        }
        multiply_add_many_kernel<<<ew_blocks(n), kT, 0, (hipStream_t)stream>>>((Fr*)d_acc, jobs,
            (const Fr*)d_scalars + c0, m, n);
        H2_LAUNCH_CHECK();
    }
    return ok();
}

RustError _halo2_poly_elementwise_multiply(void* d_out, const void* d_a, const void* d_b, uint64_t length,
    void* stream)
{
    if (length == 0)
        return ok();
    elementwise_mul_kernel<<<ew_blocks(length), kT, 0, (hipStream_t)stream>>>((Fr*)d_out, (const Fr*)d_a,
        (const Fr*)d_b, length);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_poly_fill_one(void* d_out, uint64_t length, void* stream)
{
    if (length == 0)
        return ok();
    fill_one_kernel<<<ew_blocks(length), kT, 0, (hipStream_t)stream>>>((Fr*)d_out, length);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_poly_sub_short_inplace(void* d_acc, const void* d_short, uint64_t short_len, void* stream)
{
    if (short_len == 0)
        return ok();
    sub_short_kernel<<<ew_blocks(short_len), kT, 0, (hipStream_t)stream>>>((Fr*)d_acc, (const Fr*)d_acc,
        (const Fr*)d_short, short_len, short_len);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_poly_sub_short_out_of_place(void* d_out, const void* d_long, const void* d_short, uint64_t short_len,
    uint64_t long_len, void* stream)
{
    if (long_len == 0)
        return ok();
    sub_short_kernel<<<ew_blocks(long_len), kT, 0, (hipStream_t)stream>>>((Fr*)d_out, (const Fr*)d_long,
        (const Fr*)d_short, short_len, long_len);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_poly_sub_scalar_at_zero(void* d_buf, const void* d_scalar, void* stream)
{
    sub_scalar_at_zero_kernel<<<1, 64, 0, (hipStream_t)stream>>>((Fr*)d_buf, (const Fr*)d_scalar);
    H2_LAUNCH_CHECK();
    return ok();
}

// --------------------------------------------------------- kate division

uint64_t _halo2_kate_division_workspace_size(uint64_t n) { return kate_scratch(n); }

RustError _halo2_kate_division_device(const void* d_a, void* d_q, const void* d_u, uint64_t n, void* scratch,
    uint64_t scratch_bytes, void* stream)
{
    if (n < 2)
        return ok();
    Workspace ws(scratch, scratch_bytes);
    return kate_solve((const Fr*)d_a, (Fr*)d_q, n - 1, n - 1, (const Fr*)d_u, ws, (hipStream_t)stream);
}

RustError _halo2_kate_division_device_padded(const void* d_a, void* d_q, const void* d_u, uint64_t n,
    uint64_t out_len, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (out_len == 0)
        return ok();
    const uint64_t qlen = n < 2 ? 0 : n - 1;
    // out_len < n - 1 keeps the first out_len quotient coefficients
    if (qlen) {
        Workspace ws(scratch, scratch_bytes);
        RustError e = kate_solve((const Fr*)d_a, (Fr*)d_q, qlen, out_len < qlen ? out_len : qlen, (const Fr*)d_u,
            ws, stream);
        if (e.code)
            return e;
    }
    if (out_len > qlen) {
        zero_fill_kernel<<<ew_blocks(out_len - qlen), kT, 0, stream>>>((Fr*)d_q, qlen, out_len);
        H2_LAUNCH_CHECK();
    }
    return ok();
}

// ------------------------------------------------------ Assigned decode

RustError _halo2_decode_assigned(void* d_nums, void* d_denoms, const void* d_raw, uint64_t n, AssignedLayout layout,
    void* stream)
{
    if (n == 0)
        return ok();
    H2_CHECK_ARG(layout.num_offset + 32 <= layout.stride_bytes && layout.denom_offset + 32 <= layout.stride_bytes
            && (layout.num_offset % 4) == 0 && (layout.denom_offset % 4) == 0 && (layout.stride_bytes % 4) == 0,
        "_halo2_decode_assigned: layout (stride %u, num %u, denom %u) does not hold two 32-byte payloads",
        layout.stride_bytes, layout.num_offset, layout.denom_offset);
    decode_assigned_kernel<<<ew_blocks(n), kT, 0, (hipStream_t)stream>>>((Fr*)d_nums, (Fr*)d_denoms,
        (const uint8_t*)d_raw, n, layout);
    H2_LAUNCH_CHECK();
    return ok();
}

// ----------------------------------------------------------- omega tables

uint64_t _halo2_power_of_omega_workspace_size(uint32_t log_n)
{
    return align_up(32, 256) + align_up(32ull * (log_n + 1), 256) + align_up(32, 256) + 256;
}

RustError _halo2_power_of_omega(const FFITraitObject* res, const FFITraitObject* omega_lut,
    const FFITraitObject* omega, uint32_t log_n, uint32_t pow, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    H2_CHECK_ARG(log_n < 32 && bit_width(pow) <= log_n,
        "_halo2_power_of_omega: pow must be < (1 << log_n) and log_n < 32");
    Workspace ws(scratch, scratch_bytes);
    Fr* d_omega = (Fr*)ws.take(32);
    Fr* d_lut = (Fr*)ws.take(32ull * (log_n + 1));
    Fr* d_res = (Fr*)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_power_of_omega: workspace too small");
    H2_TRY(hipMemcpyAsync(d_omega, omega->ptr, 32, hipMemcpyHostToDevice, stream));
    omega_log_lut_kernel<<<1, 64, 0, stream>>>(d_omega, d_lut, log_n);
    H2_LAUNCH_CHECK();
    power_from_lut_kernel<<<1, 64, 0, stream>>>(d_lut, pow, d_res);
    H2_LAUNCH_CHECK();
    H2_TRY(hipMemcpyAsync(omega_lut->ptr, d_lut, 32ull * (log_n + 1), hipMemcpyDeviceToHost, stream));
    H2_TRY(hipMemcpyAsync(res->ptr, d_res, 32, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

uint64_t _halo2_generate_omega_lut_workspace_size(uint32_t log_n)
{
    const uint64_t hi = log_n >= kDense ? 1ull << (log_n - kDense) : 1;
    return align_up(32ull * ((1ull << kDense) + hi), 256) + align_up(32, 256) + 256;
}

RustError _halo2_generate_omega_lut(const FFITraitObject* sparse_twiddle, const FFITraitObject* omega, uint32_t log_n,
    void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    H2_CHECK_ARG(log_n >= kDense && log_n < kDense + 32,
        "_halo2_generate_omega_lut: log_n must be in [%u, %u)", kDense, kDense + 32);
    const uint64_t lo_len = 1ull << kDense, hi_len = 1ull << (log_n - kDense);
    Workspace ws(scratch, scratch_bytes);
    Fr* d_lut = (Fr*)ws.take(32 * (lo_len + hi_len));
    Fr* d_omega = (Fr*)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_generate_omega_lut: workspace too small");
    H2_TRY(hipMemcpyAsync(d_omega, omega->ptr, 32, hipMemcpyHostToDevice, stream));
    powers_kernel<<<blocks_of(lo_len, kT), kT, 0, stream>>>(d_omega, 0, d_lut, lo_len);
    H2_LAUNCH_CHECK();
    powers_kernel<<<blocks_of(hi_len, kT), kT, 0, stream>>>(d_omega, kDense, d_lut + lo_len, hi_len);
    H2_LAUNCH_CHECK();
    H2_TRY(hipMemcpyAsync(sparse_twiddle->ptr, d_lut, 32 * (lo_len + hi_len), hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

uint64_t _halo2_generate_omegadelta_workspace_size(uint32_t log_n, uint32_t omega_start, uint32_t omega_end,
    uint32_t colunm_num, uint32_t colunm_offset)
{
    (void)omega_start;
    (void)omega_end;
    (void)colunm_offset;
    const uint64_t hi = log_n >= kDense ? 1ull << (log_n - kDense) : 1;
    const uint64_t dl = 1ull << ceil_log2(colunm_num < 1 ? 1 : colunm_num);
    return align_up(32ull << kDense, 256) + align_up(32 * hi, 256) + align_up(32 * dl, 256) + 2 * align_up(32, 256)
        + 256;
}

RustError _halo2_generate_omegadelta(void* omegadelta, const void* mapping, const void* omega_host,
    const void* delta_host, uint32_t log_n, uint32_t omega_start, uint32_t omega_end, uint32_t colunm_num,
    uint32_t colunm_offset, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    (void)colunm_offset; // the reference kernel receives it and does not read it
    hipStream_t stream = (hipStream_t)stream_;
    H2_CHECK_ARG(log_n >= kDense && log_n < kDense + 32,
        "_halo2_generate_omegadelta: log_n must be in [%u, %u)", kDense, kDense + 32);
    H2_CHECK_ARG(omega_start <= omega_end, "_halo2_generate_omegadelta: omega_start %u > omega_end %u", omega_start,
        omega_end);
    const uint64_t hi_len = 1ull << (log_n - kDense);
    const uint64_t dl_len = 1ull << ceil_log2(colunm_num < 1 ? 1 : colunm_num);
    Workspace ws(scratch, scratch_bytes);
    Fr* d_lo = (Fr*)ws.take(32ull << kDense);
    Fr* d_hi = (Fr*)ws.take(32 * hi_len);
    Fr* d_dl = (Fr*)ws.take(32 * dl_len);
    Fr* d_omega = (Fr*)ws.take(32);
    Fr* d_delta = (Fr*)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_generate_omegadelta: workspace too small");
    H2_TRY(hipMemcpyAsync(d_omega, omega_host, 32, hipMemcpyHostToDevice, stream));
    H2_TRY(hipMemcpyAsync(d_delta, delta_host, 32, hipMemcpyHostToDevice, stream));
    powers_kernel<<<blocks_of(1u << kDense, kT), kT, 0, stream>>>(d_omega, 0, d_lo, 1u << kDense);
    H2_LAUNCH_CHECK();
    powers_kernel<<<blocks_of(hi_len, kT), kT, 0, stream>>>(d_omega, kDense, d_hi, hi_len);
    H2_LAUNCH_CHECK();
    powers_kernel<<<blocks_of(dl_len, kT), kT, 0, stream>>>(d_delta, 0, d_dl, dl_len);
    H2_LAUNCH_CHECK();
    const uint64_t cnt = (uint64_t)omega_end - omega_start + 1;
    omegadelta_kernel<<<blocks_of(cnt, kT), kT, 0, stream>>>((Fr*)omegadelta, (const uint32_t*)mapping, d_lo, d_hi,
        d_dl, omega_start, omega_end);
    H2_LAUNCH_CHECK();
    return ok();
}

} // extern "C"
