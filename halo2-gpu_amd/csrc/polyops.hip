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
// q_j = a_{j+1} + u q_{j+1}, a linear recurrence solved in runs of 64 — each
// run is solved locally with a zero boundary, the run heads obey the same
// recurrence in u^64 (solved recursively), and a fix-up adds u^(e-j) q_e to
// every output of a run ending at e.
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

// outputs per thread in one recurrence level: each level is a chain of kRun
// dependent products on one lane, so short runs (more threads, one more level at
// 2^20) cut the latency that bounds the division
constexpr uint64_t kRun = 32;
constexpr int kMaxLevels = 13; // 32^13 > 2^64

struct KateLevel {
    uint64_t len; // outputs at this level
    uint64_t runs; // ceil(len / kRun)
};

// levels of the recursion for `outputs` outputs (>= 1)
int kate_levels(uint64_t outputs, KateLevel* lv)
{
    int k = 0;
    uint64_t len = outputs;
    while (true) {
        lv[k].len = len;
        lv[k].runs = (len + kRun - 1) / kRun;
        ++k;
        if (lv[k - 1].runs <= 1 || k == kMaxLevels)
            break;
        len = lv[k - 1].runs;
    }
    return k;
}

template <class WS>
void kate_carve(WS& ws, const KateLevel* lv, int levels, Fr** pw, Fr** heads, Fr** solved)
{
    for (int i = 0; i < levels; ++i) {
        pw[i] = (Fr*)ws.take(32 * (kRun + 1));
        heads[i] = (Fr*)ws.take(32 * (lv[i].runs + 1));
        solved[i] = (Fr*)ws.take(32 * (lv[i].runs + 1));
    }
}

uint64_t kate_scratch(uint64_t n)
{
    if (n < 2)
        return 0;
    KateLevel lv[kMaxLevels];
    Fr *pw[kMaxLevels], *heads[kMaxLevels], *solved[kMaxLevels];
    const int levels = kate_levels(n - 1, lv);
    Sizer z;
    kate_carve(z, lv, levels, pw, heads, solved);
    return z.s.bytes + 256;
}

// pw[i] = u^i, i <= kRun
__global__ void kate_powers_kernel(const Fr* __restrict__ d_u, Fr* __restrict__ pw)
{
    for (uint32_t i = threadIdx.x; i <= kRun; i += blockDim.x)
        Fr::pow_u64(Fr::load(d_u), i).store(pw + i);
}

// Local solve of q_j = src(j + 1) + u q_{j + 1} on run r = [r kRun, e) with a
// zero boundary q_e = 0: writes loc_j for j < cap into q and the run head
// loc_{r kRun} into heads[r]. src(i) is read for i in 1..len.
// The recurrence runs on the 29-bit form (field29.h): u enters in the 2^261 scale
// (a 5-bit shift of its stored word), so mul(loc, u) is the stored word of loc * u;
// loc stays below 2.01m (a normalised product below 1.01m plus a canonical word).
__global__ void kate_local_kernel(const Fr* __restrict__ src, Fr* __restrict__ q, uint64_t len, uint64_t cap,
    uint64_t runs, const Fr* __restrict__ pw, Fr* __restrict__ heads)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= runs)
        return;
    const Fr29 u = Fr29::shl5(Fr29::load32(pw + 1));
    const uint64_t lo = r * kRun;
    const uint64_t e = lo + kRun < len ? lo + kRun : len;
    Fr29 loc = Fr29::load32(src + e);
    if (e - 1 < cap)
        Fr29::reduce_full(loc).store32(q + e - 1);
    for (uint64_t j = e - 1; j-- > lo;) {
        loc = Fr29::add(Fr29::load32(src + j + 1), Fr29::mul(loc, u));
        if (j < cap)
            Fr29::reduce_full(Fr29::normalize(loc)).store32(q + j);
    }
    Fr29::reduce_full(Fr29::normalize(loc)).store32(heads + r);
}

// With Q_s the true value at the head of run s: q_j += u^(e - j) Q_{r + 1},
// e = (r + 1) kRun, for every j < cap outside the last run.
__global__ void kate_fix_kernel(Fr* __restrict__ q, uint64_t cap, uint64_t runs, const Fr* __restrict__ pw,
    const Fr* __restrict__ solved)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cap)
        return;
    const uint64_t r = j / kRun;
    if (r + 1 >= runs)
        return;
    const uint64_t e = (r + 1) * kRun;
    const Fr29 t = Fr29::mul(Fr29::shl5(Fr29::load32(pw + (e - j))), Fr29::load32(solved + r + 1));
    Fr29::reduce_full(Fr29::normalize(Fr29::add(Fr29::load32(q + j), t))).store32(q + j);
}

// q_j = src(j + 1) + u q_{j + 1} for j < len with q_len = 0, written for j < cap.
// Level i + 1 solves the heads H_r of level i: Q_r = H_r + u^kRun Q_{r + 1}
// (run r < runs - 1 is full, and Q_runs = 0), i.e. the same recurrence with
// src'(r + 1) = H_r and u' = u^kRun; the solved heads then fix level i.
RustError kate_solve(const Fr* src, Fr* q, uint64_t len, uint64_t cap, const Fr* d_u, Workspace& ws,
    hipStream_t stream)
{
    KateLevel lv[kMaxLevels];
    Fr *pw[kMaxLevels], *heads[kMaxLevels], *solved[kMaxLevels];
    const int levels = kate_levels(len, lv);
    kate_carve(ws, lv, levels, pw, heads, solved);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "kate division: workspace too small for %lu outputs",
            (unsigned long)len);
    const Fr* cur_src = src;
    Fr* cur_q = q;
    uint64_t cur_cap = cap;
    const Fr* cur_u = d_u;
    for (int i = 0; i < levels; ++i) {
        kate_powers_kernel<<<1, kT, 0, stream>>>(cur_u, pw[i]);
        H2_LAUNCH_CHECK();
        kate_local_kernel<<<blocks_of(lv[i].runs, 64), 64, 0, stream>>>(cur_src, cur_q, lv[i].len, cur_cap,
            lv[i].runs, pw[i], heads[i] + 1);
        H2_LAUNCH_CHECK();
        cur_src = heads[i];
        cur_q = solved[i];
        cur_cap = lv[i].runs;
        cur_u = pw[i] + kRun;
    }
    for (int i = levels - 2; i >= 0; --i) {
        Fr* qi = i == 0 ? q : solved[i - 1];
        const uint64_t ci = i == 0 ? cap : lv[i].len;
        if (ci == 0)
            continue;
        kate_fix_kernel<<<blocks_of(ci, kT), kT, 0, stream>>>(qi, ci, lv[i].runs, pw[i], solved[i]);
        H2_LAUNCH_CHECK();
    }
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
