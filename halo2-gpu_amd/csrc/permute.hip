// Lookup permutation pair (permuted input, permuted table) on the device,
// behind the reference's entry point cuda/src/commitment/permute_expression_pair.cu
// (culib.rs:127-141); the result matches permute_expression_pair_seq of
// src/plonk/lookup/prover.rs byte for byte:
//   permuted_input = sort(input[0, U)) under Fr's order (canonical value);
//   at the first row of each run of equal inputs, permuted_table = that value
//   (one table occurrence consumed); the leftover table multiset, ascending,
//   fills the remaining rows from the last one backwards; rows [U, n) are zero.
//
// Sorting: 256-bit canonical keys, merge sort — a tile of 2048 keys is sorted
// in LDS (8 keys per thread sorted in registers, then 8 merge rounds), then
// global merge passes, each work-group merging one 2048-key output tile found
// by a merge-path split. Both arrays (input, table) are sorted by the same
// launches (grid y). Then: not-head flags over the sorted input, leftover flags
// over the sorted table (a table run head whose value occurs in the input is
// consumed), two scans, and one output pass that pairs the r-th repeated row
// with leftover m - 1 - r.
#include "common.h"
#include "field.h"
#include "scan_u32.h"

using namespace h2;

namespace {

constexpr unsigned kThreads = 256; // element-wise kernels
constexpr unsigned kTile = 2048; // keys per sorted tile and per merge work-group, 64 KB of LDS

struct Key {
    uint64_t w[4]; // canonical value, little-endian 64-bit words
};

__device__ __forceinline__ Key key_load(const Key* p)
{
    const uint4* q = (const uint4*)p;
    const uint4 a = q[0], b = q[1];
    Key k;
    k.w[0] = ((uint64_t)a.y << 32) | a.x;
    k.w[1] = ((uint64_t)a.w << 32) | a.z;
    k.w[2] = ((uint64_t)b.y << 32) | b.x;
    k.w[3] = ((uint64_t)b.w << 32) | b.z;
    return k;
}

__device__ __forceinline__ void key_store(Key* p, const Key& k)
{
    uint4* q = (uint4*)p;
    q[0] = make_uint4((uint32_t)k.w[0], (uint32_t)(k.w[0] >> 32), (uint32_t)k.w[1], (uint32_t)(k.w[1] >> 32));
    q[1] = make_uint4((uint32_t)k.w[2], (uint32_t)(k.w[2] >> 32), (uint32_t)k.w[3], (uint32_t)(k.w[3] >> 32));
}

// a < b as 256-bit integers (most significant word first)
__device__ __forceinline__ bool key_lt(const Key& a, const Key& b)
{
    if (a.w[3] != b.w[3])
        return a.w[3] < b.w[3];
    if (a.w[2] != b.w[2])
        return a.w[2] < b.w[2];
    if (a.w[1] != b.w[1])
        return a.w[1] < b.w[1];
    return a.w[0] < b.w[0];
}

__device__ __forceinline__ bool key_eq(const Key& a, const Key& b)
{
    return ((a.w[0] ^ b.w[0]) | (a.w[1] ^ b.w[1]) | (a.w[2] ^ b.w[2]) | (a.w[3] ^ b.w[3])) == 0;
}

__device__ __forceinline__ Key key_max()
{
    Key k;
    k.w[0] = k.w[1] = k.w[2] = k.w[3] = ~0ull; // above every canonical value (< r < 2^254)
    return k;
}

__device__ __forceinline__ Key key_from_fr(const Fr& v)
{
    Key k;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        k.w[i] = ((uint64_t)v.v[2 * i + 1] << 32) | v.v[2 * i];
    return k;
}

__device__ __forceinline__ Fr fr_from_key(const Key& k)
{
    Fr v;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v.v[2 * i] = (uint32_t)k.w[i];
        v.v[2 * i + 1] = (uint32_t)(k.w[i] >> 32);
    }
    return v;
}

__device__ __forceinline__ void cmp_swap(Key& a, Key& b)
{
    if (key_lt(b, a)) {
        const Key t = a;
        a = b;
        b = t;
    }
}

// Number of elements taken from A among the first `diag` outputs of the
// stable merge of A (length na) and B (length nb); ties go to A.
template <class GetA, class GetB>
__device__ __forceinline__ uint64_t merge_path(uint64_t na, uint64_t nb, uint64_t diag, GetA getA, GetB getB)
{
    uint64_t lo = diag > nb ? diag - nb : 0;
    uint64_t hi = diag < na ? diag : na;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (!key_lt(getB(diag - 1 - mid), getA(mid)))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Merges up to P outputs starting at diagonal `diag` of A|B held in LDS.
template <unsigned P>
__device__ __forceinline__ void merge_run(const Key* A, uint32_t na, const Key* B, uint32_t nb, uint32_t diag,
    Key* out)
{
    const uint32_t i0 = (uint32_t)merge_path(na, nb, diag, [A](uint64_t k) { return A[k]; },
        [B](uint64_t k) { return B[k]; });
    uint32_t i = i0, j = diag - i0;
    const uint32_t total = na + nb;
#pragma unroll
    for (unsigned k = 0; k < P; ++k) {
        if (diag + k < total) {
            const bool takeA = j >= nb || (i < na && !key_lt(B[j], A[i]));
            out[k] = takeA ? A[i] : B[j];
            if (takeA)
                ++i;
            else
                ++j;
        }
    }
}

// Sorts each 2048-key tile of keys[y * stride + (0, U)) into out: eight keys
// per thread sorted in registers, then eight merge rounds in LDS.
constexpr unsigned kSortThreads = 256;
constexpr unsigned kSortPer = kTile / kSortThreads;

__global__ void __launch_bounds__(kSortThreads) tile_sort_kernel(const Key* __restrict__ keys,
    Key* __restrict__ out, uint64_t U, uint64_t stride)
{
    __shared__ Key sh[kTile];
    const Key* src = keys + blockIdx.y * stride;
    Key* dst = out + blockIdx.y * stride;
    const uint64_t base = (uint64_t)blockIdx.x * kTile;
    const uint32_t t = threadIdx.x;
    for (uint32_t k = t; k < kTile; k += kSortThreads)
        sh[k] = base + k < U ? key_load(src + base + k) : key_max();
    __syncthreads();
    Key r[kSortPer];
#pragma unroll
    for (unsigned k = 0; k < kSortPer; ++k)
        r[k] = sh[t * kSortPer + k];
    // odd-even transposition sort of the thread's keys in registers
#pragma unroll
    for (unsigned phase = 0; phase < kSortPer; ++phase) {
#pragma unroll
        for (unsigned k = phase & 1; k + 1 < kSortPer; k += 2)
            cmp_swap(r[k], r[k + 1]);
    }
    __syncthreads();
#pragma unroll
    for (unsigned k = 0; k < kSortPer; ++k)
        sh[t * kSortPer + k] = r[k];
    __syncthreads();
    for (uint32_t w = kSortPer; w < kTile; w <<= 1) {
        const uint32_t d = t * kSortPer;
        const uint32_t p = d / (2 * w);
        const Key* A = sh + p * 2 * w;
        merge_run<kSortPer>(A, w, A + w, w, d - p * 2 * w, r);
        __syncthreads();
#pragma unroll
        for (unsigned k = 0; k < kSortPer; ++k)
            sh[d + k] = r[k];
        __syncthreads();
    }
    for (uint32_t k = t; k < kTile; k += kSortThreads)
        if (base + k < U)
            key_store(dst + base + k, sh[k]);
}

// Geometry of output tile `tile` in a merge pass over runs of length w: the
// pair of runs it falls in (A at `pair`, B after it) and its diagonal range.
struct TileSpan {
    uint64_t pair, na, nb, d0, d1;
};

__device__ __forceinline__ TileSpan tile_span(uint64_t tile, uint64_t U, uint64_t w)
{
    TileSpan t;
    const uint64_t g0 = tile * kTile;
    t.pair = g0 / (2 * w) * (2 * w);
    const uint64_t rest = U - t.pair;
    t.na = rest < w ? rest : w;
    t.nb = rest > w ? (rest - w < w ? rest - w : w) : 0;
    t.d0 = g0 - t.pair;
    t.d1 = t.d0 + kTile < t.na + t.nb ? t.d0 + kTile : t.na + t.nb;
    return t;
}

// Merge-path splits of one merge pass: for each key array y, output tile and
// end (0 = first output, 1 = one past the last), the number of keys the tile
// takes from run A, at splits[(y * tiles + tile) * 2 + end]. One wave per
// split narrows the range with 64 probes per round (a 64-ary search), so a
// run of 2^21 keys takes four rounds of dependent loads instead of 21.
__global__ void __launch_bounds__(256) run_split_kernel(const Key* __restrict__ keys, uint64_t* __restrict__ splits,
    uint64_t U, uint64_t stride, uint64_t w, uint32_t tiles)
{
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (wave >= tiles * 4u) // uniform per wave
        return;
    const uint32_t y = wave / (tiles * 2u);
    const uint32_t rem = wave % (tiles * 2u);
    const TileSpan t = tile_span(rem >> 1, U, w);
    const uint64_t d = (rem & 1u) ? t.d1 : t.d0;
    const Key* A = keys + y * stride + t.pair;
    const Key* B = A + t.na;
    // first i in [lo, hi) with B[d - 1 - i] < A[i] (ties go to A), else hi
    uint64_t lo = d > t.nb ? d - t.nb : 0;
    uint64_t hi = d < t.na ? d : t.na;
    while (lo < hi) {
        const uint64_t step = (hi - lo + 63) >> 6;
        const uint64_t i = lo + lane * step;
        const bool p = i < hi && !key_lt(key_load(B + (d - 1 - i)), key_load(A + i));
        const uint32_t c = (uint32_t)__popcll(__ballot(p)); // the predicate holds on a prefix of the probes
        if (c == 0) {
            hi = lo;
        } else {
            const uint64_t first_false = lo + c * step;
            lo = lo + (c - 1) * step + 1;
            if (first_false < hi)
                hi = first_false;
        }
    }
    if (lane == 0)
        splits[wave] = lo;
}

// One merge pass: runs of length w (sorted) merged pairwise into runs of 2w,
// 2048 outputs per work-group of 1024 threads (two each) between the splits
// run_split_kernel found.
constexpr unsigned kMergeThreads = 1024;
constexpr unsigned kMergePer = kTile / kMergeThreads;

__global__ void __launch_bounds__(kMergeThreads) merge_pass_kernel(const Key* __restrict__ keys,
    Key* __restrict__ out, const uint64_t* __restrict__ splits, uint64_t U, uint64_t stride, uint64_t w)
{
    __shared__ Key sh[kTile];
    const Key* src = keys + blockIdx.y * stride;
    Key* dst = out + blockIdx.y * stride;
    const TileSpan ts = tile_span(blockIdx.x, U, w);
    const Key* A = src + ts.pair;
    const Key* B = A + ts.na;
    const uint64_t* sp = splits + ((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * 2;
    const uint64_t a0 = sp[0], a1 = sp[1];
    const uint64_t b0 = ts.d0 - a0, b1 = ts.d1 - a1;
    const uint32_t ta = (uint32_t)(a1 - a0), tb = (uint32_t)(b1 - b0);
    const uint32_t t = threadIdx.x;
    for (uint32_t k = t; k < ta + tb; k += kMergeThreads)
        sh[k] = k < ta ? key_load(A + a0 + k) : key_load(B + b0 + (k - ta));
    __syncthreads();
    Key r[kMergePer];
#pragma unroll
    for (unsigned k = 0; k < kMergePer; ++k)
        r[k] = key_max();
    const uint32_t d = t * kMergePer;
    const uint32_t total = ta + tb;
    merge_run<kMergePer>(sh, ta, sh + ta, tb, d < total ? d : total, r);
    __syncthreads();
#pragma unroll
    for (unsigned k = 0; k < kMergePer; ++k)
        if (d + k < total)
            sh[d + k] = r[k];
    __syncthreads();
    const uint64_t g0 = (uint64_t)blockIdx.x * kTile;
    for (uint32_t k = t; k < ta + tb; k += kMergeThreads)
        key_store(dst + g0 + k, sh[k]);
}

// keys[i] = canonical(input[i]), keys[stride + i] = canonical(table[i]), i < U
__global__ void canonical_keys_kernel(const Fr* __restrict__ input, const Fr* __restrict__ table, Key* __restrict__ keys,
    uint64_t U, uint64_t stride)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= U)
        return;
    key_store(keys + i, key_from_fr(Fr::from_mont(Fr::load(input + i))));
    key_store(keys + stride + i, key_from_fr(Fr::from_mont(Fr::load(table + i))));
}

__device__ bool sorted_contains(const Key* __restrict__ sorted, uint64_t n, const Key& x)
{
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (key_lt(key_load(sorted + mid), x))
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < n && key_eq(key_load(sorted + lo), x);
}

// repeat[i] = 1 when sorted input row i repeats row i - 1; leftover[j] = 0 when
// sorted table row j starts a run whose value occurs in the input
__global__ void flags_kernel(const Key* __restrict__ sin, const Key* __restrict__ stab, uint32_t* __restrict__ repeat,
    uint32_t* __restrict__ leftover, uint64_t U)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= U)
        return;
    repeat[i] = (i > 0 && key_eq(key_load(sin + i), key_load(sin + i - 1))) ? 1u : 0u;
    const Key v = key_load(stab + i);
    const bool head = i == 0 || !key_eq(v, key_load(stab + i - 1));
    leftover[i] = (head && sorted_contains(sin, U, v)) ? 0u : 1u;
}

__global__ void leftover_compact_kernel(const Key* __restrict__ stab, const uint32_t* __restrict__ leftover,
    const uint32_t* __restrict__ lpos, Key* __restrict__ out, uint64_t U)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < U && leftover[j])
        key_store(out + lpos[j], key_load(stab + j));
}

__global__ void permute_output_kernel(const Key* __restrict__ sin, const uint32_t* __restrict__ repeat,
    const uint32_t* __restrict__ rpos, const Key* __restrict__ left, const uint32_t* __restrict__ lpos_total,
    Fr* __restrict__ pin, Fr* __restrict__ ptab, uint64_t U, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (i >= U) {
        Fr::zero().store(pin + i);
        Fr::zero().store(ptab + i);
        return;
    }
    const Fr v = Fr::to_mont(fr_from_key(key_load(sin + i)));
    v.store(pin + i);
    if (!repeat[i]) {
        v.store(ptab + i);
        return;
    }
    // the r-th repeated row (r = rpos[i]) takes leftover m - 1 - r; with an input
    // that is not a sub-multiset of the table the counts differ and rows without a
    // leftover stay zero
    const int64_t idx = (int64_t)*lpos_total - 1 - (int64_t)rpos[i];
    (idx >= 0 ? Fr::to_mont(fr_from_key(key_load(left + idx))) : Fr::zero()).store(ptab + i);
}

struct PermBufs {
    Key* keys[2];
    uint32_t* repeat;
    uint32_t* leftover;
    uint32_t* rpos;
    uint32_t* lpos;
    uint64_t* splits; // merge-path splits of one pass: 2 arrays x tiles x 2
    void* scan;
};

template <class WS>
PermBufs perm_carve(WS& ws, uint64_t U)
{
    PermBufs b;
    b.keys[0] = (Key*)ws.take(2 * 32 * U);
    b.keys[1] = (Key*)ws.take(2 * 32 * U);
    b.repeat = (uint32_t*)ws.take(4 * U);
    b.leftover = (uint32_t*)ws.take(4 * U);
    b.rpos = (uint32_t*)ws.take(4 * (U + 1));
    b.lpos = (uint32_t*)ws.take(4 * (U + 1));
    b.splits = (uint64_t*)ws.take(8 * 4 * ((U + kTile - 1) / kTile));
    b.scan = ws.take(scan_u32_scratch(U));
    return b;
}

struct PermSizer {
    WorkspaceSizer s;
    void* take(uint64_t n)
    {
        s.take(n);
        return nullptr;
    }
};

unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

} // namespace

extern "C" {

uint64_t _halo2_permute_expression_pair_workspace_size(uint64_t n, uint64_t usable_rows)
{
    (void)n;
    if (usable_rows == 0)
        return 0;
    PermSizer z;
    perm_carve(z, usable_rows);
    return z.s.bytes + 256;
}

RustError _halo2_permute_expression_pair(const void* d_compressed_input, const void* d_compressed_table,
    void* d_permuted_input, void* d_permuted_table, uint64_t n, uint64_t usable_rows, void* scratch,
    uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (n == 0)
        return ok();
    const uint64_t U = usable_rows;
    H2_CHECK_ARG(U <= n, "_halo2_permute_expression_pair: usable_rows %lu > n %lu", (unsigned long)U,
        (unsigned long)n);
    H2_CHECK_ARG(U < (1ull << 32), "_halo2_permute_expression_pair: usable_rows must fit 32 bits");
    Workspace ws(scratch, scratch_bytes);
    PermBufs b {};
    if (U) {
        b = perm_carve(ws, U);
        if (ws.overflow)
            return fail((int)hipErrorInvalidValue, "_halo2_permute_expression_pair: workspace of %lu bytes is too small",
                (unsigned long)scratch_bytes);
    }
    if (U) {
        canonical_keys_kernel<<<blocks_for(U, kThreads), kThreads, 0, stream>>>((const Fr*)d_compressed_input,
            (const Fr*)d_compressed_table, b.keys[0], U, U);
        H2_LAUNCH_CHECK();
        const unsigned tiles = blocks_for(U, kTile);
        tile_sort_kernel<<<dim3(tiles, 2), kSortThreads, 0, stream>>>(b.keys[0], b.keys[1], U, U);
        H2_LAUNCH_CHECK();
        int cur = 1;
        for (uint64_t w = kTile; w < U; w <<= 1) {
            run_split_kernel<<<blocks_for(4ull * tiles * 64, 256), 256, 0, stream>>>(b.keys[cur], b.splits, U, U, w,
                tiles);
            H2_LAUNCH_CHECK();
            merge_pass_kernel<<<dim3(tiles, 2), kMergeThreads, 0, stream>>>(b.keys[cur], b.keys[cur ^ 1], b.splits, U, U,
                w);
            H2_LAUNCH_CHECK();
            cur ^= 1;
        }
        const Key* sin = b.keys[cur];
        const Key* stab = b.keys[cur] + U;
        Key* left = b.keys[cur ^ 1];
        flags_kernel<<<blocks_for(U, kThreads), kThreads, 0, stream>>>(sin, stab, b.repeat, b.leftover, U);
        H2_LAUNCH_CHECK();
        RustError e = scan_u32_exclusive(b.repeat, b.rpos, U, b.scan, stream);
        if (e.code)
            return e;
        e = scan_u32_exclusive(b.leftover, b.lpos, U, b.scan, stream);
        if (e.code)
            return e;
        leftover_compact_kernel<<<blocks_for(U, kThreads), kThreads, 0, stream>>>(stab, b.leftover, b.lpos, left, U);
        H2_LAUNCH_CHECK();
        permute_output_kernel<<<blocks_for(n, kThreads), kThreads, 0, stream>>>(sin, b.repeat, b.rpos, left,
            b.lpos + U, (Fr*)d_permuted_input, (Fr*)d_permuted_table, U, n);
        H2_LAUNCH_CHECK();
    } else {
        permute_output_kernel<<<blocks_for(n, kThreads), kThreads, 0, stream>>>(nullptr, nullptr, nullptr, nullptr,
            nullptr, (Fr*)d_permuted_input, (Fr*)d_permuted_table, 0, n);
        H2_LAUNCH_CHECK();
    }
    return ok();
}

} // extern "C"
