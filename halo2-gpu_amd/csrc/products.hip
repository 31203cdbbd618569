// Batch inversion, the multiplicative prefix scan (grand product) and the
// permutation / lookup product numerators and denominators, mirroring
// cuda/src/polynomial/operations.cu, cuda/src/commitment/grand_product.cu,
// permutation_commit.cu and commit_product.cu of the reference.
//
// Batch inversion: Montgomery's trick at two levels. Each thread folds 8
// elements (zeros are skipped and stay zero, ff::BatchInvert semantics), the
// 256 thread products of a work-group are inverted through a product tree in
// LDS, and one field inversion per 2048 elements finishes the work-group.
//
// Grand product: in-place multiplicative scan. A level scans chunks of 2048
// elements locally (thread runs, then the exclusive prefixes of the thread
// totals from the LDS product tree); the last element of every chunk (the chunk
// product) forms a strided view that is scanned recursively in place, and a
// fix-up pass multiplies each chunk by the scanned product of the chunks before
// it. No scratch memory is needed, matching the scratch-free device entry point.
//
// Everything here runs on the 29-bit limb form of field29.h.
#include "common.h"
#include "field.h"
#include "field29.h"

using namespace h2;

namespace {

constexpr unsigned T = 256; // threads per work-group
constexpr unsigned E = 8; // elements per thread
constexpr unsigned C = T * E; // elements per work-group

using R29 = Fr29;

__device__ __forceinline__ R29 ld29(const Fr* p) { return R29::load32(p); }
// any limbs below 2^32 and value below 2^261 -> the canonical stored word
__device__ __forceinline__ void st29(const R29& x, Fr* p) { R29::reduce_full(R29::normalize(x)).store32(p); }

// Inverses of the T thread products through a product tree in LDS (heap layout:
// node 1 is the root, node i has children 2i and 2i + 1, leaf t is node T + t):
// 255 products up, then the root's inverse is pushed down with 2 products per
// node, in place (a node's slot takes its inverse once its parent's is known) --
// 765 products per work-group instead of the 16 per thread of two Hillis-Steele
// scans. tree_up: the caller has written node[T + tid]; returns after the last
// barrier with node[1] = the product of all leaves.
template <class F>
__device__ __forceinline__ void tree_up(F* node, unsigned tid)
{
    for (unsigned half = T / 2; half >= 1; half >>= 1) {
        __syncthreads();
        if (tid < half) {
            const unsigned i = half + tid;
            node[i] = F::mul(node[2 * i], node[2 * i + 1]);
        }
    }
    __syncthreads();
}

// The caller has replaced node[1] by the root's inverse (and synchronised);
// returns the inverse of leaf tid.
template <class F>
__device__ __forceinline__ F tree_down(F* node, unsigned tid)
{
    for (unsigned half = 1; half < T; half <<= 1) {
        if (tid < half) {
            const unsigned i = half + tid;
            const F v = node[i];
            const F a = node[2 * i], b = node[2 * i + 1];
            node[2 * i] = F::mul(v, b);
            node[2 * i + 1] = F::mul(v, a);
        }
        __syncthreads();
    }
    return node[T + tid];
}

// Batch inversion in the 29-bit-limb form (field29.h). Stored words X = x * 2^256
// are read as 2^261-scale representatives of u = X / 2^261; the trick computes
// u^-1 * 2^261 = 2^522 / X, and one factor 2^-10 (a product with 2^251) applied to
// the work-group's inverse turns every output into 2^512 / X = (x^-1) * 2^256.
template <class P>
__global__ void __launch_bounds__(T) batch_invert_kernel(uint8_t* __restrict__ data, uint64_t n)
{
    using F = F29<P>;
    __shared__ F node[2 * T];
    const unsigned tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * C;
    F part[E];
    F acc = F::one();
#pragma unroll
    for (unsigned k = 0; k < E; ++k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        part[k] = acc;
        if (i < n) {
            const F a = F::load32(data + 32 * i);
            if (!a.is_zero())
                acc = F::mul(acc, a);
        }
    }
    node[T + tid] = acc;
    tree_up(node, tid);
    if (tid == 0) {
        F k251 = F::zero();
        k251.l[8] = 1u << 19; // 2^251 = 2^(232 + 19)
        node[1] = F::mul(F::inverse_vt(node[1]), k251);
    }
    __syncthreads();
    // inverse of this thread's product
    F inv = tree_down(node, tid);
#pragma unroll
    for (int k = (int)E - 1; k >= 0; --k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        if (i < n) {
            const F a = F::load32(data + 32 * i);
            if (!a.is_zero()) {
                F::reduce_full(F::mul(inv, part[k])).store32(data + 32 * i);
                inv = F::mul(inv, a);
            }
        }
    }
}

// ------------------------------------- with a workspace: few inversions
//
// With more than one work-group, the per-group inversion above (a dependent chain
// on one lane) dominates. Callers that hold a
// workspace take three steps instead: (1) every group writes the product of its
// elements to tot[g]; (2) the totals are inverted by the kernel above (one
// inversion per 2048 totals, in parallel groups); (3) every group finishes
// with its inverted total. The values match the single-kernel path bit for bit.

// (1) tot[g] = product of the non-zero elements of group g (2^261-scale word)
template <class P>
__global__ void __launch_bounds__(T) group_total_kernel(const uint8_t* __restrict__ data, uint64_t n,
    uint8_t* __restrict__ tot)
{
    using F = F29<P>;
    __shared__ F red[T];
    const unsigned tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * C;
    F acc = F::one();
#pragma unroll
    for (unsigned k = 0; k < E; ++k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        if (i < n) {
            const F a = F::load32(data + 32 * i);
            if (!a.is_zero())
                acc = F::mul(acc, a);
        }
    }
    red[tid] = acc;
    for (unsigned d = T / 2; d > 0; d >>= 1) {
        __syncthreads();
        if (tid < d)
            red[tid] = F::mul(red[tid], red[tid + d]);
    }
    __syncthreads();
    if (tid == 0)
        F::reduce_full(red[0]).store32(tot + 32ull * blockIdx.x);
}

// (3) as batch_invert_kernel, with the group's inverse total read from inv[g]:
// the kernel above turns the word t * 2^261 into t^-1 * 2^-10 in the 2^261 scale,
// which is exactly the single-group kernel's inv_total.
template <class P>
__global__ void __launch_bounds__(T) group_finish_kernel(uint8_t* __restrict__ data, uint64_t n,
    const uint8_t* __restrict__ inv_tot)
{
    using F = F29<P>;
    __shared__ F node[2 * T];
    const unsigned tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * C;
    F part[E];
    F acc = F::one();
#pragma unroll
    for (unsigned k = 0; k < E; ++k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        part[k] = acc;
        if (i < n) {
            const F a = F::load32(data + 32 * i);
            if (!a.is_zero())
                acc = F::mul(acc, a);
        }
    }
    node[T + tid] = acc;
    tree_up(node, tid);
    if (tid == 0)
        node[1] = F::load32(inv_tot + 32ull * blockIdx.x);
    __syncthreads();
    F inv = tree_down(node, tid);
#pragma unroll
    for (int k = (int)E - 1; k >= 0; --k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        if (i < n) {
            const F a = F::load32(data + 32 * i);
            if (!a.is_zero()) {
                F::reduce_full(F::mul(inv, part[k])).store32(data + 32 * i);
                inv = F::mul(inv, a);
            }
        }
    }
}

// ------------------------------- without a workspace: totals in spare bits
//
// The bare entry point has no workspace, so the group totals travel in the two
// spare top bits of the elements themselves (inputs are canonical, < m < 2^254):
// slot s (s < 8) of group j is the 256 bits held by the 128 elements
// [jC + 128 s, jC + 128 (s + 1)) and carries the total of group 8j + s, then its
// inverse. Readers mask the spare bits; every spare bit is written by one thread.
// The middle step inverts the totals in blocks of kMidPer (one inversion per
// block, the blocks in parallel); the finishing groups run host-last (group j
// hosts 8j .. 8j + 7), so no group overwrites a slot another still has to read.
constexpr unsigned kSlot = 128; // elements per slot, 2 bits each
constexpr unsigned kMidThreads = 64;
constexpr unsigned kMidE = 2; // totals per middle thread: short chains, many blocks
constexpr unsigned kMidPer = kMidThreads * kMidE; // totals per middle block

template <class P>
__device__ __forceinline__ F29<P> load_masked(const uint8_t* p)
{
    const uint4* q = (const uint4*)p;
    const uint4 a = q[0], b = q[1];
    const uint32_t v[8] = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w & 0x3fffffffu };
    return F29<P>::from32(v);
}

__device__ __forceinline__ uint32_t* spare_word(uint8_t* data, uint64_t elem)
{
    return (uint32_t*)(data + 32 * elem) + 7;
}

__device__ __forceinline__ uint64_t slot_base(uint64_t g) { return (g >> 3) * C + (g & 7) * kSlot; }

// the 256-bit value in the slot at element `base`
__device__ __forceinline__ void slot_get(uint8_t* data, uint64_t base, uint32_t (&w)[8])
{
#pragma unroll
    for (unsigned q = 0; q < 8; ++q) {
        uint32_t v = 0;
#pragma unroll
        for (unsigned r = 0; r < 16; ++r)
            v |= (*spare_word(data, base + 16 * q + r) >> 30) << (2 * r);
        w[q] = v;
    }
}

// store w into the slot at element `base`: per 16 elements, all loads before the stores
__device__ __forceinline__ void slot_put(uint8_t* data, uint64_t base, const uint32_t (&w)[8])
{
#pragma unroll
    for (unsigned q = 0; q < 8; ++q) {
        uint32_t old[16];
#pragma unroll
        for (unsigned r = 0; r < 16; ++r)
            old[r] = *spare_word(data, base + 16 * q + r);
#pragma unroll
        for (unsigned r = 0; r < 16; ++r)
            *spare_word(data, base + 16 * q + r) = (old[r] & 0x3fffffffu) | (((w[q] >> (2 * r)) & 3u) << 30);
    }
}

// (1) group totals into the host slots
template <class P>
__global__ void __launch_bounds__(T) spare_total_kernel(uint8_t* __restrict__ data, uint64_t n)
{
    using F = F29<P>;
    __shared__ F red[T];
    __shared__ uint32_t word[8];
    const unsigned tid = threadIdx.x;
    const uint64_t base = (uint64_t)blockIdx.x * C;
    F acc = F::one();
#pragma unroll
    for (unsigned k = 0; k < E; ++k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        if (i < n) {
            const F a = load_masked<P>(data + 32 * i);
            if (!a.is_zero())
                acc = F::mul(acc, a);
        }
    }
    red[tid] = acc;
    for (unsigned d = T / 2; d > 0; d >>= 1) {
        __syncthreads();
        if (tid < d)
            red[tid] = F::mul(red[tid], red[tid + d]);
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t v[8];
        F::reduce_full(red[0]).to32(v);
#pragma unroll
        for (int q = 0; q < 8; ++q)
            word[q] = v[q];
    }
    __syncthreads();
    if (tid < kSlot) { // one element's two bits per thread
        const uint32_t bits = (word[tid >> 4] >> (2 * (tid & 15))) & 3u;
        uint32_t* sw = spare_word(data, slot_base(blockIdx.x) + tid);
        *sw = (*sw & 0x3fffffffu) | (bits << 30);
    }
}

// The totals of the middle step: in the elements' spare bits, or in a workspace array.
struct SpareTotals {
    uint8_t* data;
    __device__ void get(uint64_t g, uint32_t (&w)[8]) const { slot_get(data, slot_base(g), w); }
    __device__ void put(uint64_t g, const uint32_t (&w)[8]) const { slot_put(data, slot_base(g), w); }
};
struct PlainTotals {
    uint8_t* tot;
    __device__ void get(uint64_t g, uint32_t (&w)[8]) const
    {
        const uint4* q = (const uint4*)(tot + 32 * g);
        const uint4 a = q[0], b = q[1];
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w;
        w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
    }
    __device__ void put(uint64_t g, const uint32_t (&w)[8]) const
    {
        uint4* q = (uint4*)(tot + 32 * g);
        q[0] = make_uint4(w[0], w[1], w[2], w[3]);
        q[1] = make_uint4(w[4], w[5], w[6], w[7]);
    }
};

// (2) block b inverts totals [b kMidPer, (b + 1) kMidPer): each becomes t^-1 * 2^-10
// in the 2^261 scale, the single-group kernel's inv_total (the same value the
// batch inversion of the totals array gives)
template <class P, class A>
__global__ void __launch_bounds__(kMidThreads) totals_inverse_kernel(A acc_totals, uint64_t groups)
{
    using F = F29<P>;
    __shared__ F pre[kMidThreads];
    __shared__ F suf[kMidThreads];
    __shared__ F inv_all;
    const unsigned tid = threadIdx.x;
    const uint64_t g0 = (uint64_t)blockIdx.x * kMidPer + (uint64_t)tid * kMidE;
    F part[kMidE];
    F acc = F::one();
#pragma unroll
    for (unsigned k = 0; k < kMidE; ++k) {
        part[k] = acc;
        if (g0 + k < groups) {
            uint32_t w[8];
            acc_totals.get(g0 + k, w);
            acc = F::mul(acc, F::from32(w)); // a total is never zero
        }
    }
    pre[tid] = acc;
    suf[tid] = acc;
    for (unsigned d = 1; d < kMidThreads; d <<= 1) {
        __syncthreads();
        const F a = tid >= d ? pre[tid - d] : F::one();
        const F b = tid + d < kMidThreads ? suf[tid + d] : F::one();
        __syncthreads();
        if (tid >= d)
            pre[tid] = F::mul(a, pre[tid]);
        if (tid + d < kMidThreads)
            suf[tid] = F::mul(b, suf[tid]);
    }
    __syncthreads();
    if (tid == 0) {
        F k251 = F::zero();
        k251.l[8] = 1u << 19; // 2^251 = 2^(232 + 19)
        inv_all = F::mul(F::inverse_vt(pre[kMidThreads - 1]), k251);
    }
    __syncthreads();
    F inv = inv_all;
    if (tid > 0)
        inv = F::mul(inv, pre[tid - 1]);
    if (tid + 1 < kMidThreads)
        inv = F::mul(inv, suf[tid + 1]);
#pragma unroll
    for (int k = (int)kMidE - 1; k >= 0; --k) {
        if (g0 + k < groups) {
            uint32_t w[8], o[8];
            acc_totals.get(g0 + k, w);
            F::reduce_full(F::mul(inv, part[k])).to32(o);
            inv = F::mul(inv, F::from32(w));
            acc_totals.put(g0 + k, o);
        }
    }
}

// (3) groups [g0, g0 + gridDim.x): inverse total from the host slot, then as the
// single-group kernel (outputs are canonical, so their spare bits end up zero)
template <class P>
__global__ void __launch_bounds__(T) spare_finish_kernel(uint8_t* __restrict__ data, uint64_t n, uint64_t g0)
{
    using F = F29<P>;
    __shared__ F node[2 * T];
    __shared__ uint32_t word[8];
    const unsigned tid = threadIdx.x;
    const uint64_t grp = g0 + blockIdx.x;
    const uint64_t base = grp * C;
    if (tid < 8)
        word[tid] = 0;
    __syncthreads();
    if (tid < kSlot)
        atomicOr(&word[tid >> 4], (*spare_word(data, slot_base(grp) + tid) >> 30) << (2 * (tid & 15)));
    F part[E];
    F acc = F::one();
#pragma unroll
    for (unsigned k = 0; k < E; ++k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        part[k] = acc;
        if (i < n) {
            const F a = load_masked<P>(data + 32 * i);
            if (!a.is_zero())
                acc = F::mul(acc, a);
        }
    }
    node[T + tid] = acc;
    tree_up(node, tid); // its first barrier also completes word[]
    if (tid == 0) {
        uint32_t w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            w[q] = word[q];
        node[1] = F::from32(w);
    }
    __syncthreads();
    F inv = tree_down(node, tid);
#pragma unroll
    for (int k = (int)E - 1; k >= 0; --k) {
        const uint64_t i = base + (uint64_t)k * T + tid;
        if (i < n) {
            const F a = load_masked<P>(data + 32 * i);
            if (!a.is_zero()) {
                F::reduce_full(F::mul(inv, part[k])).store32(data + 32 * i);
                inv = F::mul(inv, a);
            } else {
                F::zero().store32(data + 32 * i); // clears spare bits a slot left on a zero
            }
        }
    }
}

template <class F>
struct Form29;
template <>
struct Form29<Fr> {
    using P = Fr29P;
};
template <>
struct Form29<Fq> {
    using P = Fq29P;
};

// bytes of workspace batch_invert uses for n elements (0 for one group)
uint64_t batch_invert_ws_bytes(uint64_t n)
{
    const uint64_t groups = (n + C - 1) / C;
    return groups > 1 ? align_up(32 * groups, 256) : 0;
}

// In place. With `ws` of batch_invert_ws_bytes(n) bytes the totals sit in ws;
// without, they travel in the elements' spare bits.
template <class F>
RustError batch_invert(F* data, uint64_t n, hipStream_t stream, void* ws = nullptr, uint64_t ws_bytes = 0)
{
    using P = typename Form29<F>::P;
    if (n == 0)
        return ok();
    const uint64_t groups = (n + C - 1) / C;
    H2_CHECK_ARG(groups < (1ull << 31), "batch invert of %lu elements is too large", (unsigned long)n);
    if (groups == 1) {
        batch_invert_kernel<P><<<1, T, 0, stream>>>((uint8_t*)data, n);
        H2_LAUNCH_CHECK();
        return ok();
    }
    if (ws == nullptr || ws_bytes < 32 * groups) {
        uint8_t* d = (uint8_t*)data;
        spare_total_kernel<P><<<(unsigned)groups, T, 0, stream>>>(d, n);
        H2_LAUNCH_CHECK();
        totals_inverse_kernel<P><<<(unsigned)((groups + kMidPer - 1) / kMidPer), kMidThreads, 0, stream>>>(
            SpareTotals { d }, groups);
        H2_LAUNCH_CHECK();
        // hosts after the groups they host: [ceil(g/8), g), then [ceil(g/64), ceil(g/8)), ...
        for (uint64_t hi = groups; hi > 0;) {
            const uint64_t lo = hi > 1 ? (hi + 7) / 8 : 0;
            spare_finish_kernel<P><<<(unsigned)(hi - lo), T, 0, stream>>>(d, n, lo);
            H2_LAUNCH_CHECK();
            hi = lo;
        }
        return ok();
    }
    uint8_t* tot = (uint8_t*)ws;
    group_total_kernel<P><<<(unsigned)groups, T, 0, stream>>>((const uint8_t*)data, n, tot);
    H2_LAUNCH_CHECK();
    totals_inverse_kernel<P><<<(unsigned)((groups + kMidPer - 1) / kMidPer), kMidThreads, 0, stream>>>(
        PlainTotals { tot }, groups);
    H2_LAUNCH_CHECK();
    group_finish_kernel<P><<<(unsigned)groups, T, 0, stream>>>((uint8_t*)data, n, tot);
    H2_LAUNCH_CHECK();
    return ok();
}

// ------------------------------------------------------------------- scan
// Strided view A[t] = base[t * s], t < m.

// The scan runs on the 29-bit form (field29.h). Running products stay in the
// 2^261 scale (every loaded word x * 2^256 enters as shl5, x * 2^261), the exclusive
// prefix of the thread totals comes from the product tree in LDS (up-sweep, then the
// prefixes pushed down: a left child inherits its parent's prefix, a right child the
// parent's prefix times its left sibling), and the output product v * p takes p in
// the 2^256 scale (the tree prefix times the caller's prefix word, or times 2^256
// mod r), so it lands back in the stored scale.

// 2^256 mod r in 29-bit limbs: mul(v * 2^261, ONE256) = v * 2^256
__device__ __forceinline__ R29 one256()
{
    constexpr uint32_t k[9] = { 0x0ffffffbu, 0x04b1a0e2u, 0x18334a6bu, 0x18ed2b3eu, 0x1462e36fu,
        0x11b7bc3cu, 0x1cbd99bau, 0x183340fbu, 0x000e0a77u };
    R29 r;
#pragma unroll
    for (int i = 0; i < 9; ++i)
        r.l[i] = k[i];
    return r;
}

// The caller has replaced node[1] by the identity of the prefixes (and
// synchronised); returns the product of the leaves before leaf tid.
template <class F>
__device__ __forceinline__ F tree_down_prefix(F* node, unsigned tid)
{
    for (unsigned half = 1; half < T; half <<= 1) {
        if (tid < half) {
            const unsigned i = half + tid;
            const F p = node[i];
            const F left = node[2 * i];
            node[2 * i] = p;
            node[2 * i + 1] = F::mul(p, left);
        }
        __syncthreads();
    }
    return node[T + tid];
}

// Local inclusive scan of each chunk of C view elements. With `prefix` non-null
// (single-chunk top level) every element is also multiplied by *prefix.
__global__ void __launch_bounds__(T) scan_local_kernel(Fr* __restrict__ base, uint64_t s, uint64_t m,
    const Fr* __restrict__ prefix)
{
    __shared__ R29 node[2 * T];
    const unsigned tid = threadIdx.x;
    const uint64_t c0 = (uint64_t)blockIdx.x * C + (uint64_t)tid * E;
    R29 v[E];
    R29 acc = R29::one(); // 2^261 scale
#pragma unroll
    for (unsigned k = 0; k < E; ++k) {
        const uint64_t t = c0 + k;
        if (t < m)
            acc = R29::mul(acc, R29::shl5(ld29(base + t * s)));
        v[k] = acc;
    }
    node[T + tid] = acc;
    tree_up(node, tid);
    if (tid == 0)
        node[1] = R29::one();
    __syncthreads();
    const R29 pre = tree_down_prefix(node, tid); // 2^261 scale
    const R29 p = R29::mul(pre, prefix ? ld29(prefix) : one256()); // 2^256 scale
#pragma unroll
    for (unsigned k = 0; k < E; ++k) {
        const uint64_t t = c0 + k;
        if (t < m)
            st29(R29::mul(v[k], p), base + t * s);
    }
}

// Chunk u >= 1 is multiplied by the scanned chunk product A[u*C - 1]; chunk 0
// by *prefix. The last element of every non-final chunk is already final.
__global__ void __launch_bounds__(T) scan_fixup_kernel(Fr* __restrict__ base, uint64_t s, uint64_t m,
    const Fr* __restrict__ prefix)
{
    const uint64_t u = blockIdx.x;
    const uint64_t nchunks = (m + C - 1) / C;
    const R29 f = ld29(u == 0 ? prefix : base + (u * C - 1) * s);
    const uint64_t first = u * C;
    uint64_t last = first + C; // exclusive
    if (u + 1 < nchunks)
        last -= 1; // skip the chunk product, scanned at the level above
    if (last > m)
        last = m;
    for (uint64_t t = first + threadIdx.x; t < last; t += T)
        st29(R29::mul(R29::shl5(ld29(base + t * s)), f), base + t * s);
}

RustError scan_view(Fr* base, uint64_t s, uint64_t m, const Fr* prefix, hipStream_t stream)
{
    if (m == 0)
        return ok();
    if (m <= C) {
        scan_local_kernel<<<1, T, 0, stream>>>(base, s, m, prefix);
        H2_LAUNCH_CHECK();
        return ok();
    }
    const uint64_t nchunks = (m + C - 1) / C;
    scan_local_kernel<<<(unsigned)nchunks, T, 0, stream>>>(base, s, m, nullptr);
    H2_LAUNCH_CHECK();
    // chunk products of chunks 0..nchunks-2 sit at view positions u*C + C - 1
    RustError e = scan_view(base + (C - 1) * s, s * C, nchunks - 1, prefix, stream);
    if (e.code)
        return e;
    scan_fixup_kernel<<<(unsigned)nchunks, T, 0, stream>>>(base, s, m, prefix);
    H2_LAUNCH_CHECK();
    return ok();
}

// ------------------------------------------------- permutation / lookup

constexpr unsigned kEwBlock = 256;
unsigned ew_grid(uint64_t n) { return grid_for(n, kEwBlock, 256u * 16u); }

// The element-wise kernels below run on the 29-bit form: stored words (x * 2^256)
// load as normalised limbs; the product of two stored words is
// mul(shl5(a), b) (one operand moved to the 2^261 scale by a 5-bit shift, which is
// cheaper than a product); sums of two canonical words (< 2m, limbs < 2^30) enter a
// product as they are; results are reduced to canonical words before they are stored.
// omega^i = lo[i mod 2^kSplit] * hi[i >> kSplit]: the two tables (2^kSplit and
// ceil(n / 2^kSplit) entries) stay in L2, so no thread raises omega to a power
constexpr uint32_t kSplit = 11;

uint64_t split_hi(uint64_t n) { return (n + (1ull << kSplit) - 1) >> kSplit; }

uint64_t split_bytes(uint64_t n) { return align_up(32ull << kSplit, 256) + align_up(32 * split_hi(n), 256); }

__global__ void split_powers_kernel(Fr* __restrict__ lo, Fr* __restrict__ hi, const Fr* __restrict__ d_omega,
    uint64_t nhi)
{
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const Fr w = Fr::load(d_omega);
    if (j < (1u << kSplit))
        Fr::pow_u64(w, j).store(lo + j);
    if (j < nhi)
        Fr::pow_u64(w, j << kSplit).store(hi + j);
}

__global__ void mul_into_kernel(Fr* __restrict__ a, const Fr* __restrict__ b, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        st29(R29::mul(R29::shl5(ld29(a + i)), ld29(b + i)), a + i);
}

// (A' + beta)(S' + gamma)                             (commit_product.h:12-38)
__global__ void lookup_denominator_kernel(Fr* __restrict__ out, const Fr* __restrict__ pi, const Fr* __restrict__ pt,
    const Fr* __restrict__ d_beta, const Fr* __restrict__ d_gamma, uint64_t n)
{
    const R29 beta32 = R29::shl5(ld29(d_beta)), gamma = ld29(d_gamma);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        // 32 (A' + beta) < 64m and S' + gamma < 2m: both limb sets < 2^30
        const R29 x = R29::add(R29::shl5(ld29(pi + i)), beta32);
        st29(R29::mul(x, R29::add(ld29(pt + i), gamma)), out + i);
    }
}

// out *= (A + beta)(S + gamma)                        (commit_product.h:40-66)
__global__ void lookup_numerator_kernel(Fr* __restrict__ out, const Fr* __restrict__ ci, const Fr* __restrict__ ct,
    const Fr* __restrict__ d_beta, const Fr* __restrict__ d_gamma, uint64_t n)
{
    const R29 beta = ld29(d_beta), gamma = ld29(d_gamma);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const R29 v = R29::mul(R29::shl5(ld29(out + i)), R29::add(ld29(ci + i), beta)); // < 1.01m, normalised
        st29(R29::mul(R29::shl5(v), R29::add(ld29(ct + i), gamma)), out + i);
    }
}

// Up to kPermCols columns in one pass (permutation_commit.h:12-75, which runs one
// denominator and one numerator kernel per column): per element, for each column j in
// order, den *= beta perm_j + gamma + value_j and num (*)= deltaomega_j omega^i beta +
// gamma + value_j with deltaomega_j = deltaomega delta^j, every product reduced to its
// canonical word as a per-column pass would store it; den and num are read and written
// once per launch instead of once per column, and omega^i is formed once.
constexpr unsigned kPermCols = 4;
struct PermCols {
    const Fr* perm[kPermCols];
    const Fr* val[kPermCols];
};

__global__ void __launch_bounds__(kEwBlock) perm_columns_kernel(Fr* __restrict__ den, Fr* __restrict__ num,
    const PermCols cols, unsigned ncols, const Fr* __restrict__ d_beta, const Fr* __restrict__ d_gamma,
    const Fr* __restrict__ d_delta, const Fr* __restrict__ lo, const Fr* __restrict__ hi,
    const Fr* __restrict__ d_deltaomega, uint64_t n, int first)
{
    const R29 beta = ld29(d_beta), beta32 = R29::shl5(beta), gamma = ld29(d_gamma);
    // (deltaomega_j beta) in the 2^261 scale; deltaomega_(j+1) the canonical word of
    // deltaomega_j delta (what the per-column path's scalar product stores)
    R29 c32[kPermCols];
    {
        const R29 delta32 = R29::shl5(ld29(d_delta));
        R29 dw = ld29(d_deltaomega);
#pragma unroll
        for (unsigned j = 0; j < kPermCols; ++j) {
            c32[j] = R29::shl5(R29::mul(R29::shl5(dw), beta));
            dw = R29::reduce_full(R29::mul(dw, delta32));
        }
    }
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const R29 w = R29::mul(R29::shl5(ld29(lo + (i & ((1u << kSplit) - 1)))), ld29(hi + (i >> kSplit)));
        R29 d = ld29(den + i);
        R29 nm = first ? R29::zero() : ld29(num + i);
#pragma unroll
        for (unsigned j = 0; j < kPermCols; ++j) {
            if (j < ncols) {
                const R29 v = ld29(cols.val[j] + i);
                // beta * perm (normalised, < 1.01m) + gamma + value: < 3.01m, limbs < 2^31
                const R29 td = R29::add(R29::add(R29::mul(beta32, ld29(cols.perm[j] + i)), gamma), v);
                d = R29::reduce_full(R29::normalize(R29::mul(R29::shl5(d), td)));
                const R29 tn = R29::add(R29::add(R29::mul(c32[j], w), gamma), v); // < 3.01m
                nm = R29::reduce_full(R29::normalize(first && j == 0 ? tn : R29::mul(R29::shl5(nm), tn)));
            }
        }
        d.store32(den + i);
        nm.store32(num + i);
    }
}

// deltaomega *= delta, k times (the per-column scalar product, in one launch)
__global__ void advance_kernel(Fr* __restrict__ a, const Fr* __restrict__ b, unsigned k)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        Fr x = Fr::load(a);
        const Fr y = Fr::load(b);
        for (unsigned i = 0; i < k; ++i)
            x = x * y;
        x.store(a);
    }
}

// Folds `batch` columns into the denominator and numerator accumulators; the
// first column of a call with `init_num` initialises the numerator.
// `lo`/`hi` hold the split omega powers (split_powers_kernel) for this call.
RustError permutation_product_columns(Fr* den, Fr* num, const FFITraitObject* perms, const FFITraitObject* values,
    const Fr* d_beta, const Fr* d_gamma, const Fr* d_delta, const Fr* lo, const Fr* hi, Fr* d_deltaomega, uint64_t n,
    uint64_t offset, uint64_t batch, bool init_num, hipStream_t stream)
{
    for (uint64_t c0 = 0; c0 < batch; c0 += kPermCols) {
        const unsigned m = batch - c0 < kPermCols ? (unsigned)(batch - c0) : kPermCols;
        PermCols cols;
        for (unsigned j = 0; j < kPermCols; ++j) {
            cols.perm[j] = j < m ? (const Fr*)perms[c0 + j].ptr + offset : nullptr;
            cols.val[j] = j < m ? (const Fr*)values[c0 + j].ptr + offset : nullptr;
        }
        perm_columns_kernel<<<ew_grid(n), kEwBlock, 0, stream>>>(den, num, cols, m, d_beta, d_gamma, d_delta, lo, hi,
            d_deltaomega, n, init_num && c0 == 0);
        H2_LAUNCH_CHECK();
        advance_kernel<<<1, 64, 0, stream>>>(d_deltaomega, d_delta, m);
        H2_LAUNCH_CHECK();
    }
    return ok();
}

// den <- num / den; `inv_ws` holds batch_invert_ws_bytes(n) bytes (or is null)
RustError permutation_product_finish(Fr* den, Fr* num, uint64_t n, void* inv_ws, hipStream_t stream)
{
    RustError e = batch_invert<Fr>(den, n, stream, inv_ws, batch_invert_ws_bytes(n));
    if (e.code)
        return e;
    mul_into_kernel<<<ew_grid(n), kEwBlock, 0, stream>>>(den, num, n);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError split_powers(Fr* lo, Fr* hi, const Fr* d_omega, uint64_t n, hipStream_t stream)
{
    const uint64_t nhi = split_hi(n);
    const uint64_t m = nhi > (1u << kSplit) ? nhi : (1u << kSplit);
    split_powers_kernel<<<(unsigned)((m + 255) / 256), 256, 0, stream>>>(lo, hi, d_omega, nhi);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError permutation_product_device(Fr* den, Fr* num, Fr* lo, Fr* hi, void* inv_ws, const FFITraitObject* perms,
    const FFITraitObject* values, const Fr* d_beta, const Fr* d_gamma, const Fr* d_delta, const Fr* d_omega,
    Fr* d_deltaomega, uint64_t n, uint64_t offset, uint64_t batch, hipStream_t stream)
{
    H2_CHECK_ARG(batch > 0, "permutation product needs at least one column");
    RustError e = split_powers(lo, hi, d_omega, n, stream);
    if (e.code)
        return e;
    e = permutation_product_columns(den, num, perms, values, d_beta, d_gamma, d_delta, lo, hi, d_deltaomega, n,
        offset, batch, true, stream);
    if (e.code)
        return e;
    return permutation_product_finish(den, num, n, inv_ws, stream);
}

} // namespace

extern "C" {

RustError _halo2_batch_invert(void* d_data, uint32_t field_type, size_t length, void* stream)
{
    if (field_type != 0 && field_type != 1)
        return fail((int)hipErrorInvalidValue, "Invalid scalar type");
    if (length == 0)
        return fail((int)hipErrorInvalidValue, "_halo2_batch_invert: length must be >= 1");
    if (field_type == 0)
        return batch_invert<Fr>((Fr*)d_data, length, (hipStream_t)stream);
    return batch_invert<Fq>((Fq*)d_data, length, (hipStream_t)stream);
}

// ------------------------------------------------------- grand product

size_t _halo2_grand_product_max_len(size_t length, uint64_t free_bytes)
{
    for (uint64_t l = length; l > 0; l >>= 1)
        if (32 * l + 32 < free_bytes)
            return l;
    return 0;
}

uint64_t _halo2_grand_product_workspace_size(uint64_t poly_length) { return align_up(32 * poly_length + 32, 256) + 256; }

RustError _halo2_grand_product(const FFITraitObject* output, const FFITraitObject* input, const FFITraitObject* prefix,
    size_t length, size_t offset, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (length == 0)
        return ok();
    Workspace ws(scratch, scratch_bytes);
    Fr* d = (Fr*)ws.take(32 * length);
    Fr* p = (Fr*)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_grand_product: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    H2_TRY(hipMemcpyAsync(d, (const uint8_t*)input->ptr + 32 * offset, 32 * length, hipMemcpyHostToDevice, stream));
    H2_TRY(hipMemcpyAsync(p, prefix->ptr, 32, hipMemcpyHostToDevice, stream));
    RustError e = scan_view(d, 1, length, p, stream);
    if (e.code)
        return e;
    H2_TRY(hipMemcpyAsync((uint8_t*)output->ptr + 32 * offset, d, 32 * length, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

RustError _halo2_grand_product_device_inputs(const FFITraitObject* d_inout, const FFITraitObject* d_prefix,
    size_t length, size_t offset, void* stream)
{
    if (length == 0)
        return ok();
    return scan_view((Fr*)d_inout->ptr + offset, 1, length, (const Fr*)d_prefix->ptr, (hipStream_t)stream);
}

// ------------------------------------------------- permutation product

size_t _halo2_permutation_product_max_len(size_t length, uint64_t free_bytes)
{
    for (uint64_t l = length; l > 0; l >>= 1)
        if (_halo2_permutation_product_workspace_size(l) < free_bytes)
            return l;
    return 0;
}

uint64_t _halo2_permutation_product_workspace_size(uint64_t poly_length)
{
    // denominators, numerators, one staged (permutation, value) column pair, omega tables,
    // group totals of the inversion
    return 4 * align_up(32 * poly_length, 256) + split_bytes(poly_length) + batch_invert_ws_bytes(poly_length) + 512;
}

RustError _halo2_permutation_product(const FFITraitObject* modified_values, const FFITraitObject* permutations,
    const FFITraitObject* values, const void* beta_device, const void* gamma_device, const void* delta_device,
    const void* omega_device, void* deltaomega_device, size_t length, size_t offset, size_t batch, void* scratch,
    uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (length == 0)
        return ok();
    Workspace ws(scratch, scratch_bytes);
    Fr* den = (Fr*)ws.take(32 * length);
    Fr* num = (Fr*)ws.take(32 * length);
    Fr* perm = (Fr*)ws.take(32 * length);
    Fr* val = (Fr*)ws.take(32 * length);
    Fr* lo = (Fr*)ws.take(32ull << kSplit);
    Fr* hi = (Fr*)ws.take(32 * split_hi(length));
    void* inv_ws = ws.take(batch_invert_ws_bytes(length));
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_permutation_product: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    H2_CHECK_ARG(batch > 0, "_halo2_permutation_product: batch must be >= 1");
    {
        RustError e = split_powers(lo, hi, (const Fr*)omega_device, length, stream);
        if (e.code)
            return e;
    }
    // the running accumulator is the caller's chunk buffer itself (permutation_commit.cu:84)
    H2_TRY(hipMemcpyAsync(den, modified_values->ptr, 32 * length, hipMemcpyHostToDevice, stream));
    for (size_t c = 0; c < batch; ++c) {
        H2_TRY(hipMemcpyAsync(perm, (const uint8_t*)permutations[c].ptr + 32 * offset, 32 * length,
            hipMemcpyHostToDevice, stream));
        H2_TRY(hipMemcpyAsync(val, (const uint8_t*)values[c].ptr + 32 * offset, 32 * length, hipMemcpyHostToDevice,
            stream));
        FFITraitObject p { (uint64_t*)perm }, v { (uint64_t*)val };
        // one column at a time through the device routine (offset already applied)
        RustError e = permutation_product_columns(den, num, &p, &v, (const Fr*)beta_device, (const Fr*)gamma_device,
            (const Fr*)delta_device, lo, hi, (Fr*)deltaomega_device, length, 0, 1, c == 0, stream);
        if (e.code)
            return e;
    }
    RustError e = permutation_product_finish(den, num, length, inv_ws, stream);
    if (e.code)
        return e;
    H2_TRY(hipMemcpyAsync(modified_values->ptr, den, 32 * length, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

uint64_t _halo2_permutation_product_device_inputs_workspace_size(uint64_t poly_length)
{
    return align_up(32 * poly_length, 256) + split_bytes(poly_length) + batch_invert_ws_bytes(poly_length) + 512;
}

RustError _halo2_permutation_product_device_inputs(const FFITraitObject* modified_values_device,
    const FFITraitObject* permutations_device, const FFITraitObject* values_device, const void* beta_device,
    const void* gamma_device, const void* delta_device, const void* omega_device, void* deltaomega_device,
    size_t length, size_t offset, size_t batch, void* scratch, uint64_t scratch_bytes, void* stream)
{
    if (length == 0)
        return ok();
    Workspace ws(scratch, scratch_bytes);
    Fr* num = (Fr*)ws.take(32 * length);
    Fr* lo = (Fr*)ws.take(32ull << kSplit);
    Fr* hi = (Fr*)ws.take(32 * split_hi(length));
    void* inv_ws = ws.take(batch_invert_ws_bytes(length));
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue,
            "_halo2_permutation_product_device_inputs: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    return permutation_product_device((Fr*)modified_values_device->ptr, num, lo, hi, inv_ws, permutations_device,
        values_device,
        (const Fr*)beta_device, (const Fr*)gamma_device, (const Fr*)delta_device, (const Fr*)omega_device,
        (Fr*)deltaomega_device, length, offset, batch, (hipStream_t)stream);
}

// ------------------------------------------------------ lookup product

size_t _halo2_commit_product_max_len(size_t length, uint64_t free_bytes)
{
    for (uint64_t l = length; l > 0; l >>= 1)
        if (_halo2_commit_product_workspace_size(l) < free_bytes)
            return l;
    return 0;
}

uint64_t _halo2_commit_product_workspace_size(uint64_t poly_length)
{
    return 5 * align_up(32 * poly_length, 256) + batch_invert_ws_bytes(poly_length) + 512;
}

// `inv_ws` holds batch_invert_ws_bytes(n) bytes, or is null (device-input entry point:
// no workspace argument, single-kernel inversion)
static RustError commit_product_device(Fr* lp, const Fr* pi, const Fr* pt, const Fr* ci, const Fr* ct, const Fr* beta,
    const Fr* gamma, uint64_t n, void* inv_ws, hipStream_t stream)
{
    lookup_denominator_kernel<<<ew_grid(n), kEwBlock, 0, stream>>>(lp, pi, pt, beta, gamma, n);
    H2_LAUNCH_CHECK();
    RustError e = batch_invert<Fr>(lp, n, stream, inv_ws, inv_ws ? batch_invert_ws_bytes(n) : 0);
    if (e.code)
        return e;
    lookup_numerator_kernel<<<ew_grid(n), kEwBlock, 0, stream>>>(lp, ci, ct, beta, gamma, n);
    H2_LAUNCH_CHECK();
    return ok();
}

RustError _halo2_commit_product(const FFITraitObject* lookup_product, const FFITraitObject* permuted_input,
    const FFITraitObject* permuted_table, const FFITraitObject* compressed_input,
    const FFITraitObject* compressed_table, const void* beta_device, const void* gamma_device, size_t length,
    size_t offset, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (length == 0)
        return ok();
    Workspace ws(scratch, scratch_bytes);
    Fr* d[5];
    for (int i = 0; i < 5; ++i)
        d[i] = (Fr*)ws.take(32 * length);
    void* inv_ws = ws.take(batch_invert_ws_bytes(length));
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_commit_product: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    const FFITraitObject* src[4] = { permuted_input, permuted_table, compressed_input, compressed_table };
    for (int i = 0; i < 4; ++i)
        H2_TRY(hipMemcpyAsync(d[i + 1], (const uint8_t*)src[i]->ptr + 32 * offset, 32 * length, hipMemcpyHostToDevice,
            stream));
    RustError e = commit_product_device(d[0], d[1], d[2], d[3], d[4], (const Fr*)beta_device, (const Fr*)gamma_device,
        length, inv_ws, stream);
    if (e.code)
        return e;
    H2_TRY(hipMemcpyAsync((uint8_t*)lookup_product->ptr + 32 * offset, d[0], 32 * length, hipMemcpyDeviceToHost,
        stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

RustError _halo2_commit_product_device_inputs(const FFITraitObject* d_lookup_product,
    const FFITraitObject* d_permuted_input, const FFITraitObject* d_permuted_table,
    const FFITraitObject* d_compressed_input, const FFITraitObject* d_compressed_table, const void* beta_device,
    const void* gamma_device, size_t length, size_t offset, void* stream)
{
    if (length == 0)
        return ok();
    return commit_product_device((Fr*)d_lookup_product->ptr + offset, (const Fr*)d_permuted_input->ptr + offset,
        (const Fr*)d_permuted_table->ptr + offset, (const Fr*)d_compressed_input->ptr + offset,
        (const Fr*)d_compressed_table->ptr + offset, (const Fr*)beta_device, (const Fr*)gamma_device, length, nullptr,
        (hipStream_t)stream);
}

} // extern "C"
