// BN254 G1 multi-scalar multiplication (Pippenger) for gfx950, behind the
// reference's `_halo2_multiexp*` entry points (cuda/src/multiexp.cu:166-525).
//
// Pipeline (all on one stream, caller-owned workspace, no host round trip):
//  0. bases: copied into the workspace in the 2^261 Montgomery scale of the
//     29-bit-limb field form (curve29.h), in which every bucket operation
//     below runs; the result is converted back to the 2^256 scale.
//  1. scalars: Montgomery -> canonical (one Fr multiplication by 1).
//  2. signed c-bit digits (carry-propagated, |d| <= 2^(c-1)), W = ceil(255/c)
//     windows, B = 2^(c-1) buckets per window; a zero digit adds nothing.
//  3. sort of the (point index, sign) entries by bucket in two counting
//     rounds whose ranks come from LDS atomics only:
//       round 1 bins entries by (window, top bucket bits) per tile of points,
//       round 2 splits each bin by the low FB bucket bits, chunk by chunk.
//     Per-tile / per-chunk histograms are laid out so that one linear
//     exclusive scan yields every entry's destination, whatever the skew of
//     the scalars (a bucket holding every point costs the same sort work).
//  4. bucket sums with load balancing: every bucket is cut into pieces of at
//     most S0 entries, one thread per piece accumulates XYZZ mixed additions.
//     A bucket with one piece is final at once; 2..LIGHT pieces are summed by
//     one thread; more pieces (dense buckets) get a work-group tree reduction.
//  5. per window, sum_b b * B_b with running sums over segments of L1
//     buckets, then levels of Lg segments: each level carries pairs (T, S)
//     with  answer = sum T + sum_m m * S_m,  so no per-segment scalar
//     multiplication is needed (the weights become log2 L doublings of S).
//  6. Horner over windows (c doublings per window), conversion to affine.
// The result is written as (x, y, 1) or the identity (0, 1, 0) in Montgomery
// form, the layout of the reference's `cukernel_to_affine` output.
#include <memory>

#include "common.h"
#include "curve29.h"
#include "profile.h"
#include "scan_u32.h"

using namespace h2;

namespace {

constexpr uint32_t SCALAR_BITS = 254;
constexpr uint32_t MAX_FB = 11; // low bucket bits sorted in round 2 (LDS histogram of 2^11)
constexpr uint32_t R1_THREADS = 1024;
constexpr uint32_t R1_PER = 4;
constexpr uint32_t R1_TILE = R1_THREADS * R1_PER; // points per round-1 tile
constexpr uint32_t R2_THREADS = 256;
constexpr uint32_t R2_CHUNK = 16384; // entries per round-2 chunk
constexpr uint32_t LIGHT = 32; // pieces summed by one thread
constexpr uint32_t HEAVY_THREADS = 256;
constexpr uint32_t HEAVY_BLOCKS = 512;
constexpr uint32_t RED_L1 = 8; // buckets per first-level reduction segment
constexpr uint32_t RED_LG = 16; // segments per group at later levels
constexpr uint32_t MAX_RED_LEVELS = 12;

struct MsmPlan {
    uint64_t n;
    uint32_t c, W, B;
    uint64_t nb; // W * B buckets
    uint32_t FB, CB; // round-2 (fine) and round-1 (coarse) bucket bits, FB + CB = c - 1
    uint32_t NC; // coarse bins = W << CB
    uint32_t ntiles; // round-1 tiles
    uint64_t max_chunks; // upper bound on round-2 chunks
    uint32_t S0; // entries per level-0 piece
    uint64_t ub0; // upper bound on pieces
    uint32_t L1; // buckets per first reduction segment
    uint32_t nlev; // reduction levels after the first
    uint64_t Mlev[MAX_RED_LEVELS + 1]; // segments per window: Mlev[0] = B / L1, Mlev[i] after level i
    uint32_t Llev[MAX_RED_LEVELS + 1]; // group size of level i (i >= 1)
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
    p.FB = p.c - 1 < MAX_FB ? p.c - 1 : MAX_FB;
    p.CB = p.c - 1 - p.FB;
    p.NC = p.W << p.CB;
    p.ntiles = (uint32_t)((n + R1_TILE - 1) / R1_TILE);
    const uint64_t work = n * p.W;
    p.max_chunks = work / R2_CHUNK + p.NC;
    uint64_t s0 = (work + (1ull << 21) - 1) >> 21;
    if (s0 < 4)
        s0 = 4;
    p.S0 = (uint32_t)s0;
    p.ub0 = work / p.S0 + p.nb;
    p.L1 = p.B < RED_L1 ? p.B : RED_L1;
    p.Mlev[0] = p.B / p.L1;
    uint32_t lev = 0;
    while (p.Mlev[lev] > 1) {
        const uint32_t lg = p.Mlev[lev] < RED_LG ? (uint32_t)p.Mlev[lev] : RED_LG;
        ++lev;
        p.Llev[lev] = lg;
        p.Mlev[lev] = p.Mlev[lev - 1] / lg;
    }
    p.nlev = lev;
    return p;
}

struct MsmBuffers {
    uint8_t* pts; // n affine bases in the 2^261 scale (64 B each)
    uint32_t* canon; // n x 8 limbs
    uint32_t* h1; // NC * ntiles
    uint32_t* h1off; // NC * ntiles + 1
    uint32_t* val1; // n * W (round-1 output: index | sign << 31)
    uint16_t* key1; // n * W (round-1 output: low FB bucket bits)
    uint32_t* bin_start; // NC + 1
    uint32_t* chunk_base; // NC + 1
    uint32_t* h2; // max_chunks * 2^FB
    uint32_t* h2off; // max_chunks * 2^FB + 1
    uint32_t* entries; // n * W, sorted by bucket
    uint32_t* off0; // nb + 1
    uint32_t* pc; // nb
    uint32_t* poff; // nb + 1
    uint32_t* piece_bucket; // ub0
    uint32_t* heavy; // nb heavy bucket ids + 1 counter
    G1Xyzz* part; // ub0
    G1Xyzz* bsum; // nb
    G1Xyzz* red[4]; // (T, S) ping-pong: W * Mlev[0] each
    void* scan_scratch;
    uint64_t* out; // 12 limbs
};

template <class WS>
void carve(WS& ws, const MsmPlan& p, MsmBuffers* b)
{
    const uint64_t nh1 = (uint64_t)p.NC * p.ntiles;
    const uint64_t nh2 = p.max_chunks << p.FB;
    uint64_t scan_n = nh1 + 1;
    if (nh2 + 1 > scan_n)
        scan_n = nh2 + 1;
    if (p.nb + 1 > scan_n)
        scan_n = p.nb + 1;
    const uint64_t nred = (uint64_t)p.W * p.Mlev[0];
    void* v[24];
    int k = 0;
    v[k++] = ws.take(64 * p.n);
    v[k++] = ws.take(32 * p.n);
    v[k++] = ws.take(4 * nh1);
    v[k++] = ws.take(4 * (nh1 + 1));
    v[k++] = ws.take(4 * p.n * p.W);
    v[k++] = ws.take(2 * p.n * p.W);
    v[k++] = ws.take(4 * (p.NC + 1));
    v[k++] = ws.take(4 * (p.NC + 1));
    v[k++] = ws.take(4 * nh2);
    v[k++] = ws.take(4 * (nh2 + 1));
    v[k++] = ws.take(4 * p.n * p.W);
    v[k++] = ws.take(4 * (p.nb + 1));
    v[k++] = ws.take(4 * p.nb);
    v[k++] = ws.take(4 * (p.nb + 1));
    v[k++] = ws.take(4 * p.ub0);
    v[k++] = ws.take(4 * (p.nb + 1));
    v[k++] = ws.take(128 * p.ub0);
    v[k++] = ws.take(128 * p.nb);
    for (int i = 0; i < 4; ++i)
        v[k++] = ws.take(128 * nred);
    v[k++] = ws.take(scan_u32_scratch(scan_n));
    v[k++] = ws.take(96);
    if (b) {
        k = 0;
        b->pts = (uint8_t*)v[k++];
        b->canon = (uint32_t*)v[k++];
        b->h1 = (uint32_t*)v[k++];
        b->h1off = (uint32_t*)v[k++];
        b->val1 = (uint32_t*)v[k++];
        b->key1 = (uint16_t*)v[k++];
        b->bin_start = (uint32_t*)v[k++];
        b->chunk_base = (uint32_t*)v[k++];
        b->h2 = (uint32_t*)v[k++];
        b->h2off = (uint32_t*)v[k++];
        b->entries = (uint32_t*)v[k++];
        b->off0 = (uint32_t*)v[k++];
        b->pc = (uint32_t*)v[k++];
        b->poff = (uint32_t*)v[k++];
        b->piece_bucket = (uint32_t*)v[k++];
        b->heavy = (uint32_t*)v[k++];
        b->part = (G1Xyzz*)v[k++];
        b->bsum = (G1Xyzz*)v[k++];
        for (int i = 0; i < 4; ++i)
            b->red[i] = (G1Xyzz*)v[k++];
        b->scan_scratch = v[k++];
        b->out = (uint64_t*)v[k++];
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

struct SortParams {
    uint64_t n;
    uint32_t c, W, FB, CB, NC, ntiles;
};

// Round 1, pass 1: per tile of R1_TILE points, LDS histogram over the coarse
// bins (window, top CB bucket bits); stored bin-major so the linear scan
// orders entries by bin, then by tile.
__global__ void __launch_bounds__(R1_THREADS) r1_count_kernel(const uint32_t* __restrict__ canon, SortParams sp,
    uint32_t* __restrict__ h1)
{
    extern __shared__ uint32_t hist[];
    for (uint32_t i = threadIdx.x; i < sp.NC; i += R1_THREADS)
        hist[i] = 0;
    __syncthreads();
    const uint64_t p0 = (uint64_t)blockIdx.x * R1_TILE;
    for (uint32_t k = 0; k < R1_PER; ++k) {
        const uint64_t i = p0 + (uint64_t)k * R1_THREADS + threadIdx.x;
        if (i >= sp.n)
            break;
        uint32_t s[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            s[q] = canon[8 * i + q];
        uint32_t carry = 0;
        for (uint32_t w = 0; w < sp.W; ++w) {
            const int32_t d = next_digit(s, w, sp.c, carry);
            if (d != 0) {
                const uint32_t bi = (uint32_t)(d < 0 ? -d : d) - 1u;
                atomicAdd(&hist[(w << sp.CB) + (bi >> sp.FB)], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < sp.NC; b += R1_THREADS)
        h1[(uint64_t)b * sp.ntiles + blockIdx.x] = hist[b];
}

// Round 1, pass 2: the same tile places its entries at the scanned offsets.
__global__ void __launch_bounds__(R1_THREADS) r1_scatter_kernel(const uint32_t* __restrict__ canon, SortParams sp,
    const uint32_t* __restrict__ h1off, uint32_t* __restrict__ val1, uint16_t* __restrict__ key1)
{
    extern __shared__ uint32_t cursor[];
    for (uint32_t b = threadIdx.x; b < sp.NC; b += R1_THREADS)
        cursor[b] = h1off[(uint64_t)b * sp.ntiles + blockIdx.x];
    __syncthreads();
    const uint32_t fmask = (1u << sp.FB) - 1u;
    const uint64_t p0 = (uint64_t)blockIdx.x * R1_TILE;
    for (uint32_t k = 0; k < R1_PER; ++k) {
        const uint64_t i = p0 + (uint64_t)k * R1_THREADS + threadIdx.x;
        if (i >= sp.n)
            break;
        uint32_t s[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            s[q] = canon[8 * i + q];
        uint32_t carry = 0;
        for (uint32_t w = 0; w < sp.W; ++w) {
            const int32_t d = next_digit(s, w, sp.c, carry);
            if (d != 0) {
                const uint32_t bi = (uint32_t)(d < 0 ? -d : d) - 1u;
                const uint32_t pos = atomicAdd(&cursor[(w << sp.CB) + (bi >> sp.FB)], 1u);
                val1[pos] = (uint32_t)i | (d < 0 ? 0x80000000u : 0u);
                key1[pos] = (uint16_t)(bi & fmask);
            }
        }
    }
}

// One work-group: bin boundaries from the round-1 scan, and the number of
// round-2 chunks per bin (at least one, so every bucket offset gets written).
__global__ void __launch_bounds__(1024) bin_info_kernel(const uint32_t* __restrict__ h1off, SortParams sp,
    uint32_t* __restrict__ bin_start, uint32_t* __restrict__ chunk_base, uint32_t* __restrict__ off0_end)
{
    __shared__ uint32_t s[1024];
    constexpr uint32_t PER = 4; // NC <= 4096
    uint32_t k[PER];
    uint32_t sum = 0;
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t b = threadIdx.x * PER + q;
        k[q] = 0;
        if (b < sp.NC) {
            const uint32_t lo = h1off[(uint64_t)b * sp.ntiles];
            const uint32_t hi = h1off[(uint64_t)(b + 1) * sp.ntiles];
            bin_start[b] = lo;
            const uint32_t sz = hi - lo;
            k[q] = sz == 0 ? 1u : (sz + R2_CHUNK - 1) / R2_CHUNK;
        }
        sum += k[q];
    }
    s[threadIdx.x] = sum;
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        __syncthreads();
        const uint32_t t = threadIdx.x >= d ? s[threadIdx.x - d] : 0u;
        __syncthreads();
        s[threadIdx.x] += t;
    }
    __syncthreads();
    uint32_t run = threadIdx.x ? s[threadIdx.x - 1] : 0u;
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t b = threadIdx.x * PER + q;
        if (b < sp.NC)
            chunk_base[b] = run;
        run += k[q];
    }
    if (threadIdx.x == 0) {
        const uint32_t total = h1off[(uint64_t)sp.NC * sp.ntiles];
        bin_start[sp.NC] = total;
        chunk_base[sp.NC] = s[1023];
        *off0_end = total;
    }
}

// largest b in [0, nbins) with base[b] <= t (base strictly increasing, base[0] = 0)
__device__ __forceinline__ uint32_t find_bin(const uint32_t* __restrict__ base, uint32_t nbins, uint32_t t)
{
    uint32_t lo = 0, hi = nbins;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Round 2, pass 1: LDS histogram of the low bucket bits of one chunk of one
// bin; stored at (bin, low, chunk) order. Blocks beyond the real chunk count
// zero their slice so the scan over the upper bound stays exact.
__global__ void __launch_bounds__(R2_THREADS) r2_count_kernel(const uint16_t* __restrict__ key1, SortParams sp,
    const uint32_t* __restrict__ bin_start, const uint32_t* __restrict__ chunk_base, uint32_t* __restrict__ h2)
{
    __shared__ uint32_t hist[1u << MAX_FB];
    const uint32_t FBN = 1u << sp.FB;
    const uint32_t t = blockIdx.x;
    if (t >= chunk_base[sp.NC]) {
        for (uint32_t i = threadIdx.x; i < FBN; i += R2_THREADS)
            h2[((uint64_t)t << sp.FB) + i] = 0;
        return;
    }
    for (uint32_t i = threadIdx.x; i < FBN; i += R2_THREADS)
        hist[i] = 0;
    const uint32_t b = find_bin(chunk_base, sp.NC, t);
    const uint32_t j = t - chunk_base[b];
    const uint32_t kb = chunk_base[b + 1] - chunk_base[b];
    const uint32_t start = bin_start[b] + j * R2_CHUNK;
    uint32_t end = start + R2_CHUNK;
    if (end > bin_start[b + 1])
        end = bin_start[b + 1];
    __syncthreads();
    for (uint32_t e = start + threadIdx.x; e < end; e += R2_THREADS)
        atomicAdd(&hist[key1[e]], 1u);
    __syncthreads();
    const uint64_t base = (uint64_t)chunk_base[b] << sp.FB;
    for (uint32_t low = threadIdx.x; low < FBN; low += R2_THREADS)
        h2[base + (uint64_t)low * kb + j] = hist[low];
}

// Round 2, pass 2: final placement; chunk 0 of each bin also publishes the
// bucket offsets (bucket id = bin * 2^FB + low).
__global__ void __launch_bounds__(R2_THREADS) r2_scatter_kernel(const uint16_t* __restrict__ key1,
    const uint32_t* __restrict__ val1, SortParams sp, const uint32_t* __restrict__ bin_start,
    const uint32_t* __restrict__ chunk_base, const uint32_t* __restrict__ h2off, uint32_t* __restrict__ entries,
    uint32_t* __restrict__ off0)
{
    __shared__ uint32_t cursor[1u << MAX_FB];
    const uint32_t FBN = 1u << sp.FB;
    const uint32_t t = blockIdx.x;
    if (t >= chunk_base[sp.NC])
        return;
    const uint32_t b = find_bin(chunk_base, sp.NC, t);
    const uint32_t j = t - chunk_base[b];
    const uint32_t kb = chunk_base[b + 1] - chunk_base[b];
    const uint64_t base = (uint64_t)chunk_base[b] << sp.FB;
    for (uint32_t low = threadIdx.x; low < FBN; low += R2_THREADS) {
        const uint32_t o = h2off[base + (uint64_t)low * kb + j];
        cursor[low] = o;
        if (j == 0)
            off0[((uint64_t)b << sp.FB) + low] = o;
    }
    const uint32_t start = bin_start[b] + j * R2_CHUNK;
    uint32_t end = start + R2_CHUNK;
    if (end > bin_start[b + 1])
        end = bin_start[b + 1];
    __syncthreads();
    for (uint32_t e = start + threadIdx.x; e < end; e += R2_THREADS) {
        const uint32_t pos = atomicAdd(&cursor[key1[e]], 1u);
        entries[pos] = val1[e];
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

__global__ void piece_map_kernel(const uint32_t* __restrict__ poff, uint64_t nb, uint32_t* __restrict__ piece_bucket)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb)
        return;
    for (uint32_t p = poff[b]; p < poff[b + 1]; ++p)
        piece_bucket[p] = (uint32_t)b;
}

// bases -> 2^261 scale (x * 2^261 = Montgomery product of x * 2^256 and 2^261);
// the identity (0, 0) stays (0, 0)
__global__ void to_scale261_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Fq k;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        k.v[q] = Fq29Params::TO261[q];
    (Fq::load(in + 64 * i) * k).store(out + 64 * i);
    (Fq::load(in + 64 * i + 32) * k).store(out + 64 * i + 32);
}

// level 0: one thread per piece of <= S entries of one bucket; affine inputs.
// A single-piece bucket is final (bsum); otherwise the partial goes to part.
__global__ void __launch_bounds__(256) sum_pieces_kernel(const uint32_t* __restrict__ entries,
    const uint8_t* __restrict__ points, const uint32_t* __restrict__ off, const uint32_t* __restrict__ poff,
    const uint32_t* __restrict__ piece_bucket, uint64_t nb, uint32_t S, G1Xyzz* __restrict__ part,
    G1Xyzz* __restrict__ bsum)
{
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= poff[nb])
        return;
    const uint32_t b = piece_bucket[p];
    const uint32_t first = poff[b];
    const uint64_t start = off[b] + (uint64_t)(p - first) * S;
    uint64_t end = start + S;
    if (end > off[b + 1])
        end = off[b + 1];
    G1X29 acc = G1X29::inf();
    for (uint64_t e = start; e < end; ++e) {
        const uint32_t ent = entries[e];
        G1A29 q = G1A29::load(points + 64ull * (ent & 0x7fffffffu));
        if (ent >> 31)
            q = q.neg();
        acc = G1X29::add_affine(acc, q);
    }
    if (poff[b + 1] - first == 1)
        acc.store(bsum + b);
    else
        acc.store(part + p);
}

// buckets with 2..LIGHT pieces: one thread sums them; more: queued for the
// work-group reduction. heavy[nb] is the queue length (zeroed beforehand).
__global__ void __launch_bounds__(256) merge_light_kernel(const G1Xyzz* __restrict__ part,
    const uint32_t* __restrict__ poff, uint64_t nb, G1Xyzz* __restrict__ bsum, uint32_t* __restrict__ heavy)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb)
        return;
    const uint32_t lo = poff[b], hi = poff[b + 1];
    const uint32_t k = hi - lo;
    if (k <= 1)
        return;
    if (k > LIGHT) {
        const uint32_t slot = atomicAdd(&heavy[nb], 1u);
        heavy[slot] = (uint32_t)b;
        return;
    }
    G1X29 acc = G1X29::load(part + lo);
    for (uint32_t q = lo + 1; q < hi; ++q)
        acc = G1X29::add(acc, G1X29::load(part + q));
    acc.store(bsum + b);
}

// one work-group per queued dense bucket: strided partial sums, LDS tree
__global__ void __launch_bounds__(HEAVY_THREADS) merge_heavy_kernel(const G1Xyzz* __restrict__ part,
    const uint32_t* __restrict__ poff, uint64_t nb, const uint32_t* __restrict__ heavy, G1Xyzz* __restrict__ bsum)
{
    __shared__ G1X29 sh[HEAVY_THREADS];
    const uint32_t count = heavy[nb];
    for (uint32_t h = blockIdx.x; h < count; h += gridDim.x) {
        const uint32_t b = heavy[h];
        const uint32_t lo = poff[b], hi = poff[b + 1];
        G1X29 acc = G1X29::inf();
        for (uint32_t q = lo + threadIdx.x; q < hi; q += HEAVY_THREADS)
            acc = G1X29::add(acc, G1X29::load(part + q));
        sh[threadIdx.x] = acc;
        for (uint32_t d = HEAVY_THREADS / 2; d > 0; d >>= 1) {
            __syncthreads();
            if (threadIdx.x < d)
                sh[threadIdx.x] = G1X29::add(sh[threadIdx.x], sh[threadIdx.x + d]);
        }
        __syncthreads();
        if (threadIdx.x == 0)
            sh[0].store(bsum + b);
        __syncthreads();
    }
}

__device__ __forceinline__ G1X29 dbl_times(G1X29 x, uint32_t k)
{
    for (uint32_t i = 0; i < k; ++i)
        x = G1X29::dbl(x);
    return x;
}

// first reduction level: segment m of window w covers buckets (1-based)
// m*L+1 .. m*L+L.  T = sum_j j * B_{mL+j} (local weights), S = L * sum_j B.
// An empty bucket (no entries) is the identity and its bsum slot is not read.
__global__ void __launch_bounds__(256) reduce_first_kernel(const G1Xyzz* __restrict__ bsum,
    const uint32_t* __restrict__ off, uint32_t W, uint32_t B, uint32_t L, uint32_t log_l, uint64_t M,
    G1Xyzz* __restrict__ T, G1Xyzz* __restrict__ S)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)W * M)
        return;
    const uint64_t w = t / M, m = t % M;
    const uint64_t b0 = w * B + m * L;
    G1X29 run = G1X29::inf(), acc = G1X29::inf();
    for (int j = (int)L - 1; j >= 0; --j) {
        const uint64_t b = b0 + (uint32_t)j;
        if (off[b + 1] != off[b])
            run = G1X29::add(run, G1X29::load(bsum + b));
        acc = G1X29::add(acc, run);
    }
    acc.store(T + t);
    if (M > 1)
        dbl_times(run, log_l).store(S + t);
}

// later levels: answer = sum_m T_m + sum_m m * S_m over Mi segments; group g
// of Lg segments gives T' = sum_i T_{gLg+i} + sum_i i * S_{gLg+i} and
// S' = Lg * sum_i S_{gLg+i}, so the form is kept with Mi / Lg segments.
__global__ void __launch_bounds__(256) reduce_level_kernel(const G1Xyzz* __restrict__ Ti,
    const G1Xyzz* __restrict__ Si, uint32_t W, uint64_t Mi, uint32_t Lg, uint32_t log_lg, G1Xyzz* __restrict__ To,
    G1Xyzz* __restrict__ So)
{
    const uint64_t Mo = Mi / Lg;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)W * Mo)
        return;
    const uint64_t w = t / Mo, g = t % Mo;
    const uint64_t base = w * Mi + g * Lg;
    G1X29 a = G1X29::inf(), run = G1X29::inf(), acc = G1X29::inf();
    for (int i = (int)Lg - 1; i >= 0; --i) {
        a = G1X29::add(a, G1X29::load(Ti + base + i));
        run = G1X29::add(run, G1X29::load(Si + base + i));
        if (i > 0)
            acc = G1X29::add(acc, run);
    }
    G1X29::add(a, acc).store(To + t);
    if (Mo > 1)
        dbl_times(run, log_lg).store(So + t);
}

__global__ void __launch_bounds__(64) finalize_kernel(const G1Xyzz* __restrict__ win, uint32_t W, uint32_t c, uint64_t* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    G1X29 acc = G1X29::load(win + (W - 1));
    for (int w = (int)W - 2; w >= 0; --w) {
        for (uint32_t k = 0; k < c; ++k)
            acc = G1X29::dbl(acc);
        acc = G1X29::add(acc, G1X29::load(win + w));
    }
    Fq one = Fq::one();
    if (acc.is_inf()) {
        Fq::zero().store(out);
        one.store(out + 4);
        Fq::zero().store(out + 8);
    } else {
        G1Affine a = acc.to_xyzz256().to_affine();
        a.x.store(out);
        a.y.store(out + 4);
        one.store(out + 8);
    }
}

unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

uint32_t ilog2(uint32_t x) { return 31 - __builtin_clz(x); }

// Full device pipeline; the 96-byte result lands in b.out (device).
RustError msm_device(const MsmPlan& p, const MsmBuffers& b, const Fr* d_scalars, const uint8_t* d_points,
    hipStream_t stream)
{
    const unsigned T = 256;
    ProfScope prof_all("msm_total", stream);
    std::unique_ptr<ProfScope> prof_sort(new ProfScope("msm_sort", stream));
    to_canonical_kernel<<<blocks_for(p.n, T), T, 0, stream>>>(d_scalars, b.canon, p.n);
    H2_LAUNCH_CHECK();
    to_scale261_kernel<<<blocks_for(p.n, T), T, 0, stream>>>(d_points, b.pts, p.n);
    H2_LAUNCH_CHECK();
    SortParams sp { p.n, p.c, p.W, p.FB, p.CB, p.NC, p.ntiles };
    const uint64_t nh1 = (uint64_t)p.NC * p.ntiles;
    r1_count_kernel<<<p.ntiles, R1_THREADS, 4 * p.NC, stream>>>(b.canon, sp, b.h1);
    H2_LAUNCH_CHECK();
    RustError e = scan_u32_exclusive(b.h1, b.h1off, nh1, b.scan_scratch, stream);
    if (e.code)
        return e;
    r1_scatter_kernel<<<p.ntiles, R1_THREADS, 4 * p.NC, stream>>>(b.canon, sp, b.h1off, b.val1, b.key1);
    H2_LAUNCH_CHECK();
    bin_info_kernel<<<1, 1024, 0, stream>>>(b.h1off, sp, b.bin_start, b.chunk_base, b.off0 + p.nb);
    H2_LAUNCH_CHECK();
    r2_count_kernel<<<(unsigned)p.max_chunks, R2_THREADS, 0, stream>>>(b.key1, sp, b.bin_start, b.chunk_base, b.h2);
    H2_LAUNCH_CHECK();
    e = scan_u32_exclusive(b.h2, b.h2off, p.max_chunks << p.FB, b.scan_scratch, stream);
    if (e.code)
        return e;
    r2_scatter_kernel<<<(unsigned)p.max_chunks, R2_THREADS, 0, stream>>>(b.key1, b.val1, sp, b.bin_start,
        b.chunk_base, b.h2off, b.entries, b.off0);
    H2_LAUNCH_CHECK();
    prof_sort.reset();

    {
        ProfScope prof("msm_buckets", stream);
        piece_count_kernel<<<blocks_for(p.nb, T), T, 0, stream>>>(b.off0, p.nb, p.S0, b.pc);
        H2_LAUNCH_CHECK();
        e = scan_u32_exclusive(b.pc, b.poff, p.nb, b.scan_scratch, stream);
        if (e.code)
            return e;
        piece_map_kernel<<<blocks_for(p.nb, T), T, 0, stream>>>(b.poff, p.nb, b.piece_bucket);
        H2_LAUNCH_CHECK();
        sum_pieces_kernel<<<blocks_for(p.ub0, T), T, 0, stream>>>(b.entries, b.pts, b.off0, b.poff,
            b.piece_bucket, p.nb, p.S0, b.part, b.bsum);
        H2_LAUNCH_CHECK();
        H2_TRY(hipMemsetAsync(b.heavy + p.nb, 0, 4, stream));
        merge_light_kernel<<<blocks_for(p.nb, T), T, 0, stream>>>(b.part, b.poff, p.nb, b.bsum, b.heavy);
        H2_LAUNCH_CHECK();
        merge_heavy_kernel<<<HEAVY_BLOCKS, HEAVY_THREADS, 0, stream>>>(b.part, b.poff, p.nb, b.heavy, b.bsum);
        H2_LAUNCH_CHECK();
    }

    ProfScope prof_red("msm_reduce", stream);
    const uint64_t M0 = p.Mlev[0];
    reduce_first_kernel<<<blocks_for((uint64_t)p.W * M0, T), T, 0, stream>>>(b.bsum, b.off0, p.W, p.B, p.L1,
        ilog2(p.L1), M0, b.red[0], b.red[1]);
    H2_LAUNCH_CHECK();
    int cur = 0;
    for (uint32_t lev = 1; lev <= p.nlev; ++lev) {
        const uint64_t Mi = p.Mlev[lev - 1];
        const uint32_t lg = p.Llev[lev];
        reduce_level_kernel<<<blocks_for((uint64_t)p.W * (Mi / lg), T), T, 0, stream>>>(b.red[2 * cur],
            b.red[2 * cur + 1], p.W, Mi, lg, ilog2(lg), b.red[2 * (cur ^ 1)], b.red[2 * (cur ^ 1) + 1]);
        H2_LAUNCH_CHECK();
        cur ^= 1;
    }
    finalize_kernel<<<1, 64, 0, stream>>>(b.red[2 * cur], p.W, p.c, b.out);
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
