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
//       round 2 splits each bin by the low FB bucket bits, chunk by chunk,
//       staging the chunk in LDS so the final stores come out in runs.
//     Per-tile / per-chunk histograms are laid out so that one linear
//     exclusive scan yields every entry's destination, whatever the skew of
//     the scalars (a bucket holding every point costs the same sort work).
//  4. bucket sums, balanced: thread t adds the S sorted entries
//     [tS, tS + S) with XYZZ mixed additions, whatever the bucket boundaries
//     (the next point is fetched while the current one is added). A bucket
//     wholly inside one thread's range is final at once; the pieces of a
//     bucket split between threads are merged afterwards (one thread for up
//     to LIGHT pieces, a work-group tree for denser buckets).
//  5. per window, sum_b b * B_b: running sums over segments of L1 buckets
//     give pairs (T, S) with  answer = sum T + sum_m m * S_m;  work-group
//     trees over 256 pairs keep that form (T = T_L + T_R + S_R,
//     S = 2 (S_L + S_R), S carrying the block size as a weight) until one
//     pair per window is left.
//  6. on the host: Horner over the exported window sums (c doublings per
//     window) and conversion to affine, a serial chain a CPU core finishes
//     faster than one GPU lane.
// The result is written as (x, y, 1) or the identity (0, 1, 0) in Montgomery
// form, the layout of the reference's `cukernel_to_affine` output.
#include <atomic>
#include <mutex>
#include <optional>
#include <vector>

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
constexpr uint32_t R2_THREADS = 1024;
constexpr uint32_t R2_CHUNK = 16384; // entries per round-2 chunk (staged in LDS)
constexpr uint32_t LIGHT = 32; // pieces of a split bucket summed by one thread
constexpr uint32_t HEAVY_THREADS = 256;
constexpr uint32_t HEAVY_BLOCKS = 512;
constexpr uint32_t RED_L1 = 16; // at most this many buckets per first-level reduction segment
constexpr uint64_t RED_THREADS = 1ull << 17; // first-level segments wanted: 2 waves per SIMD
constexpr uint32_t TREE = 256; // threads (tree leaves) per work-group at later levels
constexpr uint32_t TREE_K = 2; // pairs folded sequentially into each leaf
constexpr uint32_t MAX_RED_LEVELS = 8;
constexpr uint32_t MAX_JOBS = 64; // MSMs in one batched pipeline

// One MSM of a batch: its points occupy the virtual range [voff, voff + n)
// (voff a multiple of R1_TILE, so a round-1 tile belongs to one job) and the
// round-1 tiles [tile0, tile0 + tiles).
struct BatchJob {
    uint64_t voff, n;
    uint32_t tile0, tiles;
};

// 128-byte slot holding one XYZZ point (four 32-byte coordinates)
struct PointSlot {
    uint4 w[8];
};

struct MsmPlan {
    uint64_t n; // points (for a batch: the virtual range of all jobs)
    uint64_t nbase; // bases converted to the 2^261 scale (a batch shares one base array)
    bool glv; // scalars split k = k1 + lambda k2: 2 nbase virtual points of 128-bit scalars
    uint64_t pts_count; // bases held in the 2^261 scale: nbase, or 2 nbase with the endomorphism images
    uint32_t J; // MSMs in the pipeline (1 unless batched); each has its own W bucket sets (1 with a table)
    uint32_t c, W, B;
    uint32_t Wb; // bucket windows: J W, or J when the windows share buckets (precomputed bases)
    uint64_t table_n; // precomputed bases: row stride of the table (0 = plain bases)
    uint64_t nb; // Wb * B buckets
    uint32_t FB, CB; // round-2 (fine) and round-1 (coarse) bucket bits, FB + CB = c - 1
    uint32_t NCj; // coarse bins per MSM = (W, or 1 for shared buckets) << CB
    uint32_t NC; // coarse bins = J * NCj
    uint32_t ntiles; // round-1 tiles
    uint64_t max_chunks; // upper bound on round-2 chunks
    uint32_t S; // sorted entries per accumulation thread
    uint64_t nthr; // upper bound on accumulation threads
    uint32_t L1; // buckets per first reduction segment
    uint32_t nlev; // tree levels after the first
    uint64_t Mlev[MAX_RED_LEVELS + 1]; // pairs per window: Mlev[0] = B / L1, then / (leaves * K) per level
    uint32_t Klev[MAX_RED_LEVELS + 1]; // pairs per leaf at level i (i >= 1)
};

uint32_t log2_floor(uint64_t n)
{
    uint32_t l = 0;
    while ((2ull << l) <= n)
        ++l;
    return l;
}

// Window width: plain bases need W bucket sets (one per window) and a Horner
// over windows, so c grows more slowly than log2 n; precomputed bases
// (rows 2^(c w) P) share one bucket set, so the width can follow log2 n.
// Window widths chosen by a tuner (0 = use the default rule): entry i < 32 for
// plain MSMs of floor(log2 n) = i, entry 32 + i for precomputed tables of
// floor(log2 table_n) = i. Set with h2amd_msm_set_widths before sizing,
// launching and building tables.
// Atomic entries: a setter on one host thread may race planners on others.
std::atomic<uint8_t> g_width_table[64] = {};

// n (virtual) points of `bits`-bit signed-digit scalars
uint32_t default_width(uint64_t n, bool shared, uint32_t bits)
{
    int c = (int)log2_floor(n < 1 ? 1 : n) - (shared ? 0 : 4);
    if (c < 4)
        c = 4;
    if (c > 20) // shared buckets too: measured 2^22 tables 7.6 ms at 20 bits vs 8.3 at 22, 2^24 equal
        c = 20;
    if (!shared && c > 19) // plain 2^24: 26.4 ms at 19 bits (14 windows) vs 27.2 at 20 (13 windows)
        c = 19;
    if (shared && c == 20 && n < (1ull << 21)) // 2^20 tables: 15 rows of 17 bits, an eighth of the
        c = 17; // buckets to reduce: 2.21 vs 2.28 ms per MSM, a k = 20 proof 42.0 vs 43.7 ms
    // same window count, narrowest width: the top window is as full as the others
    const uint32_t w0 = (bits + (uint32_t)c - 1) / (uint32_t)c;
    return (bits + w0 - 1) / w0;
}

uint32_t plan_width(uint64_t n, bool shared)
{
    const uint32_t tuned = g_width_table[(log2_floor(n < 1 ? 1 : n) & 31) + (shared ? 32 : 0)].load(std::memory_order_relaxed);
    if (tuned)
        return tuned;
    return default_width(n, shared, 255);
}

// GLV: 2n virtual points with scalars below 2^128 (signed digits need 129 bits).
// It halves the windows, so the per-window reduction and the Horner over
// windows (a serial chain of ~254 doublings without it): the win below 2^21
// points, where they dominate. Above, the additions are the cost and they do
// not shrink (2n points x 129/c windows), while the sort and bucket traffic
// grow by the rounding of the window count, so plain scalars stay.
constexpr uint32_t GLV_BITS = 129;
constexpr uint64_t kGlvBelow = 1ull << 21;

bool use_glv(uint64_t npoints) { return npoints < kGlvBelow; }

uint32_t plain_msm_width(uint64_t npoints)
{
    const uint32_t tuned = g_width_table[log2_floor(npoints < 1 ? 1 : npoints) & 31].load(std::memory_order_relaxed);
    if (tuned)
        return tuned;
    if (!use_glv(npoints))
        return default_width(npoints, false, 255);
    // GLV from 2^19 points: one bit narrower than the rule (the tuner measured 2^20 at
    // c = 16: 3.82 ms against 4.23 ms at c = 17; 2^16..2^18 are within noise of the rule)
    if (npoints >= (1ull << 19))
        return log2_floor(2 * npoints) - 5;
    // 2^18 points: 1.30 ms at c = 16 against 1.54 at the rule's 15 (9 windows either way,
    // but half as many entries per bucket to merge)
    if (npoints >= (1ull << 18))
        return 16;
    return default_width(2 * npoints, false, GLV_BITS);
}

// n points (virtual range), window width c, J MSMs; table_n > 0 selects
// precomputed bases (each MSM's windows share its bucket set).
MsmPlan plan_core(uint64_t n, uint32_t c, uint32_t J, uint64_t table_n, uint32_t bits = 255)
{
    MsmPlan p {};
    p.n = n;
    p.nbase = n;
    p.pts_count = n;
    p.J = J;
    p.table_n = table_n;
    p.c = c;
    p.W = (bits + p.c - 1) / p.c;
    p.Wb = table_n ? J : J * p.W;
    p.B = 1u << (p.c - 1);
    p.nb = (uint64_t)p.Wb * p.B;
    p.FB = p.c - 1 < MAX_FB ? p.c - 1 : MAX_FB;
    p.CB = p.c - 1 - p.FB;
    p.NCj = (table_n ? 1u : p.W) << p.CB;
    p.NC = J * p.NCj;
    p.ntiles = (uint32_t)((n + R1_TILE - 1) / R1_TILE);
    const uint64_t work = n * p.W;
    p.max_chunks = work / R2_CHUNK + p.NC;
    // ~2^20 accumulation threads, 16..128 entries each: longer runs than buckets
    // split fewer buckets between threads (less merging) and still fill the chip
    uint64_t s = work >> 20;
    if (s < 16)
        s = 16;
    if (s > 128)
        s = 128;
    p.S = (uint32_t)s;
    p.nthr = (work + p.S - 1) / p.S;
    // first-level segments: each is a serial chain of 2 L1 point additions, so L1 is
    // the largest power of two in [2, RED_L1] that still leaves ~RED_THREADS chains
    // to fill the chip; the tree levels take the rest
    uint32_t l1 = 2;
    while (l1 < RED_L1 && (uint64_t)p.Wb * p.B / (2 * l1) >= RED_THREADS)
        l1 *= 2;
    p.L1 = p.B < l1 ? p.B : l1;
    p.Mlev[0] = p.B / p.L1;
    uint32_t lev = 0;
    while (p.Mlev[lev] > 1) {
        const uint64_t Mi = p.Mlev[lev];
        ++lev;
        if (Mi >= (uint64_t)TREE * TREE_K) {
            p.Klev[lev] = TREE_K;
            p.Mlev[lev] = Mi / ((uint64_t)TREE * TREE_K);
        } else {
            p.Klev[lev] = Mi > TREE ? (uint32_t)(Mi / TREE) : 1u;
            p.Mlev[lev] = 1;
        }
    }
    p.nlev = lev;
    return p;
}

// n points; table_n > 0 selects precomputed bases whose table was built for
// table_n points (the window width follows the table).
MsmPlan make_plan(uint64_t n, uint64_t table_n = 0)
{
    if (table_n)
        return plan_core(n, plan_width(table_n, true), 1, table_n);
    if (!use_glv(n))
        return plan_core(n, plain_msm_width(n), 1, 0);
    MsmPlan p = plan_core(2 * n, plain_msm_width(n), 1, 0, GLV_BITS);
    p.glv = true;
    p.nbase = n;
    p.pts_count = 2 * n;
    return p;
}

uint64_t align_tile(uint64_t n) { return (n + R1_TILE - 1) / R1_TILE * R1_TILE; }

// J MSMs over one shared base array (or one precomputed table of table_n
// points): one width for all (from the mean job size, or the table's), each
// job's points at a tile-aligned virtual offset.
MsmPlan make_batch_plan(const uint64_t* lengths, uint32_t J, uint64_t table_n = 0)
{
    uint64_t nv = 0, nmax = 0;
    for (uint32_t j = 0; j < J; ++j) {
        nv += align_tile(lengths[j]);
        if (lengths[j] > nmax)
            nmax = lengths[j];
    }
    const uint64_t mean = nv / J ? nv / J : 1;
    uint32_t c = table_n ? plan_width(table_n, true) : plan_width(mean, false);
    if (!table_n && c > 16) // J * W * 2^(c-1) buckets of 128 B: 4 GiB at most for 64 jobs
        c = 16;
    MsmPlan p = plan_core(nv ? nv : 1, c, J, table_n);
    p.n = nv;
    p.ntiles = (uint32_t)(nv / R1_TILE);
    p.nbase = nmax;
    p.pts_count = nmax;
    return p;
}

struct MsmBuffers {
    uint8_t* pts; // nbase affine bases in the 2^261 scale (64 B each)
    BatchJob* jobs; // J (batches only)
    uint32_t* tile_job; // job of each round-1 tile (batches only)
    uint32_t* canon; // n x 8 limbs
    uint32_t* h1; // NC * ntiles
    uint32_t* h1off; // NC * ntiles + 1
    uint64_t* kv1; // n * W round-1 output: (index | sign << 31) | low bucket bits << 32
    uint32_t* bin_start; // NC + 1
    uint32_t* chunk_base; // NC + 1
    uint32_t* h2; // max_chunks * 2^FB
    uint32_t* h2off; // max_chunks * 2^FB + 1
    uint32_t* entries; // n * W, sorted by bucket
    uint32_t* off0; // nb + 1
    uint32_t* heavy; // nb queued split buckets + 1 counter
    PointSlot* part; // 2 * nthr partial sums of split buckets
    PointSlot* bsum; // nb
    PointSlot* red[4]; // (T, S) ping-pong: Wb * Mlev[0] pairs in 0, 1; Wb * Mlev[1] in 2, 3
    void* scan_scratch;
    uint64_t* out; // 12 limbs per MSM
};

template <class WS>
void carve(WS& ws, const MsmPlan& p, MsmBuffers* b)
{
    const uint64_t nh1 = (uint64_t)p.NCj * p.ntiles; // every tile counts the bins of its own MSM
    const uint64_t nh2 = p.max_chunks << p.FB;
    uint64_t scan_n = nh1 + 1;
    if (nh2 + 1 > scan_n)
        scan_n = nh2 + 1;
    const uint64_t nred = (uint64_t)p.Wb * p.Mlev[0]; // pairs after the first level
    const uint64_t nred1 = p.nlev ? (uint64_t)p.Wb * p.Mlev[1] : 0; // the most any later level writes
    void* v[24];
    int k = 0;
    v[k++] = ws.take(p.table_n ? 0 : 64 * p.pts_count);
    v[k++] = ws.take(p.J > 1 ? sizeof(BatchJob) * p.J : 0);
    v[k++] = ws.take(p.J > 1 ? 4ull * p.ntiles : 0);
    v[k++] = ws.take(32 * p.n);
    v[k++] = ws.take(4 * nh1);
    v[k++] = ws.take(4 * (nh1 + 1));
    v[k++] = ws.take(8 * p.n * p.W);
    v[k++] = ws.take(4 * (p.NC + 1));
    v[k++] = ws.take(4 * (p.NC + 1));
    v[k++] = ws.take(4 * nh2);
    v[k++] = ws.take(4 * (nh2 + 1));
    v[k++] = ws.take(4 * p.n * p.W);
    v[k++] = ws.take(4 * (p.nb + 1));
    v[k++] = ws.take(4 * (p.nb + 1));
    v[k++] = ws.take(128 * 2 * p.nthr);
    v[k++] = ws.take(128 * p.nb);
    for (int i = 0; i < 4; ++i)
        v[k++] = ws.take(128 * (i < 2 ? nred : nred1));
    v[k++] = ws.take(scan_u32_scratch(scan_n));
    v[k++] = ws.take(128ull * p.Wb); // window sums exported to the host
    if (b) {
        k = 0;
        b->pts = (uint8_t*)v[k++];
        b->jobs = (BatchJob*)v[k++];
        b->tile_job = (uint32_t*)v[k++];
        b->canon = (uint32_t*)v[k++];
        b->h1 = (uint32_t*)v[k++];
        b->h1off = (uint32_t*)v[k++];
        b->kv1 = (uint64_t*)v[k++];
        b->bin_start = (uint32_t*)v[k++];
        b->chunk_base = (uint32_t*)v[k++];
        b->h2 = (uint32_t*)v[k++];
        b->h2off = (uint32_t*)v[k++];
        b->entries = (uint32_t*)v[k++];
        b->off0 = (uint32_t*)v[k++];
        b->heavy = (uint32_t*)v[k++];
        b->part = (PointSlot*)v[k++];
        b->bsum = (PointSlot*)v[k++];
        for (int i = 0; i < 4; ++i)
            b->red[i] = (PointSlot*)v[k++];
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

uint64_t msm_workspace_bytes(uint64_t n, uint64_t table_n = 0)
{
    MsmPlan p = make_plan(n, table_n);
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

// ------------------------------------------------------- GLV split
//
// BN254 G1 has the endomorphism phi(x, y) = (beta x, y) = lambda (x, y), with
// lambda and beta primitive cube roots of unity mod r and p (lambda is the
// coset generator zeta). A scalar k = k1 + lambda k2 (mod r) with |k1|, |k2| <
// 2^128 turns k P into k1 P + k2 phi(P): 2n points of 128-bit scalars, half the
// windows. The split is Babai rounding against the lattice basis
// (a1, -a2), (a2, b2) of {(x, y) : x + lambda y = 0 mod r}, found by the
// extended Euclidean algorithm on (r, lambda): c1 = floor(k g1 / 2^256),
// c2 = floor(k g2 / 2^256) with g1 = floor(b2 2^256 / r), g2 = floor(a2 2^256 / r),
// then k1 = k - c1 a1 - c2 a2 and k2 = c1 a2 - c2 b2. Any c1, c2 give a valid
// split; the rounding keeps |k1|, |k2| below 2^128 (below 2^127 in practice).
constexpr uint32_t GLV_G1[5] = { 0x00ff6565u, 0x5398fd03u, 0xa773d2d2u, 0x4ccef014u, 0x00000002u };
constexpr uint32_t GLV_G2[3] = { 0xc7e0b3d7u, 0xd91d232eu, 0x00000002u };
constexpr uint32_t GLV_A1[4] = { 0x7d4f1128u, 0x8211bbebu, 0xeeb859fcu, 0x6f4d8248u };
constexpr uint32_t GLV_A2[2] = { 0x94d213e3u, 0x89d32568u };
constexpr uint32_t GLV_B2[4] = { 0x1221250bu, 0x0be4e154u, 0xeeb859fdu, 0x6f4d8248u };
// beta * 2^256 mod p (Montgomery form)
constexpr uint32_t GLV_BETA[8] = { 0x13e80b9cu, 0x3350c88eu, 0xdb5e56b9u, 0x7dce557cu, 0xb615564au, 0x6001b4b8u,
    0x020217e0u, 0x2682e617u };

// out[0, na + nb) = a * b (schoolbook over 32-bit limbs)
template <int NA, int NB>
__device__ __forceinline__ void mul_words(const uint32_t (&a)[NA], const uint32_t (&b)[NB], uint32_t (&out)[NA + NB])
{
#pragma unroll
    for (int i = 0; i < NA + NB; ++i)
        out[i] = 0;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const uint64_t t = (uint64_t)a[i] * b[j] + out[i + j] + carry;
            out[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        out[i + NB] = (uint32_t)carry;
    }
}

// acc (N words, two's complement) -= x (M <= N words)
template <int N, int M>
__device__ __forceinline__ void sub_words(uint32_t (&acc)[N], const uint32_t (&x)[M])
{
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint64_t d = (uint64_t)acc[i] - (i < M ? x[i] : 0u) - borrow;
        acc[i] = (uint32_t)d;
        borrow = (uint32_t)(d >> 32) & 1u;
    }
}

// |v| of an N-word two's complement value into the low 4 words of out (word 7 bit 31 = sign)
template <int N>
__device__ __forceinline__ void store_half(const uint32_t (&v)[N], uint32_t* __restrict__ out)
{
    const bool neg = (v[N - 1] >> 31) != 0;
    uint32_t m[4];
    uint32_t carry = 1;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (neg) {
            const uint64_t t = (uint64_t)(~v[i]) + carry;
            m[i] = (uint32_t)t;
            carry = (uint32_t)(t >> 32);
        } else {
            m[i] = v[i];
        }
    }
    uint4* q = (uint4*)out;
    q[0] = make_uint4(m[0], m[1], m[2], m[3]);
    q[1] = make_uint4(0u, 0u, 0u, neg ? 0x80000000u : 0u);
}

// canon[2i] = (|k1|, sign), canon[2i + 1] = (|k2|, sign) for k = canonical(scal[i])
__global__ void glv_split_kernel(const Fr* __restrict__ scal, uint32_t* __restrict__ canon, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const Fr kv = Fr::from_mont(Fr::load(scal + i));
    uint32_t k[8];
#pragma unroll
    for (int q = 0; q < 8; ++q)
        k[q] = kv.v[q];
    uint32_t p1[13], p2[11];
    mul_words<8, 5>(k, GLV_G1, p1);
    mul_words<8, 3>(k, GLV_G2, p2);
    const uint32_t c1[4] = { p1[8], p1[9], p1[10], p1[11] }; // k g1 < 2^384: c1 < 2^128
    const uint32_t c2[3] = { p2[8], p2[9], p2[10] };
    uint32_t m1[8], m2[5], m3[6], m4[7];
    mul_words<4, 4>(c1, GLV_A1, m1);
    mul_words<3, 2>(c2, GLV_A2, m2);
    mul_words<4, 2>(c1, GLV_A2, m3);
    mul_words<3, 4>(c2, GLV_B2, m4);
    uint32_t k1[9] = { k[0], k[1], k[2], k[3], k[4], k[5], k[6], k[7], 0u };
    sub_words<9, 8>(k1, m1);
    sub_words<9, 5>(k1, m2);
    uint32_t k2[8] = { m3[0], m3[1], m3[2], m3[3], m3[4], m3[5], 0u, 0u };
    sub_words<8, 7>(k2, m4);
    store_half<9>(k1, canon + 16 * i);
    store_half<8>(k2, canon + 16 * i + 8);
}

// pts[2i] = g_i, pts[2i + 1] = phi(g_i) = (beta x, y), both in the 2^261 scale; (0, 0) stays (0, 0)
__global__ void glv_bases_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Fq k, beta;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        k.v[q] = Fq29Params::TO261[q];
        beta.v[q] = GLV_BETA[q];
    }
    const Fq x = Fq::load(in + 64 * i) * k, y = Fq::load(in + 64 * i + 32) * k;
    x.store(out + 128 * i);
    y.store(out + 128 * i + 32);
    (x * beta).store(out + 128 * i + 64);
    y.store(out + 128 * i + 96);
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
    uint64_t table_n; // > 0: windows share the bins and entry w of point i indexes row w of the table
    uint32_t NCj; // coarse bins per MSM
    const BatchJob* jobs; // null for a single MSM
    const uint32_t* tile_job;
};

// The MSM a round-1 tile belongs to: its virtual point range and tile range.
// The round-1 histogram h1 is job-major, then bin-major, then tile: job j's
// bin b of its tile t sits at NCj * tile0 + b * tiles + (t - tile0).
struct TileRange {
    uint64_t v0, vend;
    uint32_t tile0, tiles;
};

__device__ __forceinline__ TileRange tile_range(const SortParams& sp, uint32_t tile)
{
    if (!sp.jobs)
        return TileRange { 0, sp.n, 0, sp.ntiles };
    const BatchJob jb = sp.jobs[sp.tile_job[tile]];
    return TileRange { jb.voff, jb.voff + jb.n, jb.tile0, jb.tiles };
}

__device__ __forceinline__ uint64_t h1_index(const SortParams& sp, const TileRange& r, uint32_t b, uint32_t tile)
{
    return (uint64_t)sp.NCj * r.tile0 + (uint64_t)b * r.tiles + (tile - r.tile0);
}

// Round 1, pass 1: per tile of R1_TILE points, LDS histogram over the coarse
// bins (window, top CB bucket bits); stored bin-major so the linear scan
// orders entries by bin, then by tile.
__global__ void __launch_bounds__(R1_THREADS) r1_count_kernel(const uint32_t* __restrict__ canon, SortParams sp,
    uint32_t* __restrict__ h1)
{
    extern __shared__ uint32_t hist[];
    const TileRange tr = tile_range(sp, blockIdx.x);
    for (uint32_t i = threadIdx.x; i < sp.NCj; i += R1_THREADS)
        hist[i] = 0;
    __syncthreads();
    const uint64_t p0 = (uint64_t)blockIdx.x * R1_TILE;
    for (uint32_t k = 0; k < R1_PER; ++k) {
        const uint64_t i = p0 + (uint64_t)k * R1_THREADS + threadIdx.x;
        if (i >= tr.vend)
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
                atomicAdd(&hist[(sp.table_n ? 0u : w << sp.CB) + (bi >> sp.FB)], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < sp.NCj; b += R1_THREADS)
        h1[h1_index(sp, tr, b, blockIdx.x)] = hist[b];
}

// Round 1, pass 2: the same tile places its entries at the scanned offsets.
// Window by window the tile's entries are first ordered by coarse bin in LDS
// (LDS ranks, a scan of the 2^CB bin counts), then written out in that order,
// so consecutive threads store consecutive addresses of one bin's run instead
// of isolated 8-byte words. With a bin per (window, bucket bits) a bin is
// written in one window iteration only; with shared bins (precomputed bases)
// the tile's cursors advance after every window.
uint32_t r1_scatter_lds(uint32_t NC, uint32_t CB)
{
    return 4 * ((NC + 1) & ~1u) + 4 * (((1u << CB) + 2) & ~1u) + 10 * R1_TILE; // < 56 KiB for c <= 20
}

__global__ void __launch_bounds__(R1_THREADS) r1_scatter_kernel(const uint32_t* __restrict__ canon, SortParams sp,
    const uint32_t* __restrict__ h1off, uint64_t* __restrict__ kv1)
{
    extern __shared__ uint32_t sm[];
    const uint32_t BW = 1u << sp.CB;
    uint32_t* cursor = sm; // NCj: this tile's start in every (window, bin) run of its MSM
    uint32_t* base = cursor + ((sp.NCj + 1) & ~1u); // BW + 1: per-window bin counts, then their exclusive scan
    uint64_t* stage = (uint64_t*)(base + ((BW + 2) & ~1u)); // R1_TILE entries, bin-ordered
    uint16_t* sbin = (uint16_t*)(stage + R1_TILE); // the bin of each staged entry
    __shared__ uint32_t wsum[R1_THREADS / 64];
    const TileRange tr = tile_range(sp, blockIdx.x);
    for (uint32_t b = threadIdx.x; b < sp.NCj; b += R1_THREADS)
        cursor[b] = h1off[h1_index(sp, tr, b, blockIdx.x)];
    const uint32_t fmask = (1u << sp.FB) - 1u;
    const uint64_t p0 = (uint64_t)blockIdx.x * R1_TILE;
    uint32_t s[R1_PER][8];
    uint32_t carry[R1_PER];
#pragma unroll
    for (uint32_t k = 0; k < R1_PER; ++k) {
        const uint64_t i = p0 + (uint64_t)k * R1_THREADS + threadIdx.x;
        carry[k] = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            s[k][q] = i < tr.vend ? canon[8 * i + q] : 0u; // a zero scalar has only zero digits
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t w = 0; w < sp.W; ++w) {
        for (uint32_t b = threadIdx.x; b <= BW; b += R1_THREADS)
            base[b] = 0;
        __syncthreads();
        uint32_t bin[R1_PER], rank[R1_PER];
        uint64_t ent[R1_PER];
#pragma unroll
        for (uint32_t k = 0; k < R1_PER; ++k) {
            const uint64_t i = p0 + (uint64_t)k * R1_THREADS + threadIdx.x;
            const int32_t d = next_digit(s[k], w, sp.c, carry[k]);
            bin[k] = BW; // no entry
            if (d != 0) {
                const uint32_t bi = (uint32_t)(d < 0 ? -d : d) - 1u;
                bin[k] = bi >> sp.FB;
                const uint32_t idx = (uint32_t)((sp.table_n ? w * sp.table_n : 0) + (i - tr.v0));
                // the point's own sign (a negative GLV half) rides in bit 31 of scalar word 7
                const bool neg = (d < 0) != ((s[k][7] >> 31) != 0);
                ent[k] = ((uint64_t)(bi & fmask) << 32) | idx | (neg ? 0x80000000u : 0u);
                rank[k] = atomicAdd(&base[bin[k]], 1u);
            }
        }
        __syncthreads();
        // exclusive scan of base[0, BW) in place, base[BW] = total: each thread
        // owns a contiguous run of ceil(BW / R1_THREADS) counters
        const uint32_t per = (BW + R1_THREADS - 1) / R1_THREADS;
        const uint32_t b0 = threadIdx.x * per;
        uint32_t run = 0;
        for (uint32_t q = 0; q < per; ++q)
            if (b0 + q < BW)
                run += base[b0 + q];
        uint32_t incl = run;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t t = __shfl_up(incl, d, 64);
            if (lane >= d)
                incl += t;
        }
        if (lane == 63)
            wsum[wave] = incl;
        __syncthreads();
        uint32_t woff = 0;
        for (uint32_t q = 0; q < wave; ++q)
            woff += wsum[q];
        uint32_t acc = woff + incl - run;
        for (uint32_t q = 0; q < per; ++q) {
            if (b0 + q < BW) {
                const uint32_t c = base[b0 + q];
                base[b0 + q] = acc;
                acc += c;
            }
        }
        if (threadIdx.x == R1_THREADS - 1)
            base[BW] = acc;
        __syncthreads();
#pragma unroll
        for (uint32_t k = 0; k < R1_PER; ++k) {
            if (bin[k] < BW) {
                const uint32_t idx = base[bin[k]] + rank[k];
                stage[idx] = ent[k];
                sbin[idx] = (uint16_t)bin[k];
            }
        }
        __syncthreads();
        const uint32_t total = base[BW];
        const uint32_t wbin = sp.table_n ? 0u : w << sp.CB;
        for (uint32_t idx = threadIdx.x; idx < total; idx += R1_THREADS) {
            const uint32_t b = sbin[idx];
            kv1[cursor[wbin + b] + (idx - base[b])] = stage[idx];
        }
        __syncthreads();
        if (sp.table_n) { // shared bins: the next window continues after this one's entries
            for (uint32_t b = threadIdx.x; b < BW; b += R1_THREADS)
                cursor[b] += base[b + 1] - base[b];
            __syncthreads();
        }
    }
}

// Start of global coarse bin g (g <= NC) in the round-1 order: job-major,
// then bin, then tile (bin g = j * NCj + b of job j).
__device__ __forceinline__ uint32_t bin_lo(const uint32_t* __restrict__ h1off, const SortParams& sp, uint32_t g)
{
    if (!sp.jobs)
        return h1off[(uint64_t)g * sp.ntiles];
    if (g >= sp.NC)
        return h1off[(uint64_t)sp.NCj * sp.ntiles];
    const BatchJob jb = sp.jobs[g / sp.NCj];
    return h1off[(uint64_t)sp.NCj * jb.tile0 + (uint64_t)(g % sp.NCj) * jb.tiles];
}

// One work-group: bin boundaries from the round-1 scan, and the number of
// round-2 chunks per bin (at least one, so every bucket offset gets written).
// Thread t owns the contiguous bins [t * per, (t + 1) * per).
__global__ void __launch_bounds__(1024) bin_info_kernel(const uint32_t* __restrict__ h1off, SortParams sp,
    uint32_t* __restrict__ bin_start, uint32_t* __restrict__ chunk_base, uint32_t* __restrict__ off0_end)
{
    __shared__ uint32_t s[1024];
    const uint32_t per = (sp.NC + 1023) / 1024;
    const uint32_t g0 = threadIdx.x * per;
    auto chunks = [&](uint32_t g, uint32_t& lo) -> uint32_t {
        lo = bin_lo(h1off, sp, g);
        const uint32_t sz = bin_lo(h1off, sp, g + 1) - lo;
        return sz == 0 ? 1u : (sz + R2_CHUNK - 1) / R2_CHUNK;
    };
    uint32_t sum = 0;
    for (uint32_t q = 0; q < per; ++q) {
        uint32_t lo;
        if (g0 + q < sp.NC)
            sum += chunks(g0 + q, lo);
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
    for (uint32_t q = 0; q < per; ++q) {
        const uint32_t g = g0 + q;
        if (g < sp.NC) {
            uint32_t lo;
            const uint32_t k = chunks(g, lo);
            bin_start[g] = lo;
            chunk_base[g] = run;
            run += k;
        }
    }
    if (threadIdx.x == 0) {
        const uint32_t total = bin_lo(h1off, sp, sp.NC);
        bin_start[sp.NC] = total;
        chunk_base[sp.NC] = s[1023];
        *off0_end = total;
    }
}

// largest b in [0, nbins) with base[b] <= t (base nondecreasing, base[0] = 0)
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
__global__ void __launch_bounds__(R2_THREADS) r2_count_kernel(const uint64_t* __restrict__ kv1, SortParams sp,
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
        atomicAdd(&hist[(uint32_t)(kv1[e] >> 32)], 1u);
    __syncthreads();
    const uint64_t base = (uint64_t)chunk_base[b] << sp.FB;
    for (uint32_t low = threadIdx.x; low < FBN; low += R2_THREADS)
        h2[base + (uint64_t)low * kb + j] = hist[low];
}

// Round 2, pass 2: the chunk is counting-sorted in LDS (ranks by LDS
// atomics), then written out in LDS order, so consecutive threads store into
// the same bucket run. Chunk 0 of each bin also publishes the bucket offsets
// (bucket id = bin * 2^FB + low).
__global__ void __launch_bounds__(R2_THREADS) r2_scatter_kernel(const uint64_t* __restrict__ kv1, SortParams sp,
    const uint32_t* __restrict__ bin_start, const uint32_t* __restrict__ chunk_base,
    const uint32_t* __restrict__ h2off, uint32_t* __restrict__ entries, uint32_t* __restrict__ off0)
{
    __shared__ uint32_t cnt[1u << MAX_FB]; // histogram, then global run starts
    __shared__ uint32_t loc[1u << MAX_FB]; // local exclusive offsets
    __shared__ uint32_t part[R2_THREADS];
    __shared__ uint32_t sval[R2_CHUNK];
    __shared__ uint16_t skey[R2_CHUNK];
    const uint32_t FBN = 1u << sp.FB;
    const uint32_t t = blockIdx.x;
    if (t >= chunk_base[sp.NC])
        return;
    const uint32_t b = find_bin(chunk_base, sp.NC, t);
    const uint32_t j = t - chunk_base[b];
    const uint32_t kb = chunk_base[b + 1] - chunk_base[b];
    const uint64_t base = (uint64_t)chunk_base[b] << sp.FB;
    const uint32_t start = bin_start[b] + j * R2_CHUNK;
    uint32_t end = start + R2_CHUNK;
    if (end > bin_start[b + 1])
        end = bin_start[b + 1];
    for (uint32_t low = threadIdx.x; low < FBN; low += R2_THREADS) {
        cnt[low] = 0;
        if (j == 0)
            off0[((uint64_t)b << sp.FB) + low] = h2off[base + (uint64_t)low * kb];
    }
    __syncthreads();
    // the chunk stays in registers between the histogram and the placement; the
    // histogram's atomics also give each entry its rank among its key's entries
    constexpr uint32_t PER_T = R2_CHUNK / R2_THREADS;
    uint64_t kv[PER_T];
    uint32_t rank[PER_T];
#pragma unroll
    for (uint32_t q = 0; q < PER_T; ++q) {
        const uint32_t e = start + q * R2_THREADS + threadIdx.x;
        kv[q] = e < end ? kv1[e] : 0;
        rank[q] = e < end ? atomicAdd(&cnt[(uint32_t)(kv[q] >> 32)], 1u) : 0u;
    }
    __syncthreads();
    // exclusive scan of cnt into loc: FBN / R2_THREADS consecutive keys per thread
    const uint32_t per = (FBN + R2_THREADS - 1) / R2_THREADS;
    uint32_t sum = 0;
    for (uint32_t q = 0; q < per; ++q) {
        const uint32_t k = threadIdx.x * per + q;
        if (k < FBN)
            sum += cnt[k];
    }
    part[threadIdx.x] = sum;
    for (uint32_t d = 1; d < R2_THREADS; d <<= 1) {
        __syncthreads();
        const uint32_t v = threadIdx.x >= d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
    }
    __syncthreads();
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t q = 0; q < per; ++q) {
        const uint32_t k = threadIdx.x * per + q;
        if (k < FBN) {
            loc[k] = run;
            run += cnt[k];
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PER_T; ++q) {
        if (start + q * R2_THREADS + threadIdx.x < end) {
            const uint32_t low = (uint32_t)(kv[q] >> 32);
            const uint32_t r = loc[low] + rank[q];
            sval[r] = (uint32_t)kv[q];
            skey[r] = (uint16_t)low;
        }
    }
    __syncthreads();
    // cnt is free again: it now holds, per low key, the global start of this
    // chunk's run minus its local start
    for (uint32_t low = threadIdx.x; low < FBN; low += R2_THREADS)
        cnt[low] = h2off[base + (uint64_t)low * kb + j] - loc[low];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < end - start; i += R2_THREADS)
        entries[cnt[skey[i]] + i] = sval[i];
}

// The nonempty bucket holding entry e, given off[lo] <= e (lo is the bucket
// after the one that just ended at e): the largest x >= lo with off[x] <= e.
// Galloping, then bisection: one step when bucket lo is nonempty (the common
// case), O(log gap) over a run of empty buckets (small scalars leave whole
// windows empty, hundreds of thousands of buckets in a row).
__device__ __forceinline__ uint32_t next_bucket(const uint32_t* __restrict__ off, uint32_t lo, uint32_t nb, uint32_t e)
{
    uint32_t hi = nb; // off[nb] = E > e
    for (uint32_t step = 1;; step <<= 1) {
        const uint64_t probe = (uint64_t)lo + step;
        if (probe >= nb)
            break;
        if (off[probe] > e) {
            hi = (uint32_t)probe;
            break;
        }
        lo = (uint32_t)probe;
    }
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Thread t adds the sorted entries [tS, min(tS + S, E)). A bucket wholly in
// the range goes to bsum; the piece of a bucket that started before the
// range goes to part[2t], the piece of one that runs past it to part[2t+1].
// The next entry's point is fetched (raw words) before the current addition.
__global__ void __launch_bounds__(256) accumulate_kernel(const uint32_t* __restrict__ entries,
    const uint8_t* __restrict__ pts, const uint32_t* __restrict__ off, uint64_t nb, uint32_t S,
    PointSlot* __restrict__ part, PointSlot* __restrict__ bsum)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t E = off[nb];
    const uint64_t e0 = t * S;
    if (e0 >= E)
        return;
    const uint32_t e1 = (uint32_t)(e0 + S < E ? e0 + S : E);
    uint32_t b = find_bin(off, (uint32_t)nb, (uint32_t)e0); // the bucket holding e0 (empty ones skipped)
    uint32_t bend = off[b + 1];
    bool first = off[b] < e0; // the current bucket began before this range
    G1X29 acc = G1X29::inf();
    uint32_t ent = entries[e0];
    const uint4* src = (const uint4*)(pts + 64ull * (ent & 0x7fffffffu));
    uint4 r0 = src[0], r1 = src[1], r2 = src[2], r3 = src[3];
    for (uint32_t e = (uint32_t)e0; e < e1; ++e) {
        const uint32_t en = entries[e + 1 < e1 ? e + 1 : e];
        const uint4* sn = (const uint4*)(pts + 64ull * (en & 0x7fffffffu));
        const uint4 n0 = sn[0], n1 = sn[1], n2 = sn[2], n3 = sn[3];
        if (e == bend) {
            // previous bucket complete within the range (or the range's first piece)
            acc.store(first ? (void*)(part + 2 * t) : (void*)(bsum + b));
            first = false;
            acc = G1X29::inf();
            b = next_bucket(off, b + 1, (uint32_t)nb, e);
            bend = off[b + 1];
        }
        const uint32_t xw[8] = { r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w };
        const uint32_t yw[8] = { r2.x, r2.y, r2.z, r2.w, r3.x, r3.y, r3.z, r3.w };
        G1A29 q { Fq29::from32(xw), Fq29::from32(yw) };
        if (ent >> 31)
            q = q.neg();
        acc = G1X29::add_affine(acc, q);
        ent = en;
        r0 = n0;
        r1 = n1;
        r2 = n2;
        r3 = n3;
    }
    if (bend > e1)
        acc.store(first ? (void*)(part + 2 * t) : (void*)(part + 2 * t + 1));
    else
        acc.store(first ? (void*)(part + 2 * t) : (void*)(bsum + b));
}

// A bucket split between threads t0 < t1 (t0 holds its first entry, so the
// bucket did not begin before t0's range) has its pieces in part[2 t0 + 1]
// and part[2 t] for t0 < t <= t1.
// Up to LIGHT pieces: one thread sums them; more: queued (heavy[nb] counts).
__global__ void __launch_bounds__(256) merge_split_kernel(const PointSlot* __restrict__ part,
    const uint32_t* __restrict__ off, uint64_t nb, uint32_t S, PointSlot* __restrict__ bsum,
    uint32_t* __restrict__ heavy)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb)
        return;
    const uint32_t lo = off[b], hi = off[b + 1];
    if (lo == hi)
        return;
    const uint32_t t0 = lo / S, t1 = (hi - 1) / S;
    if (t0 == t1)
        return;
    if (t1 - t0 + 1 > LIGHT) {
        const uint32_t slot = atomicAdd(&heavy[nb], 1u);
        heavy[slot] = (uint32_t)b;
        return;
    }
    G1X29 acc = G1X29::load(part + 2 * (uint64_t)t0 + 1);
    for (uint32_t t = t0 + 1; t <= t1; ++t)
        acc = G1X29::add(acc, G1X29::load(part + 2 * (uint64_t)t));
    acc.store(bsum + b);
}

// one work-group per queued dense bucket: strided partial sums, LDS tree
__global__ void __launch_bounds__(HEAVY_THREADS) merge_heavy_kernel(const PointSlot* __restrict__ part,
    const uint32_t* __restrict__ off, uint64_t nb, uint32_t S, const uint32_t* __restrict__ heavy,
    PointSlot* __restrict__ bsum)
{
    __shared__ G1X29 sh[HEAVY_THREADS];
    const uint32_t count = heavy[nb];
    for (uint32_t h = blockIdx.x; h < count; h += gridDim.x) {
        const uint32_t b = heavy[h];
        const uint32_t t0 = off[b] / S, t1 = (off[b + 1] - 1) / S;
        G1X29 acc = G1X29::inf();
        for (uint32_t q = t0 + threadIdx.x; q <= t1; q += HEAVY_THREADS)
            acc = G1X29::add(acc, G1X29::load(part + 2 * (uint64_t)q + (q == t0 ? 1 : 0)));
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
__global__ void __launch_bounds__(256) reduce_first_kernel(const PointSlot* __restrict__ bsum,
    const uint32_t* __restrict__ off, uint32_t W, uint32_t B, uint32_t L, uint32_t log_l, uint64_t M,
    PointSlot* __restrict__ T, PointSlot* __restrict__ S)
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

// Tree level: work-group g of window w reduces the pairs
// [g * cnt * K, (g + 1) * cnt * K) of the Mi pairs of that window
// (answer = sum T + sum_m m S_m, S in units of the current segment size).
// Leaf i folds K consecutive pairs (local weights 0..K-1, S scaled by K);
// adjacent blocks of equal size then combine as T = T_L + T_R + S_R,
// S = 2 (S_L + S_R); after log2(cnt) rounds the block's (T, S) is in the same
// form with cnt * K times larger segments.
// The rounds are a chain of dependent point operations, so the two halves of
// every step run on different waves: threads [0, TREE) form T, threads
// [TREE, 2 TREE) form S of the same node (2 point operations per round instead
// of 4). Round d's nodes sit on the first cnt / 2d threads of each half, so the
// waves with no node left skip the round instead of issuing it with idle lanes.
__global__ void __launch_bounds__(2 * TREE) reduce_tree_kernel(const PointSlot* __restrict__ Ti,
    const PointSlot* __restrict__ Si, uint64_t Mi, uint32_t K, uint32_t log_k, PointSlot* __restrict__ To,
    PointSlot* __restrict__ So)
{
    __shared__ G1X29 sT[TREE];
    __shared__ G1X29 sS[TREE];
    const uint64_t per_blk = Mi >= (uint64_t)TREE * K ? (uint64_t)TREE * K : Mi;
    const uint64_t Mo = Mi / per_blk;
    const uint32_t cnt = (uint32_t)(per_blk / K);
    const uint64_t w = blockIdx.x / Mo, g = blockIdx.x % Mo;
    const uint64_t base = w * Mi + g * per_blk;
    const bool s_role = threadIdx.x >= TREE; // whole waves: TREE is a multiple of 64
    // S node i on thread TREE + (i + 64) mod TREE: the waves of a work-group are dealt
    // to the CU's 4 SIMDs in turn, so the late rounds' T wave (threads 0..63) and S
    // wave (threads TREE + 64..) issue on different SIMDs instead of sharing one
    const uint32_t i = s_role ? (threadIdx.x - TREE + TREE - 64) % TREE : threadIdx.x;
    const bool keep_s = cnt > 1 || Mo > 1;
    if (i < cnt) {
        const uint64_t q0 = base + (uint64_t)i * K;
        if (!s_role) { // T = sum_j T_j + sum_j j S_j
            G1X29 a = G1X29::load(Ti + q0 + K - 1), run = G1X29::inf(), acc = G1X29::inf();
            for (int j = (int)K - 1; j > 0; --j) {
                run = G1X29::add(run, G1X29::load(Si + q0 + (uint32_t)j));
                acc = G1X29::add(acc, run);
                a = G1X29::add(a, G1X29::load(Ti + q0 + (uint32_t)j - 1));
            }
            sT[i] = G1X29::add(a, acc);
        } else if (keep_s) { // S = K sum_j S_j
            G1X29 run = G1X29::load(Si + q0);
            for (uint32_t j = 1; j < K; ++j)
                run = G1X29::add(run, G1X29::load(Si + q0 + j));
            sS[i] = dbl_times(run, log_k);
        }
    }
    for (uint32_t d = 1; d < cnt; d <<= 1) {
        __syncthreads();
        const uint32_t l = i * 2 * d, r = l + d;
        const bool need_s = 2 * d < cnt || Mo > 1;
        const bool act = r < cnt && (!s_role || need_s);
        G1X29 v;
        if (act) {
            if (!s_role)
                v = G1X29::add(G1X29::add(sT[l], sT[r]), sS[r]);
            else
                v = G1X29::dbl(G1X29::add(sS[l], sS[r]));
        }
        __syncthreads();
        if (act) {
            if (!s_role)
                sT[l] = v;
            else
                sS[l] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sT[0].store(To + blockIdx.x);
        if (Mo > 1)
            sS[0].store(So + blockIdx.x);
    }
}

// The window sums leave the device as XYZZ points in the boundary's 2^256 scale
// (4 x 32 B each, out[t] for window sum t < Wb): the Horner over windows (a chain
// of c (W - 1) doublings) and the affine conversion run on the host, where one
// core finishes the chain in ~0.1 ms while a single GPU lane needs ~1.5 ms for it.
__global__ void __launch_bounds__(64) export_windows_kernel(const PointSlot* __restrict__ win, uint32_t Wb,
    uint64_t* __restrict__ out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Wb)
        return;
    const G1Xyzz x = G1X29::load(win + t).to_xyzz256();
    x.X.store(out + 16ull * t);
    x.Y.store(out + 16ull * t + 4);
    x.ZZ.store(out + 16ull * t + 8);
    x.ZZZ.store(out + 16ull * t + 12);
}

// ------------------------------------------------ host side: window Horner
//
// Fq in 4 x 64-bit Montgomery form (R = 2^256, the boundary's representation),
// CIOS multiplication; XYZZ points (add-2008-s, dbl-2008-s-1, a = 0) as on the
// device.
struct HFq {
    uint64_t v[4];
};

constexpr uint64_t kHP[4] = { 0x3c208c16d87cfd47ull, 0x97816a916871ca8dull, 0xb85045b68181585dull,
    0x30644e72e131a029ull };
constexpr uint64_t kHN0 = 0x87d20782e4866389ull; // -p^-1 mod 2^64
constexpr uint64_t kHOne[4] = { 0xd35d438dc58f0d9dull, 0x0a78eb28f5c70b3dull, 0x666ea36f7879462cull,
    0x0e0a77c19a07df2full }; // 2^256 mod p

bool hfq_geq_p(const uint64_t (&a)[4])
{
    for (int i = 3; i >= 0; --i)
        if (a[i] != kHP[i])
            return a[i] > kHP[i];
    return true;
}

HFq hfq_sub_p(const HFq& a)
{
    HFq r;
    unsigned __int128 b = 0;
    for (int i = 0; i < 4; ++i) {
        const unsigned __int128 d = (unsigned __int128)a.v[i] - kHP[i] - (uint64_t)b;
        r.v[i] = (uint64_t)d;
        b = (d >> 64) ? 1 : 0;
    }
    return r;
}

HFq hfq_add(const HFq& a, const HFq& b)
{
    HFq r;
    unsigned __int128 c = 0;
    for (int i = 0; i < 4; ++i) {
        c += (unsigned __int128)a.v[i] + b.v[i];
        r.v[i] = (uint64_t)c;
        c >>= 64;
    }
    return (c || hfq_geq_p(r.v)) ? hfq_sub_p(r) : r; // a + b < 2p < 2^256
}

HFq hfq_sub(const HFq& a, const HFq& b)
{
    HFq r;
    unsigned __int128 br = 0;
    for (int i = 0; i < 4; ++i) {
        const unsigned __int128 d = (unsigned __int128)a.v[i] - b.v[i] - (uint64_t)br;
        r.v[i] = (uint64_t)d;
        br = (d >> 64) ? 1 : 0;
    }
    if (br) { // add p back
        unsigned __int128 c = 0;
        for (int i = 0; i < 4; ++i) {
            c += (unsigned __int128)r.v[i] + kHP[i];
            r.v[i] = (uint64_t)c;
            c >>= 64;
        }
    }
    return r;
}

HFq hfq_mul(const HFq& a, const HFq& b)
{
    uint64_t t[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i < 4; ++i) {
        unsigned __int128 c = 0;
        for (int j = 0; j < 4; ++j) {
            c += (unsigned __int128)a.v[j] * b.v[i] + t[j];
            t[j] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[4] = (uint64_t)c;
        t[5] = (uint64_t)(c >> 64);
        const uint64_t m = t[0] * kHN0;
        c = ((unsigned __int128)m * kHP[0] + t[0]) >> 64;
        for (int j = 1; j < 4; ++j) {
            c += (unsigned __int128)m * kHP[j] + t[j];
            t[j - 1] = (uint64_t)c;
            c >>= 64;
        }
        c += t[4];
        t[3] = (uint64_t)c;
        t[4] = t[5] + (uint64_t)(c >> 64);
    }
    HFq r { { t[0], t[1], t[2], t[3] } };
    return (t[4] || hfq_geq_p(r.v)) ? hfq_sub_p(r) : r;
}

bool hfq_is_zero(const HFq& a) { return (a.v[0] | a.v[1] | a.v[2] | a.v[3]) == 0; }

bool hfq_eq(const HFq& a, const HFq& b)
{
    return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2] && a.v[3] == b.v[3];
}

HFq hfq_inv(const HFq& a) // a^(p - 2); Montgomery in, Montgomery out
{
    uint64_t e[4] = { kHP[0] - 2, kHP[1], kHP[2], kHP[3] };
    HFq r { { kHOne[0], kHOne[1], kHOne[2], kHOne[3] } };
    for (int i = 255; i >= 0; --i) {
        r = hfq_mul(r, r);
        if ((e[i >> 6] >> (i & 63)) & 1)
            r = hfq_mul(r, a);
    }
    return r;
}

struct HXyzz {
    HFq X, Y, ZZ, ZZZ; // ZZ = 0: the identity
};

HXyzz hx_dbl(const HXyzz& p)
{
    if (hfq_is_zero(p.ZZ))
        return p;
    const HFq U = hfq_add(p.Y, p.Y);
    const HFq V = hfq_mul(U, U);
    const HFq W = hfq_mul(U, V);
    const HFq S = hfq_mul(p.X, V);
    const HFq X2 = hfq_mul(p.X, p.X);
    const HFq M = hfq_add(hfq_add(X2, X2), X2);
    HXyzz r;
    r.X = hfq_sub(hfq_mul(M, M), hfq_add(S, S));
    r.Y = hfq_sub(hfq_mul(M, hfq_sub(S, r.X)), hfq_mul(W, p.Y));
    r.ZZ = hfq_mul(V, p.ZZ);
    r.ZZZ = hfq_mul(W, p.ZZZ);
    return r;
}

HXyzz hx_add(const HXyzz& p, const HXyzz& q)
{
    if (hfq_is_zero(q.ZZ))
        return p;
    if (hfq_is_zero(p.ZZ))
        return q;
    const HFq U1 = hfq_mul(p.X, q.ZZ), U2 = hfq_mul(q.X, p.ZZ);
    const HFq S1 = hfq_mul(p.Y, q.ZZZ), S2 = hfq_mul(q.Y, p.ZZZ);
    const HFq P = hfq_sub(U2, U1), R = hfq_sub(S2, S1);
    if (hfq_is_zero(P)) {
        if (hfq_is_zero(R))
            return hx_dbl(p);
        return HXyzz {}; // ZZ = 0: the identity
    }
    const HFq PP = hfq_mul(P, P), PPP = hfq_mul(P, PP), Q = hfq_mul(U1, PP);
    HXyzz r;
    r.X = hfq_sub(hfq_sub(hfq_mul(R, R), PPP), hfq_add(Q, Q));
    r.Y = hfq_sub(hfq_mul(R, hfq_sub(Q, r.X)), hfq_mul(S1, PPP));
    r.ZZ = hfq_mul(hfq_mul(p.ZZ, q.ZZ), PP);
    r.ZZZ = hfq_mul(hfq_mul(p.ZZZ, q.ZZZ), PPP);
    return r;
}

HXyzz hx_load(const uint64_t* w)
{
    HXyzz r;
    for (int i = 0; i < 4; ++i) {
        r.X.v[i] = w[i];
        r.Y.v[i] = w[4 + i];
        r.ZZ.v[i] = w[8 + i];
        r.ZZZ.v[i] = w[12 + i];
    }
    return r;
}

// J results of 96 B ((x, y, 1) or (0, 1, 0), Montgomery) from J x Wj window sums
// of 128 B: sum over w of 2^(c w) S_w by Horner.
void host_finalize(const uint64_t* win, uint32_t Wj, uint32_t c, uint32_t J, uint64_t* out)
{
    for (uint32_t j = 0; j < J; ++j) {
        const uint64_t* wj = win + 16ull * Wj * j;
        HXyzz acc = hx_load(wj + 16ull * (Wj - 1));
        for (int w = (int)Wj - 2; w >= 0; --w) {
            for (uint32_t k = 0; k < c; ++k)
                acc = hx_dbl(acc);
            acc = hx_add(acc, hx_load(wj + 16ull * w));
        }
        uint64_t* o = out + 12ull * j;
        memset(o, 0, 96);
        if (hfq_is_zero(acc.ZZ)) {
            memcpy(o + 4, kHOne, 32);
            continue;
        }
        const HFq inv = hfq_inv(hfq_mul(acc.ZZ, acc.ZZZ)); // 1 / (ZZ ZZZ)
        const HFq x = hfq_mul(acc.X, hfq_mul(acc.ZZZ, inv)); // X / ZZ
        const HFq y = hfq_mul(acc.Y, hfq_mul(acc.ZZ, inv)); // Y / ZZZ
        memcpy(o, x.v, 32);
        memcpy(o + 4, y.v, 32);
        memcpy(o + 8, kHOne, 32);
    }
}

// Copies the J x Wj window sums of a finished pipeline to the host, waits, and
// writes the J 96-byte results to `h_out`.
// A single MSM's window sums (at most 64 windows: c >= 4) land in a stack buffer, so the
// single-MSM entry points, the scratch-free one included, allocate nothing on the host
// either; only a batch's J x W sums take a heap buffer.
RustError finish_on_host(const MsmPlan& p, const MsmBuffers& b, void* h_out, hipStream_t stream)
{
    const uint32_t Wj = p.Wb / p.J;
    constexpr uint32_t kStackWindows = 64;
    uint64_t local[16 * kStackWindows];
    std::vector<uint64_t> heap;
    uint64_t* win = local;
    if (p.Wb > kStackWindows) {
        heap.resize(16ull * p.Wb);
        win = heap.data();
    }
    H2_TRY(hipMemcpyAsync(win, b.out, 128ull * p.Wb, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    host_finalize(win, Wj, p.c, p.J, (uint64_t*)h_out);
    return ok();
}

// ------------------------------------------------- precomputed window bases
//
// Row w of the table holds 2^(c w) P_i (affine, 2^261 scale, 64 B), so the
// digit of window w of scalar i selects row w and every window shares one
// bucket set. A row is the previous row doubled c times in XYZZ, then made
// affine with one batched inversion of ZZZ: 1/Z = ZZ/ZZZ, x = X/Z^2, y = Y/ZZZ.

// xyzz[i] = 2^c * row[i]; zzz[i] = ZZZ as a 32-byte Montgomery word (2^256 scale)
__global__ void window_double_kernel(const uint8_t* __restrict__ row, PointSlot* __restrict__ xyzz,
    uint32_t* __restrict__ zzz, uint64_t n, uint32_t c)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    G1X29 acc = G1X29::dbl_affine(G1A29::load(row + 64 * i)); // the identity gives ZZ = ZZZ = 0
    for (uint32_t k = 1; k < c; ++k)
        acc = G1X29::dbl_inline(acc);
    acc.store(xyzz + i);
    Fq29 from;
#pragma unroll
    for (int q = 0; q < 9; ++q)
        from.l[q] = Fq29Params::FROM261[q];
    Fq29::reduce_full(Fq29::mul(acc.ZZZ, from)).store32(zzz + 8 * i);
}

// row[i] = affine(xyzz[i]) with inv[i] = 1/ZZZ (2^256 scale words; 0 for the identity)
__global__ void window_affine_kernel(const PointSlot* __restrict__ xyzz, const uint32_t* __restrict__ inv,
    uint8_t* __restrict__ row, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Fq k;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        k.v[q] = Fq29Params::TO261[q];
    const Fq29 iz3 = Fq29::from_fp(Fq::load(inv + 8 * i) * k); // 1/ZZZ in the 2^261 scale
    const G1X29 p = G1X29::load(xyzz + i);
    const Fq29 iz = Fq29::mul(p.ZZ, iz3); // 1/Z
    const Fq29 iz2 = Fq29::sqr(iz);
    Fq29::reduce_full(Fq29::mul(p.X, iz2)).store32(row + 64 * i);
    Fq29::reduce_full(Fq29::mul(p.Y, iz3)).store32(row + 64 * i + 32);
}

// ------------------------------------------------- Lagrange bases from g
//
// g_lagrange[i] = n^-1 sum_j omega^-ij g_j (an inverse NTT over G1 points, as
// g_to_lagrange does with best_fft): radix-2 decimation in time after a
// bit-reversed load, every twiddle applied as a scalar multiplication of the
// odd input (2-bit windows, most significant first), then the n^-1 scaling and
// one batched inversion for the affine form.

// twiddles: tw[i] = canonical(omega^i), i < count (omega in Montgomery form)
__global__ void ec_twiddle_kernel(Fr omega, uint32_t* __restrict__ tw, uint64_t count)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count)
        return;
    Fr::from_mont(Fr::pow_u64(omega, i)).store(tw + 8 * i);
}

// P[bitrev(i)] = g_i as XYZZ in the 2^261 scale
__global__ void ec_load_kernel(const uint8_t* __restrict__ g, PointSlot* __restrict__ P, uint64_t n, uint32_t k)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Fq t;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        t.v[q] = Fq29Params::TO261[q];
    const Fq x = Fq::load(g + 64 * i) * t, y = Fq::load(g + 64 * i + 32) * t; // (0, 0) stays (0, 0)
    const G1A29 a { Fq29::from_fp(x), Fq29::from_fp(y) };
    const uint64_t r = k ? (uint64_t)(__brevll(i) >> (64 - k)) : 0;
    G1X29::from_affine(a).store(P + r);
}

// s * p for a canonical scalar s (< 2^254), 2-bit windows from the top
__device__ G1X29 scalar_mul(const G1X29& p, const uint32_t (&s)[8])
{
    const G1X29 p2 = G1X29::dbl(p);
    const G1X29 p3 = G1X29::add(p2, p);
    G1X29 acc = G1X29::inf();
    for (int w = 126; w >= 0; --w) {
        acc = G1X29::dbl_inline(G1X29::dbl_inline(acc));
        const uint32_t d = (s[w >> 4] >> ((w & 15) * 2)) & 3u;
        if (d)
            acc = G1X29::add(acc, d == 1 ? p : (d == 2 ? p2 : p3));
    }
    return acc;
}

__device__ __forceinline__ G1X29 g1x_neg(const G1X29& p)
{
    return G1X29 { p.X, nsub2(Fq29::zero(), p.Y), p.ZZ, p.ZZZ };
}

// stage s of the transform: butterflies (a + w b, a - w b) at distance 2^s,
// w = omega^(j 2^(k-1-s)) for position j inside the half block
__global__ void ec_stage_kernel(PointSlot* __restrict__ P, const uint32_t* __restrict__ tw, uint64_t n, uint32_t k,
    uint32_t s)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n / 2)
        return;
    const uint64_t half = 1ull << s;
    const uint64_t j = t & (half - 1);
    const uint64_t i0 = ((t >> s) << (s + 1)) | j;
    const uint64_t i1 = i0 + half;
    const G1X29 a = G1X29::load(P + i0);
    G1X29 b = G1X29::load(P + i1);
    if (j) {
        uint32_t w[8];
        const uint4* q = (const uint4*)(tw + 8 * (j << (k - 1 - s)));
        const uint4 lo = q[0], hi = q[1];
        w[0] = lo.x; w[1] = lo.y; w[2] = lo.z; w[3] = lo.w;
        w[4] = hi.x; w[5] = hi.y; w[6] = hi.z; w[7] = hi.w;
        b = scalar_mul(b, w);
    }
    G1X29::add(a, b).store(P + i0);
    G1X29::add(a, g1x_neg(b)).store(P + i1);
}

// P[i] *= s (canonical scalar) and zzz[i] = ZZZ in the 2^256 scale for the inversion
__global__ void ec_scale_kernel(PointSlot* __restrict__ P, uint32_t* __restrict__ zzz, uint64_t n, Fr s)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const G1X29 r = scalar_mul(G1X29::load(P + i), s.v);
    r.store(P + i);
    Fq29::reduce_full(Fq29::mul(r.ZZZ, Fq29::from_limbs(Fq29P::FROM261))).store32(zzz + 8 * i);
}

// out[i] = affine(P[i]) in the boundary's 2^256 scale, with inv[i] = 1/ZZZ
// (2^256-scale words; 0 for the identity, which gives (0, 0))
__global__ void ec_affine_kernel(const PointSlot* __restrict__ P, const uint32_t* __restrict__ inv,
    uint8_t* __restrict__ out, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    Fq k;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        k.v[q] = Fq29Params::TO261[q];
    const Fq29 iz3 = Fq29::from_fp(Fq::load(inv + 8 * i) * k); // 1/ZZZ in the 2^261 scale
    const G1X29 p = G1X29::load(P + i);
    const Fq29 iz = Fq29::mul(p.ZZ, iz3); // 1/Z
    const Fq29 from = Fq29::from_limbs(Fq29P::FROM261);
    Fq29::reduce_full(Fq29::mul(Fq29::mul(p.X, Fq29::sqr(iz)), from)).store32(out + 64 * i);
    Fq29::reduce_full(Fq29::mul(Fq29::mul(p.Y, iz3), from)).store32(out + 64 * i + 32);
}

unsigned blocks_for(uint64_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

uint32_t ilog2(uint32_t x) { return 31 - __builtin_clz(x); }

// Full device pipeline; the window sums land in b.out (device, for finish_on_host). A batch
// (p.J > 1) has its canonical scalars and job tables in place already and
// passes d_scalars = null.
RustError msm_device(const MsmPlan& p, const MsmBuffers& b, const Fr* d_scalars, const uint8_t* d_points,
    hipStream_t stream)
{
    const unsigned T = 256;
    ProfScope prof_all("msm_total", stream);
    std::optional<ProfScope> prof_sort(std::in_place, "msm_sort", stream); // ends before the buckets
    if (d_scalars && p.glv) {
        glv_split_kernel<<<blocks_for(p.nbase, T), T, 0, stream>>>(d_scalars, b.canon, p.nbase);
        H2_LAUNCH_CHECK();
    } else if (d_scalars) {
        to_canonical_kernel<<<blocks_for(p.n, T), T, 0, stream>>>(d_scalars, b.canon, p.n);
        H2_LAUNCH_CHECK();
    }
    const uint8_t* pts = d_points; // a precomputed table is already in the 2^261 scale
    if (p.glv) {
        glv_bases_kernel<<<blocks_for(p.nbase, T), T, 0, stream>>>(d_points, b.pts, p.nbase);
        H2_LAUNCH_CHECK();
        pts = b.pts;
    } else if (!p.table_n) {
        to_scale261_kernel<<<blocks_for(p.nbase, T), T, 0, stream>>>(d_points, b.pts, p.nbase);
        H2_LAUNCH_CHECK();
        pts = b.pts;
    }
    SortParams sp { p.n, p.c, p.W, p.FB, p.CB, p.NC, p.ntiles, p.table_n, p.NCj, p.J > 1 ? b.jobs : nullptr,
        p.J > 1 ? b.tile_job : nullptr };
    const uint64_t nh1 = (uint64_t)p.NCj * p.ntiles; // every tile counts the bins of its own MSM
    r1_count_kernel<<<p.ntiles, R1_THREADS, 4 * p.NCj, stream>>>(b.canon, sp, b.h1);
    H2_LAUNCH_CHECK();
    RustError e = scan_u32_exclusive(b.h1, b.h1off, nh1, b.scan_scratch, stream);
    if (e.code)
        return e;
    r1_scatter_kernel<<<p.ntiles, R1_THREADS, r1_scatter_lds(p.NCj, p.CB), stream>>>(b.canon, sp, b.h1off, b.kv1);
    H2_LAUNCH_CHECK();
    bin_info_kernel<<<1, 1024, 0, stream>>>(b.h1off, sp, b.bin_start, b.chunk_base, b.off0 + p.nb);
    H2_LAUNCH_CHECK();
    r2_count_kernel<<<(unsigned)p.max_chunks, R2_THREADS, 0, stream>>>(b.kv1, sp, b.bin_start, b.chunk_base, b.h2);
    H2_LAUNCH_CHECK();
    e = scan_u32_exclusive(b.h2, b.h2off, p.max_chunks << p.FB, b.scan_scratch, stream);
    if (e.code)
        return e;
    r2_scatter_kernel<<<(unsigned)p.max_chunks, R2_THREADS, 0, stream>>>(b.kv1, sp, b.bin_start, b.chunk_base,
        b.h2off, b.entries, b.off0);
    H2_LAUNCH_CHECK();
    prof_sort.reset();

    {
        ProfScope prof("msm_buckets", stream);
        accumulate_kernel<<<blocks_for(p.nthr, T), T, 0, stream>>>(b.entries, pts, b.off0, p.nb, p.S, b.part,
            b.bsum);
        H2_LAUNCH_CHECK();
        H2_TRY(hipMemsetAsync(b.heavy + p.nb, 0, 4, stream));
        merge_split_kernel<<<blocks_for(p.nb, T), T, 0, stream>>>(b.part, b.off0, p.nb, p.S, b.bsum, b.heavy);
        H2_LAUNCH_CHECK();
        merge_heavy_kernel<<<HEAVY_BLOCKS, HEAVY_THREADS, 0, stream>>>(b.part, b.off0, p.nb, p.S, b.heavy, b.bsum);
        H2_LAUNCH_CHECK();
    }

    ProfScope prof_red("msm_reduce", stream);
    const uint64_t M0 = p.Mlev[0];
    reduce_first_kernel<<<blocks_for((uint64_t)p.Wb * M0, T), T, 0, stream>>>(b.bsum, b.off0, p.Wb, p.B, p.L1,
        ilog2(p.L1), M0, b.red[0], b.red[1]);
    H2_LAUNCH_CHECK();
    int cur = 0;
    for (uint32_t lev = 1; lev <= p.nlev; ++lev) {
        const uint64_t Mi = p.Mlev[lev - 1];
        const uint64_t Mo = p.Mlev[lev];
        reduce_tree_kernel<<<(unsigned)(p.Wb * Mo), 2 * TREE, 0, stream>>>(b.red[2 * cur], b.red[2 * cur + 1], Mi,
            p.Klev[lev], ilog2(p.Klev[lev]), b.red[2 * (cur ^ 1)], b.red[2 * (cur ^ 1) + 1]);
        H2_LAUNCH_CHECK();
        cur ^= 1;
    }
    export_windows_kernel<<<blocks_for(p.Wb, 64), 64, 0, stream>>>(b.red[2 * cur], p.Wb, b.out);
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
    // two virtual points per base (GLV): indices stay below 2^31 (bit 31 is the sign)
    H2_CHECK_ARG(n < (1ull << 30), "%s: length %lu exceeds 2^30", who, (unsigned long)n);
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
    return finish_on_host(p, b, h_out, stream);
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

// ------------------------------- scratch-free entry point (no reference counterpart)
//
// For callers that cannot allocate device memory (code under graph capture, tight
// serving loops): the MSM's scratch lives in a file-scope device array, one per GPU,
// held until the process ends. Every call on a GPU shares that GPU's array, so calls
// that overlap on one GPU race; the pipeline clears what must start at zero itself.

}  // extern "C"

namespace {

constexpr uint64_t kNoAllocMaxLength = 1ull << 16;
constexpr uint64_t kNoAllocScratchBytes = 64ull << 20; // >= msm_workspace_bytes(2^16) at default widths

__device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device
// (the array's name is a host-side stand-in and must never reach a kernel).
RustError noalloc_scratch(const char* who, void** out)
{
    int dev = 0;
    H2_TRY(hipGetDevice(&dev));
    H2_CHECK_ARG(dev >= 0 && dev < kNoAllocMaxDevices, "%s: device %d beyond the %d supported", who, dev,
        kNoAllocMaxDevices);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* ptr = nullptr;
        H2_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        H2_CHECK_ARG(ptr != nullptr, "%s: no device address for the built-in scratch", who);
        g_noalloc_ptr[dev] = ptr;
    }
    *out = g_noalloc_ptr[dev];
    return ok();
}

} // namespace

extern "C" {

uint64_t h2amd_multiexp_noalloc_max_length(void) { return kNoAllocMaxLength; }

RustError h2amd_multiexp_device_scalars_device_bases_noalloc(const void* d_scalar, const void* d_bases,
    const FFITraitObject* h_output, size_t length, void* stream)
{
    const char* who = "h2amd_multiexp_device_scalars_device_bases_noalloc";
    H2_CHECK_ARG(length <= kNoAllocMaxLength, "%s: length %lu exceeds the built-in scratch's maximum of %lu points",
        who, (unsigned long)length, (unsigned long)kNoAllocMaxLength);
    const uint64_t need = msm_workspace_bytes(length);
    H2_CHECK_ARG(need <= kNoAllocScratchBytes,
        "%s: length %lu needs %lu bytes of scratch at the current window width, the built-in scratch holds %lu", who,
        (unsigned long)length, (unsigned long)need, (unsigned long)kNoAllocScratchBytes);
    void* ws = nullptr;
    RustError e = noalloc_scratch(who, &ws);
    if (e.code)
        return e;
    return msm_entry(who, d_scalar, true, d_bases, true, h_output->ptr, length, ws, kNoAllocScratchBytes,
        (hipStream_t)stream);
}

// ----------------------------------------- fixed bases (no reference counterpart)

uint64_t h2amd_msm_table_bytes(uint64_t n)
{
    const MsmPlan p = make_plan(n, n ? n : 1);
    return 64ull * n * p.W;
}

uint64_t h2amd_msm_precompute_workspace_size(uint64_t n) { return 160 * n + 512; }

RustError h2amd_msm_precompute(void* d_table, const void* d_bases, uint64_t n, void* scratch, uint64_t scratch_bytes,
    void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (n == 0)
        return ok();
    H2_CHECK_ARG(n < (1ull << 31), "h2amd_msm_precompute: length %lu exceeds 2^31", (unsigned long)n);
    const MsmPlan p = make_plan(n, n);
    H2_CHECK_ARG(n * p.W < (1ull << 31), "h2amd_msm_precompute: %lu points x %u windows exceed 2^31 entries",
        (unsigned long)n, p.W);
    Workspace ws(scratch, scratch_bytes);
    PointSlot* xyzz = (PointSlot*)ws.take(128 * n);
    uint32_t* zzz = (uint32_t*)ws.take(32 * n);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "h2amd_msm_precompute: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    uint8_t* table = (uint8_t*)d_table;
    const unsigned T = 256;
    to_scale261_kernel<<<blocks_for(n, T), T, 0, stream>>>((const uint8_t*)d_bases, table, n);
    H2_LAUNCH_CHECK();
    for (uint32_t w = 1; w < p.W; ++w) {
        window_double_kernel<<<blocks_for(n, T), T, 0, stream>>>(table + 64 * n * (w - 1), xyzz, zzz, n, p.c);
        H2_LAUNCH_CHECK();
        RustError e = _halo2_batch_invert(zzz, 1, n, stream);
        if (e.code)
            return e;
        window_affine_kernel<<<blocks_for(n, T), T, 0, stream>>>(xyzz, zzz, table + 64 * n * w, n);
        H2_LAUNCH_CHECK();
    }
    return ok();
}

uint64_t h2amd_multiexp_precomputed_workspace_size(uint64_t length, uint64_t table_n)
{
    return msm_workspace_bytes(length, table_n ? table_n : 1);
}

RustError h2amd_multiexp_precomputed(const void* d_scalar, const void* d_table, uint64_t table_n,
    const FFITraitObject* h_output, uint64_t length, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    uint64_t* o = (uint64_t*)h_output->ptr;
    if (length == 0) {
        memset(o, 0, 96);
        Fq one = Fq::one();
        memcpy(o + 4, one.v, 32);
        return ok();
    }
    H2_CHECK_ARG(length <= table_n, "h2amd_multiexp_precomputed: length %lu exceeds the table's %lu points",
        (unsigned long)length, (unsigned long)table_n);
    const MsmPlan p = make_plan(length, table_n);
    H2_CHECK_ARG(table_n * p.W < (1ull << 31), "h2amd_multiexp_precomputed: table too large");
    Workspace ws(scratch, scratch_bytes);
    MsmBuffers b {};
    carve(ws, p, &b);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue,
            "h2amd_multiexp_precomputed: workspace of %lu bytes is too small for length %lu",
            (unsigned long)scratch_bytes, (unsigned long)length);
    RustError e = msm_device(p, b, (const Fr*)d_scalar, (const uint8_t*)d_table, stream);
    if (e.code)
        return e;
    return finish_on_host(p, b, o, stream);
}

// ------------------------------------------ batched MSMs (no reference counterpart)

uint64_t h2amd_multiexp_batch_workspace_size(const uint64_t* lengths, uint32_t count)
{
    if (count == 0 || count > MAX_JOBS)
        return 0;
    const MsmPlan p = make_batch_plan(lengths, count);
    SizerAdapter sa;
    carve(sa, p, nullptr);
    return sa.s.bytes + 4096;
}

// J MSMs in one pipeline over shared bases (table_n = 0) or one precomputed table
static RustError batch_entry(const char* who, const void* const* d_scalars, const uint64_t* lengths, uint32_t count,
    const void* d_bases, uint64_t table_n, void* h_out, void* scratch, uint64_t scratch_bytes, hipStream_t stream)
{
    H2_CHECK_ARG(count <= MAX_JOBS, "%s: %u MSMs exceed %u per batch", who, count, MAX_JOBS);
    if (count == 0)
        return ok();
    uint64_t* o = (uint64_t*)h_out;
    uint64_t nv = 0;
    for (uint32_t j = 0; j < count; ++j) {
        H2_CHECK_ARG(lengths[j] < (1ull << 31), "%s: length %lu of MSM %u exceeds 2^31", who,
            (unsigned long)lengths[j], j);
        nv += align_tile(lengths[j]);
    }
    if (nv == 0) { // every MSM is empty: identities (0, 1, 0)
        memset(o, 0, 96ull * count);
        const Fq one = Fq::one();
        for (uint32_t j = 0; j < count; ++j)
            memcpy(o + 12 * j + 4, one.v, 32);
        return ok();
    }
    const MsmPlan p = make_batch_plan(lengths, count, table_n);
    H2_CHECK_ARG(p.n * p.W < (1ull << 32), "%s: %lu padded points too many for 32-bit offsets", who,
        (unsigned long)p.n);
    H2_CHECK_ARG(table_n * p.W < (1ull << 31), "%s: table too large", who);
    Workspace ws(scratch, scratch_bytes);
    MsmBuffers b {};
    carve(ws, p, &b);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "%s: workspace of %lu bytes is too small", who,
            (unsigned long)scratch_bytes);
    std::vector<BatchJob> jobs(count);
    std::vector<uint32_t> tile_job(p.ntiles);
    uint64_t voff = 0;
    for (uint32_t j = 0; j < count; ++j) {
        const uint64_t span = align_tile(lengths[j]);
        jobs[j] = BatchJob { voff, lengths[j], (uint32_t)(voff / R1_TILE), (uint32_t)(span / R1_TILE) };
        for (uint32_t t = 0; t < jobs[j].tiles; ++t)
            tile_job[jobs[j].tile0 + t] = j;
        voff += span;
    }
    H2_TRY(hipMemcpyAsync(b.jobs, jobs.data(), sizeof(BatchJob) * count, hipMemcpyHostToDevice, stream));
    H2_TRY(hipMemcpyAsync(b.tile_job, tile_job.data(), 4ull * p.ntiles, hipMemcpyHostToDevice, stream));
    for (uint32_t j = 0; j < count; ++j) {
        if (lengths[j] == 0)
            continue;
        to_canonical_kernel<<<blocks_for(lengths[j], 256), 256, 0, stream>>>((const Fr*)d_scalars[j],
            b.canon + 8 * jobs[j].voff, lengths[j]);
        H2_LAUNCH_CHECK();
    }
    RustError e = msm_device(p, b, nullptr, (const uint8_t*)d_bases, stream);
    if (e.code)
        return e;
    // its synchronisation also keeps the host job tables alive until their copies are done
    return finish_on_host(p, b, o, stream);
}

RustError h2amd_multiexp_batch(const void* const* d_scalars, const uint64_t* lengths, uint32_t count,
    const void* d_bases, void* h_out, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    return batch_entry("h2amd_multiexp_batch", d_scalars, lengths, count, d_bases, 0, h_out, scratch, scratch_bytes,
        (hipStream_t)stream_);
}

uint64_t h2amd_multiexp_batch_precomputed_workspace_size(const uint64_t* lengths, uint32_t count, uint64_t table_n)
{
    if (count == 0 || count > MAX_JOBS || table_n == 0)
        return 0;
    const MsmPlan p = make_batch_plan(lengths, count, table_n);
    SizerAdapter sa;
    carve(sa, p, nullptr);
    return sa.s.bytes + 4096;
}

RustError h2amd_multiexp_batch_precomputed(const void* const* d_scalars, const uint64_t* lengths, uint32_t count,
    const void* d_table, uint64_t table_n, void* h_out, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    H2_CHECK_ARG(table_n > 0, "h2amd_multiexp_batch_precomputed: empty table");
    for (uint32_t j = 0; j < count; ++j)
        H2_CHECK_ARG(lengths[j] <= table_n,
            "h2amd_multiexp_batch_precomputed: length %lu of MSM %u exceeds the table's %lu points",
            (unsigned long)lengths[j], j, (unsigned long)table_n);
    return batch_entry("h2amd_multiexp_batch_precomputed", d_scalars, lengths, count, d_table, table_n, h_out,
        scratch, scratch_bytes, (hipStream_t)stream_);
}

// ----------------------------------- Lagrange bases (no reference counterpart)

uint64_t h2amd_g1_to_lagrange_workspace_size(uint32_t k)
{
    if (k > 28)
        return 0;
    const uint64_t n = 1ull << k;
    WorkspaceSizer w;
    w.take(128 * n);
    w.take(32 * n);
    w.take(32 * (n / 2 ? n / 2 : 1));
    return w.bytes + 1024;
}

RustError h2amd_g1_to_lagrange(void* d_out, const void* d_g, uint32_t k, const void* h_omega_inv, void* scratch,
    uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    H2_CHECK_ARG(k <= 28, "h2amd_g1_to_lagrange: k=%u exceeds 28", k);
    const uint64_t n = 1ull << k;
    Workspace ws(scratch, scratch_bytes);
    PointSlot* P = (PointSlot*)ws.take(128 * n);
    uint32_t* zzz = (uint32_t*)ws.take(32 * n);
    uint32_t* tw = (uint32_t*)ws.take(32 * (n / 2 ? n / 2 : 1));
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "h2amd_g1_to_lagrange: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    Fr omega;
    memcpy(omega.v, h_omega_inv, 32);
    // n^-1 = (2^-1)^k mod r, canonical, on the host
    const uint32_t* M = FrParams::M;
    uint32_t inv[8] = { 1, 0, 0, 0, 0, 0, 0, 0 };
    for (uint32_t i = 0; i < k; ++i) { // inv = inv / 2 mod r
        const bool odd = inv[0] & 1u;
        uint64_t c = 0;
        uint32_t t[9];
        for (int q = 0; q < 8; ++q) {
            c += (uint64_t)inv[q] + (odd ? M[q] : 0u);
            t[q] = (uint32_t)c;
            c >>= 32;
        }
        t[8] = (uint32_t)c;
        for (int q = 0; q < 8; ++q)
            inv[q] = (t[q] >> 1) | (t[q + 1] << 31);
    }
    Fr ninv;
    memcpy(ninv.v, inv, 32);
    const unsigned T = 256;
    if (n > 1) {
        ec_twiddle_kernel<<<blocks_for(n / 2, T), T, 0, stream>>>(omega, tw, n / 2);
        H2_LAUNCH_CHECK();
    }
    ec_load_kernel<<<blocks_for(n, T), T, 0, stream>>>((const uint8_t*)d_g, P, n, k);
    H2_LAUNCH_CHECK();
    for (uint32_t s = 0; s < k; ++s) {
        ec_stage_kernel<<<blocks_for(n / 2, T), T, 0, stream>>>(P, tw, n, k, s);
        H2_LAUNCH_CHECK();
    }
    ec_scale_kernel<<<blocks_for(n, T), T, 0, stream>>>(P, zzz, n, ninv);
    H2_LAUNCH_CHECK();
    RustError e = _halo2_batch_invert(zzz, 1, n, stream);
    if (e.code)
        return e;
    ec_affine_kernel<<<blocks_for(n, T), T, 0, stream>>>(P, zzz, (uint8_t*)d_out, n);
    H2_LAUNCH_CHECK();
    return ok();
}

// ------------------------------------ tuned window widths (no reference counterpart)

RustError h2amd_msm_set_widths(const uint8_t* widths, uint32_t count)
{
    H2_CHECK_ARG(count <= 64, "h2amd_msm_set_widths: %u entries exceed 64", count);
    for (uint32_t i = 0; i < count; ++i)
        H2_CHECK_ARG(widths[i] == 0 || (widths[i] >= 4 && widths[i] <= (i < 32 ? 20 : 22)),
            "h2amd_msm_set_widths: width %u for entry %u outside [4, %u]", widths[i], i, i < 32 ? 20u : 22u);
    for (uint32_t i = 0; i < 64; ++i)
        g_width_table[i].store(i < count ? widths[i] : 0, std::memory_order_relaxed);
    return ok();
}

uint32_t h2amd_msm_width(uint64_t n) { return plain_msm_width(n); }

uint32_t h2amd_msm_table_width(uint64_t table_n) { return plan_width(table_n ? table_n : 1, true); }

} // extern "C"
