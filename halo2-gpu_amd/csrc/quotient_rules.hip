// Custom-gate quotient evaluation: the reference's compiled expression graph
// (a list of 128-bit CalcRules plus value-part rules) evaluated per row on
// one extended-domain part, folded into h(X) with Horner in y. Mirrors
// cuda/src/quotient.cu:309-846 and kernel/quotient.h:57-303; the encoding is
// the one src/plonk/evaluation.rs:107-350 emits:
//
//   ValueSource (40 bits): src:4 | idx:20 | |rot|:15 | sign:1
//       src: Fixed=0 Instance=1 Advice=2 Intermediate=3 Constant=4 Challenge=5 Dummy=6
//   CalcRule (128 bits): low  = a:40 | c1:4 | d:4 ;  high = b:40 | c2:4
//       result = (C1 a + C2 b) * (d ? b : 1),  C = [0, 1, -1, 2]
//   value-part rule (64 bits): a ValueSource, Constant or Intermediate.
//
// Unlike the reference, the host first compiles the rules into a short
// program for the device: every operand is resolved to its final form (a
// column base pointer with the rotation already reduced modulo the part
// length, an index into one constant table, or a register slot), rules whose
// result nothing reads are dropped, the Horner steps are interleaved with the
// rules as soon as their part is ready, and intermediates get slots by a
// linear scan over their live ranges, so a gate needs a few slots instead of
// one per rule. The program is read with uniform (scalar) loads; columns are
// read in place (no transposed row matrix); slots live in LDS when they fit,
// else in a [slot][thread] global buffer.
#include <mutex>
#include <vector>

#include "common.h"
#include "field.h"

using namespace h2;

namespace {

constexpr unsigned kBlock = 64;
constexpr uint32_t kLdsMaxSlots = 16; // 16 slots x 64 threads x 32 B = 32 KiB
constexpr uint32_t kTwoRowSlots = 6; // two rows per thread up to this many slots
constexpr uint32_t kGlobalThreads = 1024 * kBlock;

enum : uint32_t {
    OP_ZERO = 0,
    OP_COPY, // a
    OP_MUL, // a * b
    OP_ADD, // a + b
    OP_SUB, // a - b
    OP_NEG, // -a
    OP_DBL, // 2a
    OP_HORNER, // acc = acc * y + a
    OP_HORNER_FIRST, // acc = a (first Horner step with no seed)
};
enum : uint32_t { K_ZERO = 0, K_COL, K_SLOT, K_CONST };

// One program step, 64 bytes (one scalar load of 16 dwords).
struct alignas(64) Instr {
    uint32_t op, dst, ka, kb;
    uint32_t ia, ib, pad0, pad1; // slot or constant-table index per operand
    const Fr* pa; // column base (K_COL)
    uint64_t ra; // rotation modulo the part length, in [0, nq)
    const Fr* pb;
    uint64_t rb;
};
static_assert(sizeof(Instr) == 64, "Instr is one 64-byte record");

struct RunArgs {
    uint64_t nq; // part length
    uint64_t row_start, nrows;
    Fr y;
    const Fr* prev; // per-row Horner seed or null
    Fr* out; // nrows
    uint4* slab; // [slot][half][thread] when the slots are not in LDS
    uint32_t ninstr;
};

template <int R, bool LDS>
__device__ __forceinline__ Fr fetch(uint32_t kind, uint32_t idx, const Fr* p, uint64_t rot, uint64_t row, uint64_t nq,
    const uint4* slots, uint64_t sstride, const Fr* __restrict__ cvals, int j)
{
    if (kind == K_COL) {
        uint64_t rr = row + rot; // row, rot < nq
        if (rr >= nq)
            rr -= nq;
        return Fr::load(p + rr);
    }
    if (kind == K_SLOT) {
        const uint4* s = slots + (uint64_t)((idx * R + j) * 2) * sstride;
        const uint4 lo = s[0], hi = s[sstride];
        Fr r;
        r.v[0] = lo.x; r.v[1] = lo.y; r.v[2] = lo.z; r.v[3] = lo.w;
        r.v[4] = hi.x; r.v[5] = hi.y; r.v[6] = hi.z; r.v[7] = hi.w;
        return r;
    }
    if (kind == K_CONST)
        return Fr::load(cvals + idx);
    return Fr::zero();
}

// R rows per thread (row j of a tile at j * kBlock + lane, coalesced per j).
template <int R, bool LDS>
__global__ void __launch_bounds__(kBlock) quotient_program_kernel(
    const Instr* __restrict__ prog, const Fr* __restrict__ cvals, RunArgs ra)
{
    extern __shared__ uint4 smem[];
    uint4* slots;
    uint64_t sstride;
    if (LDS) {
        slots = smem + threadIdx.x;
        sstride = kBlock;
    } else {
        slots = ra.slab + (uint64_t)blockIdx.x * kBlock + threadIdx.x;
        sstride = (uint64_t)gridDim.x * kBlock;
    }
    const uint64_t tile = (uint64_t)R * kBlock;
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < ra.nrows; base += (uint64_t)gridDim.x * tile) {
        uint64_t rel[R], row[R];
        Fr acc[R];
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint64_t r = base + (uint64_t)j * kBlock + threadIdx.x;
            rel[j] = r < ra.nrows ? r : ra.nrows - 1; // rows past the end recompute the last one
            row[j] = ra.row_start + rel[j];
            acc[j] = ra.prev ? Fr::load(ra.prev + rel[j]) : Fr::zero();
        }
        for (uint32_t pc = 0; pc < ra.ninstr; ++pc) {
            const Instr& in = prog[pc];
            const uint32_t op = in.op;
            const bool two = op == OP_MUL || op == OP_ADD || op == OP_SUB;
            Fr a[R], b[R];
#pragma unroll
            for (int j = 0; j < R; ++j)
                a[j] = fetch<R, LDS>(in.ka, in.ia, in.pa, in.ra, row[j], ra.nq, slots, sstride, cvals, j);
            if (two) {
#pragma unroll
                for (int j = 0; j < R; ++j)
                    b[j] = fetch<R, LDS>(in.kb, in.ib, in.pb, in.rb, row[j], ra.nq, slots, sstride, cvals, j);
            }
#pragma unroll
            for (int j = 0; j < R; ++j) {
                Fr res;
                if (op == OP_MUL || op == OP_HORNER) {
                    // one multiplier for both: a * b, or acc * y
                    const Fr x = op == OP_MUL ? a[j] : acc[j];
                    const Fr z = op == OP_MUL ? b[j] : ra.y;
                    res = x * z;
                    if (op == OP_HORNER)
                        acc[j] = res + a[j];
                } else if (op == OP_ADD) {
                    res = a[j] + b[j];
                } else if (op == OP_SUB) {
                    res = a[j] - b[j];
                } else if (op == OP_NEG) {
                    res = Fr::neg(a[j]);
                } else if (op == OP_DBL) {
                    res = Fr::dbl(a[j]);
                } else if (op == OP_COPY) {
                    res = a[j];
                } else if (op == OP_HORNER_FIRST) {
                    acc[j] = a[j];
                } else {
                    res = Fr::zero();
                }
                if (op < OP_HORNER) {
                    uint4* s = slots + (uint64_t)((in.dst * R + j) * 2) * sstride;
                    s[0] = make_uint4(res.v[0], res.v[1], res.v[2], res.v[3]);
                    s[sstride] = make_uint4(res.v[4], res.v[5], res.v[6], res.v[7]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < R; ++j) {
            const uint64_t r = base + (uint64_t)j * kBlock + threadIdx.x;
            if (r < ra.nrows)
                acc[j].store(ra.out + r);
        }
    }
}

struct Program {
    std::vector<Instr> code;
    uint32_t slots = 0;
};

struct Operand {
    uint32_t src, idx;
    int32_t rot;
};

Operand operand(uint64_t w)
{
    Operand o;
    o.src = (uint32_t)w & 0xfu;
    o.idx = (uint32_t)(w >> 4) & 0xfffffu;
    const int32_t a = (int32_t)((w >> 24) & 0x7fffu);
    o.rot = ((w >> 39) & 1u) ? -a : a;
    return o;
}

// The (c1, c2, d) combinations the encoder emits; anything else yields 0.
uint32_t rule_op(uint64_t lo, uint64_t hi)
{
    const uint32_t c1 = (uint32_t)(lo >> 40) & 0xfu, d = (uint32_t)(lo >> 44) & 0xfu;
    const uint32_t c2 = (uint32_t)(hi >> 40) & 0xfu;
    if (c1 == 1) {
        if (c2 == 0)
            return d == 1 ? OP_MUL : OP_COPY;
        if (c2 == 1)
            return OP_ADD;
        if (c2 == 2)
            return OP_SUB;
        return OP_ZERO;
    }
    if (c1 == 2)
        return OP_NEG;
    if (c1 == 3)
        return OP_DBL;
    return OP_ZERO;
}

uint32_t op_arity(uint32_t op)
{
    return (op == OP_MUL || op == OP_ADD || op == OP_SUB) ? 2 : (op == OP_ZERO ? 0 : 1);
}

// Compiles the rules; fails on an operand that names a column, constant,
// challenge or intermediate that does not exist (intermediates must come from
// an earlier rule).
RustError compile(const char* who, const uint64_t* rules, uint64_t num_rules, const uint64_t* vp, uint64_t num_vp,
    const std::vector<const Fr*>& cols, const uint64_t counts[3], uint64_t num_constants, uint64_t num_challenges,
    uint64_t nq, uint64_t rot_scale, bool has_prev, Program& prog)
{
    std::vector<uint32_t> ops(num_rules);
    std::vector<uint8_t> live(num_rules, 0);
    auto check = [&](const Operand& o, uint64_t rule) -> RustError {
        if (o.src <= 2 && o.idx >= counts[o.src])
            return fail((int)hipErrorInvalidValue, "%s: rule %lu reads column %u of %lu (kind %u)", who,
                (unsigned long)rule, o.idx, (unsigned long)counts[o.src], o.src);
        if (o.src == 3 && o.idx >= rule)
            return fail((int)hipErrorInvalidValue, "%s: rule %lu reads intermediate %u before it is computed", who,
                (unsigned long)rule, o.idx);
        if (o.src == 4 && o.idx >= num_constants)
            return fail((int)hipErrorInvalidValue, "%s: rule %lu reads constant %u of %lu", who,
                (unsigned long)rule, o.idx, (unsigned long)num_constants);
        if (o.src == 5 && o.idx >= num_challenges)
            return fail((int)hipErrorInvalidValue, "%s: rule %lu reads challenge %u of %lu", who,
                (unsigned long)rule, o.idx, (unsigned long)num_challenges);
        return ok();
    };
    for (uint64_t i = 0; i < num_rules; ++i) {
        ops[i] = rule_op(rules[2 * i], rules[2 * i + 1]);
        for (uint32_t o = 0; o < op_arity(ops[i]); ++o) {
            RustError e = check(operand(rules[2 * i + o]), i);
            if (e.code)
                return e;
        }
    }
    for (uint64_t h = 0; h < num_vp; ++h) {
        const Operand o = operand(vp[h]);
        if (o.src == 4) {
            if (o.idx >= num_constants)
                return fail((int)hipErrorInvalidValue, "%s: value part %lu reads constant %u of %lu", who,
                    (unsigned long)h, o.idx, (unsigned long)num_constants);
        } else if (o.idx >= num_rules) {
            return fail((int)hipErrorInvalidValue, "%s: value part %lu reads intermediate %u of %lu", who,
                (unsigned long)h, o.idx, (unsigned long)num_rules);
        } else {
            live[o.idx] = 1;
        }
    }
    for (uint64_t i = num_rules; i-- > 0;) {
        if (!live[i])
            continue;
        for (uint32_t o = 0; o < op_arity(ops[i]); ++o) {
            const Operand x = operand(rules[2 * i + o]);
            if (x.src == 3)
                live[x.idx] = 1;
        }
    }
    // emission order: live rules, each followed by the Horner steps it makes ready
    struct Step {
        bool horner;
        uint64_t id; // rule or value part
    };
    std::vector<Step> steps;
    uint64_t h = 0;
    auto ready = [&](uint64_t upto) { // value parts whose intermediate index < upto
        while (h < num_vp) {
            const Operand o = operand(vp[h]);
            if (o.src != 4 && o.idx >= upto)
                break;
            steps.push_back({ true, h++ });
        }
    };
    ready(0);
    for (uint64_t i = 0; i < num_rules; ++i) {
        if (live[i])
            steps.push_back({ false, i });
        ready(i + 1);
    }
    // last step reading each intermediate
    std::vector<uint64_t> last(num_rules, 0);
    auto reads = [&](const Step& s, uint32_t out[2]) -> uint32_t { // intermediates read by a step
        uint32_t n = 0;
        if (s.horner) {
            const Operand o = operand(vp[s.id]);
            if (o.src != 4)
                out[n++] = o.idx;
        } else {
            for (uint32_t o = 0; o < op_arity(ops[s.id]); ++o) {
                const Operand x = operand(rules[2 * s.id + o]);
                if (x.src == 3)
                    out[n++] = x.idx;
            }
        }
        return n;
    };
    for (uint64_t k = 0; k < steps.size(); ++k) {
        uint32_t r[2];
        const uint32_t n = reads(steps[k], r);
        for (uint32_t t = 0; t < n; ++t)
            last[r[t]] = k;
    }
    // linear-scan slot assignment
    std::vector<uint32_t> slot_of(num_rules, 0);
    std::vector<uint8_t> busy;
    bool first_horner = true;
    prog.code.clear();
    prog.code.reserve(steps.size());
    prog.slots = 0;
    auto encode = [&](uint64_t w, uint32_t& kind, uint32_t& idx, const Fr*& p, uint64_t& rot) {
        const Operand o = operand(w);
        kind = K_ZERO;
        idx = 0;
        p = nullptr;
        rot = 0;
        if (o.src <= 2) {
            kind = K_COL;
            const uint64_t base = o.src == 0 ? 0 : (o.src == 1 ? counts[0] : counts[0] + counts[1]);
            p = cols[base + o.idx];
            __int128 off = (__int128)o.rot * (__int128)rot_scale % (__int128)nq;
            if (off < 0)
                off += nq;
            rot = (uint64_t)off;
        } else if (o.src == 3) {
            kind = K_SLOT;
            idx = slot_of[o.idx];
        } else if (o.src == 4) {
            kind = K_CONST;
            idx = o.idx;
        } else if (o.src == 5) {
            kind = K_CONST;
            idx = (uint32_t)num_constants + o.idx;
        }
    };
    for (uint64_t k = 0; k < steps.size(); ++k) {
        Instr in {};
        const Step& s = steps[k];
        if (s.horner) {
            in.op = (first_horner && !has_prev) ? OP_HORNER_FIRST : OP_HORNER;
            first_horner = false;
            const Operand o = operand(vp[s.id]);
            if (o.src == 4) {
                in.ka = K_CONST;
                in.ia = o.idx;
            } else {
                in.ka = K_SLOT;
                in.ia = slot_of[o.idx];
            }
        } else {
            in.op = ops[s.id];
            const uint32_t ar = op_arity(in.op);
            if (ar >= 1)
                encode(rules[2 * s.id], in.ka, in.ia, in.pa, in.ra);
            if (ar == 2)
                encode(rules[2 * s.id + 1], in.kb, in.ib, in.pb, in.rb);
        }
        uint32_t r[2];
        const uint32_t n = reads(s, r);
        for (uint32_t t = 0; t < n; ++t)
            if (last[r[t]] == k)
                busy[slot_of[r[t]]] = 0;
        if (!s.horner) {
            uint32_t d = 0;
            while (d < busy.size() && busy[d])
                ++d;
            if (d == busy.size())
                busy.push_back(0);
            busy[d] = 1;
            slot_of[s.id] = d;
            in.dst = d;
            if (busy.size() > prog.slots)
                prog.slots = (uint32_t)busy.size();
        }
        prog.code.push_back(in);
    }
    return ok();
}

bool use_lds(uint32_t slots) { return slots <= kLdsMaxSlots; }

// Carves (or sizes) the launcher's workspace in a fixed order. `slots` is an
// upper bound of the program's slots (the rule count when sizing).
template <class WS>
void carve(WS& ws, uint64_t num_instr, uint64_t num_values, uint64_t slots, uint64_t nrows, bool host_out,
    uint64_t host_col_elems, void** p)
{
    p[0] = ws.take(sizeof(Instr) * (num_instr ? num_instr : 1));
    p[1] = ws.take(32 * (num_values ? num_values : 1));
    p[2] = slots <= kLdsMaxSlots ? nullptr : ws.take(32ull * slots * kGlobalThreads);
    p[3] = host_out ? ws.take(32 * nrows) : nullptr;
    p[4] = host_col_elems ? ws.take(32 * host_col_elems) : nullptr;
}

struct Sizer {
    WorkspaceSizer s;
    void* take(uint64_t n)
    {
        s.take(n);
        return nullptr;
    }
};

uint64_t workspace_bytes(uint64_t num_rules, uint64_t num_vp, uint64_t num_constants, uint64_t num_challenges,
    uint64_t nrows, bool host_out, uint64_t host_col_elems)
{
    Sizer sz;
    void* p[5];
    carve(sz, num_rules + num_vp, num_constants + num_challenges, num_rules, nrows, host_out, host_col_elems, p);
    return sz.s.bytes + 4096;
}

enum ColumnKind { COLS_HOST, COLS_DEVICE };

RustError quotient_impl(const char* who, ColumnKind kind, const void* fixed, uint64_t num_fixed, const void* instance,
    uint64_t num_instance, const void* advices, uint64_t num_advice, FFITraitObject challenges, uint64_t num_challenges,
    FFITraitObject constants, uint64_t num_constants, FFITraitObject expr_constants, const uint64_t* rules,
    uint64_t num_rules, const uint64_t* vp_rules, uint64_t num_vp, void* out, bool out_on_device, uint64_t num_quotient,
    const void* prev_device, uint64_t rotation_scale, uint64_t row_start, uint64_t row_end, void* scratch,
    uint64_t scratch_bytes, hipStream_t stream)
{
    H2_CHECK_ARG(row_start <= row_end && row_end <= num_quotient, "%s: bad row range [%lu, %lu) of %lu", who,
        (unsigned long)row_start, (unsigned long)row_end, (unsigned long)num_quotient);
    H2_CHECK_ARG(num_rules < (1u << 20) && num_vp < (1u << 20), "%s: too many rules", who);
    const uint64_t nrows = row_end - row_start;
    if (nrows == 0)
        return ok();
    const uint64_t ncols = num_fixed + num_instance + num_advice;
    const uint64_t host_cols = kind == COLS_HOST ? ncols * num_quotient : 0;
    const uint64_t nvalues = num_constants + num_challenges;
    Workspace ws(scratch, scratch_bytes);
    void* p[5];
    // sized with the rule count as the slot bound, as the sizing entry point does
    carve(ws, num_rules + num_vp, nvalues, num_rules, nrows, !out_on_device, host_cols, p);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "%s: workspace of %lu bytes is too small", who,
            (unsigned long)scratch_bytes);
    // column bases in program order: fixed | instance | advice
    std::vector<const Fr*> cols(ncols);
    const uint64_t counts[3] = { num_fixed, num_instance, num_advice };
    const void* srcs[3] = { fixed, instance, advices };
    uint64_t k = 0;
    for (int s = 0; s < 3; ++s) {
        for (uint64_t c = 0; c < counts[s]; ++c, ++k) {
            if (kind == COLS_DEVICE) {
                cols[k] = (const Fr*)((const void* const*)srcs[s])[c];
            } else {
                // host columns are staged into the workspace in table order
                Fr* dst = (Fr*)p[4] + k * num_quotient;
                H2_TRY(hipMemcpyAsync(dst, ((const FFITraitObject*)srcs[s])[c].ptr, 32 * num_quotient,
                    hipMemcpyHostToDevice, stream));
                cols[k] = dst;
            }
        }
    }
    Program prog;
    {
        RustError e = compile(who, rules, num_rules, vp_rules, num_vp, cols, counts, num_constants, num_challenges,
            num_quotient, rotation_scale, prev_device != nullptr, prog);
        if (e.code)
            return e;
    }
    std::vector<Fr> values(nvalues);
    if (num_constants)
        memcpy(values.data(), constants.ptr, 32 * num_constants);
    if (num_challenges)
        memcpy(values.data() + num_constants, challenges.ptr, 32 * num_challenges);
    {
        RustError e = stage_upload(p[0], prog.code.data(), sizeof(Instr) * prog.code.size(), p[1], values.data(),
            32 * nvalues, stream);
        if (e.code)
            return e;
    }

    RunArgs ra {};
    ra.nq = num_quotient;
    ra.row_start = row_start;
    ra.nrows = nrows;
    memcpy(ra.y.v, (const uint8_t*)expr_constants.ptr + 4 * 32, 32); // expr_constants = [0, 1, -1, 2, y]
    ra.prev = (const Fr*)prev_device;
    ra.out = out_on_device ? (Fr*)out : (Fr*)p[3];
    ra.slab = (uint4*)p[2];
    ra.ninstr = (uint32_t)prog.code.size();
    const Instr* d_prog = (const Instr*)p[0];
    const Fr* d_values = (const Fr*)p[1];
    const uint32_t slots = prog.slots ? prog.slots : 1;
    if (use_lds(prog.slots)) {
        const int rows = prog.slots <= kTwoRowSlots ? 2 : 1;
        uint64_t blocks = (nrows + rows * kBlock - 1) / (rows * kBlock);
        if (blocks > 65536)
            blocks = 65536;
        const unsigned lds = 32u * kBlock * rows * slots;
        if (rows == 2)
            quotient_program_kernel<2, true><<<(unsigned)blocks, kBlock, lds, stream>>>(d_prog, d_values, ra);
        else
            quotient_program_kernel<1, true><<<(unsigned)blocks, kBlock, lds, stream>>>(d_prog, d_values, ra);
    } else {
        quotient_program_kernel<1, false><<<kGlobalThreads / kBlock, kBlock, 0, stream>>>(d_prog, d_values, ra);
    }
    H2_LAUNCH_CHECK();
    if (!out_on_device) {
        H2_TRY(hipMemcpyAsync(out, p[3], 32 * nrows, hipMemcpyDeviceToHost, stream));
        H2_TRY(hipStreamSynchronize(stream));
    }
    return ok();
}

} // namespace

extern "C" {

uint64_t _halo2_quotient_workspace_size(const uint64_t* rules, uint64_t num_rules, uint64_t num_vp_rules,
    uint64_t num_quotient, uint64_t num_fixed, uint64_t num_instance, uint64_t num_advice, uint64_t num_challenges,
    uint64_t num_constants, uint64_t rotation_scale, uint64_t row_start, uint64_t row_end)
{
    (void)rules;
    (void)rotation_scale;
    if (row_end < row_start)
        return 0;
    const uint64_t ncols = num_fixed + num_instance + num_advice;
    // covers all three launchers: host columns staged, host output staged
    return workspace_bytes(num_rules, num_vp_rules, num_constants, num_challenges, row_end - row_start, true,
        ncols * num_quotient);
}

uint64_t _halo2_evaluate_h_max_rows(const uint64_t* rules, uint64_t num_rules, uint64_t num_vp_rules,
    uint64_t num_quotient, uint64_t num_fixed, uint64_t num_instance, uint64_t num_advice, uint64_t num_challenges,
    uint64_t num_constants, uint64_t rotation_scale, uint64_t row_start, uint64_t row_end, uint64_t free_bytes)
{
    (void)rules;
    (void)rotation_scale;
    uint64_t rows = row_end > row_start ? row_end - row_start : 0;
    (void)num_fixed;
    (void)num_instance;
    (void)num_advice;
    while (rows > 0) {
        if (workspace_bytes(num_rules, num_vp_rules, num_constants, num_challenges, rows, true, 0) < free_bytes)
            return rows;
        rows >>= 1;
    }
    (void)num_quotient;
    return 0;
}

RustError _halo2_quotient(const FFITraitObject* fixed, uint64_t num_fixed, const FFITraitObject* instance,
    uint64_t num_instance, const FFITraitObject* advices, uint64_t num_advice, FFITraitObject challenges,
    uint64_t num_challenges, FFITraitObject constants, uint64_t num_constants, FFITraitObject expr_constants,
    const uint64_t* rules, uint64_t num_rules, const uint64_t* value_part_rules, uint64_t num_vp_rules,
    FFITraitObject quotient_poly, uint64_t num_quotient, uint64_t rotation_scale, uint64_t row_start,
    uint64_t row_end, void* scratch, uint64_t scratch_bytes, void* stream)
{
    return quotient_impl("_halo2_quotient", COLS_HOST, fixed, num_fixed, instance, num_instance, advices, num_advice,
        challenges, num_challenges, constants, num_constants, expr_constants, rules, num_rules, value_part_rules,
        num_vp_rules, quotient_poly.ptr, false, num_quotient, nullptr, rotation_scale, row_start, row_end, scratch,
        scratch_bytes, (hipStream_t)stream);
}

RustError _halo2_quotient_device_columns(const void* const* fixed_d_ptrs, uint64_t num_fixed,
    const void* const* instance_d_ptrs, uint64_t num_instance, const void* const* advice_d_ptrs, uint64_t num_advice,
    FFITraitObject challenges, uint64_t num_challenges, FFITraitObject constants, uint64_t num_constants,
    FFITraitObject expr_constants, const uint64_t* rules, uint64_t num_rules, const uint64_t* value_part_rules,
    uint64_t num_vp_rules, FFITraitObject quotient_poly, uint64_t num_quotient, uint64_t rotation_scale,
    uint64_t row_start, uint64_t row_end, void* scratch, uint64_t scratch_bytes, void* stream)
{
    return quotient_impl("_halo2_quotient_device_columns", COLS_DEVICE, fixed_d_ptrs, num_fixed, instance_d_ptrs,
        num_instance, advice_d_ptrs, num_advice, challenges, num_challenges, constants, num_constants, expr_constants,
        rules, num_rules, value_part_rules, num_vp_rules, quotient_poly.ptr, false, num_quotient, nullptr,
        rotation_scale, row_start, row_end, scratch, scratch_bytes, (hipStream_t)stream);
}

RustError _halo2_quotient_device_columns_device_out(const void* const* fixed_d_ptrs, uint64_t num_fixed,
    const void* const* instance_d_ptrs, uint64_t num_instance, const void* const* advice_d_ptrs, uint64_t num_advice,
    FFITraitObject challenges, uint64_t num_challenges, FFITraitObject constants, uint64_t num_constants,
    FFITraitObject expr_constants, const uint64_t* rules, uint64_t num_rules, const uint64_t* value_part_rules,
    uint64_t num_vp_rules, void* quotient_poly_device_ptr, uint64_t num_quotient, const void* prev_values_device_ptr,
    uint64_t rotation_scale, uint64_t row_start, uint64_t row_end, void* scratch, uint64_t scratch_bytes,
    void* stream)
{
    return quotient_impl("_halo2_quotient_device_columns_device_out", COLS_DEVICE, fixed_d_ptrs, num_fixed,
        instance_d_ptrs, num_instance, advice_d_ptrs, num_advice, challenges, num_challenges, constants, num_constants,
        expr_constants, rules, num_rules, value_part_rules, num_vp_rules, quotient_poly_device_ptr, true, num_quotient,
        prev_values_device_ptr, rotation_scale, row_start, row_end, scratch, scratch_bytes, (hipStream_t)stream);
}

} // extern "C"
