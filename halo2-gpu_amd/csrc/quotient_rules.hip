// Custom-gate quotient evaluation: the reference's compiled expression graph
// (a list of 128-bit CalcRules plus value-part rules) interpreted per row on
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
// Unlike the reference, columns are read in place through device pointer
// tables with the rotation wrapped modulo the part length (no transposed
// copy of every column into a row matrix), and intermediates live in LDS
// when the rule count is small, else in a [rule][thread] global buffer.
#include "common.h"
#include "field.h"

using namespace h2;

namespace {

constexpr unsigned kBlock = 64;
constexpr uint32_t kLdsMaxRules = 16; // 16 rules x 64 threads x 32 B = 32 KiB
constexpr uint32_t kGlobalThreads = 1024 * kBlock;

struct QuotientArgs {
    const Fr* const* cols[3]; // fixed, instance, advice pointer tables (device)
    const Fr* constants;
    const Fr* challenges;
    const uint64_t* rules; // 2 words per rule
    const uint64_t* vp_rules;
    uint32_t num_rules, num_vp;
    uint64_t nq; // part length (rotations wrap modulo nq)
    int64_t rot_scale;
    uint64_t row_start, nrows;
    Fr y;
    const Fr* prev; // per-row Horner seed or null
    Fr* out; // nrows
    Fr* interm_global; // [rule][thread] when not in LDS
};

__device__ __forceinline__ int32_t vs_rot(uint64_t v)
{
    const int32_t a = (int32_t)((v >> 24) & 0x7fffu);
    return ((v >> 39) & 1u) ? -a : a;
}

template <bool LDS>
__global__ void __launch_bounds__(kBlock) quotient_rules_kernel(QuotientArgs qa)
{
    extern __shared__ uint4 smem[];
    const uint64_t tid_global = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    Fr* interm;
    uint64_t istride;
    if (LDS) {
        interm = (Fr*)smem + threadIdx.x;
        istride = blockDim.x;
    } else {
        interm = qa.interm_global + tid_global;
        istride = stride;
    }
    for (uint64_t r = tid_global; r < qa.nrows; r += stride) {
        const int64_t row = (int64_t)(qa.row_start + r);
        for (uint32_t i = 0; i < qa.num_rules; ++i) {
            const uint64_t lo = qa.rules[2 * i], hi = qa.rules[2 * i + 1];
            const uint32_t c1 = (uint32_t)(lo >> 40) & 0xfu, d = (uint32_t)(lo >> 44) & 0xfu;
            const uint32_t c2 = (uint32_t)(hi >> 40) & 0xfu;
            Fr v[2];
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                const uint64_t w = o ? hi : lo;
                const uint32_t src = (uint32_t)w & 0xfu;
                const uint32_t idx = (uint32_t)(w >> 4) & 0xfffffu;
                if (src <= 2) {
                    int64_t rr = (row + (int64_t)vs_rot(w) * qa.rot_scale) % (int64_t)qa.nq;
                    if (rr < 0)
                        rr += (int64_t)qa.nq;
                    v[o] = Fr::load(qa.cols[src][idx] + rr);
                } else if (src == 3) {
                    v[o] = interm[(uint64_t)idx * istride];
                } else if (src == 4) {
                    v[o] = Fr::load(qa.constants + idx);
                } else if (src == 5) {
                    v[o] = Fr::load(qa.challenges + idx);
                } else {
                    v[o] = Fr::zero();
                }
            }
            // the (c1, c2, d) combinations the encoder emits; anything else yields 0
            Fr res = Fr::zero();
            if (c1 == 1) {
                if (c2 == 0)
                    res = d == 1 ? v[0] * v[1] : v[0];
                else if (c2 == 1)
                    res = v[0] + v[1];
                else if (c2 == 2)
                    res = v[0] - v[1];
            } else if (c1 == 2) {
                res = Fr::neg(v[0]);
            } else if (c1 == 3) {
                res = Fr::dbl(v[0]);
            }
            interm[(uint64_t)i * istride] = res;
        }
        Fr val = qa.prev ? Fr::load(qa.prev + r) : Fr::zero();
        for (uint32_t i = 0; i < qa.num_vp; ++i) {
            const uint64_t w = qa.vp_rules[i];
            const uint32_t src = (uint32_t)w & 0xfu;
            const uint32_t idx = (uint32_t)(w >> 4) & 0xfffffu;
            const Fr part = src == 4 ? Fr::load(qa.constants + idx) : interm[(uint64_t)idx * istride];
            val = val * qa.y + part;
        }
        val.store(qa.out + r);
    }
}

struct Layout {
    uint64_t rules, vp, consts, chals, tables, interm, staging_out, staging_cols;
};

bool use_lds(uint64_t num_rules) { return num_rules <= kLdsMaxRules; }

// Carves (or sizes) the launcher's workspace in a fixed order.
template <class WS>
void carve(WS& ws, uint64_t num_rules, uint64_t num_vp, uint64_t num_constants, uint64_t num_challenges,
    uint64_t ncols, uint64_t nrows, bool host_out, uint64_t host_col_elems, void** p)
{
    p[0] = ws.take(16 * (num_rules ? num_rules : 1));
    p[1] = ws.take(8 * (num_vp ? num_vp : 1));
    p[2] = ws.take(32 * (num_constants ? num_constants : 1));
    p[3] = ws.take(32 * (num_challenges ? num_challenges : 1));
    p[4] = ws.take(8 * (ncols ? ncols : 1));
    p[5] = use_lds(num_rules) ? nullptr : ws.take(32ull * num_rules * kGlobalThreads);
    p[6] = host_out ? ws.take(32 * nrows) : nullptr;
    p[7] = host_col_elems ? ws.take(32 * host_col_elems) : nullptr;
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
    uint64_t ncols, uint64_t nrows, bool host_out, uint64_t host_col_elems)
{
    Sizer sz;
    void* p[8];
    carve(sz, num_rules, num_vp, num_constants, num_challenges, ncols, nrows, host_out, host_col_elems, p);
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
    Workspace ws(scratch, scratch_bytes);
    void* p[8];
    carve(ws, num_rules, num_vp, num_constants, num_challenges, ncols, nrows, !out_on_device, host_cols, p);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "%s: workspace of %lu bytes is too small", who,
            (unsigned long)scratch_bytes);
    if (num_rules)
        H2_TRY(hipMemcpyAsync(p[0], rules, 16 * num_rules, hipMemcpyHostToDevice, stream));
    if (num_vp)
        H2_TRY(hipMemcpyAsync(p[1], vp_rules, 8 * num_vp, hipMemcpyHostToDevice, stream));
    if (num_constants)
        H2_TRY(hipMemcpyAsync(p[2], constants.ptr, 32 * num_constants, hipMemcpyHostToDevice, stream));
    if (num_challenges)
        H2_TRY(hipMemcpyAsync(p[3], challenges.ptr, 32 * num_challenges, hipMemcpyHostToDevice, stream));
    // device pointer tables: fixed | instance | advice
    const void* tbl[4096];
    H2_CHECK_ARG(ncols <= 4096, "%s: %lu columns exceed 4096", who, (unsigned long)ncols);
    const uint64_t counts[3] = { num_fixed, num_instance, num_advice };
    const void* srcs[3] = { fixed, instance, advices };
    uint64_t k = 0;
    for (int s = 0; s < 3; ++s) {
        for (uint64_t c = 0; c < counts[s]; ++c, ++k) {
            if (kind == COLS_DEVICE) {
                tbl[k] = ((const void* const*)srcs[s])[c];
            } else {
                // host columns are staged into the workspace in table order
                Fr* dst = (Fr*)p[7] + k * num_quotient;
                H2_TRY(hipMemcpyAsync(dst, ((const FFITraitObject*)srcs[s])[c].ptr, 32 * num_quotient,
                    hipMemcpyHostToDevice, stream));
                tbl[k] = dst;
            }
        }
    }
    if (ncols)
        H2_TRY(hipMemcpyAsync(p[4], tbl, 8 * ncols, hipMemcpyHostToDevice, stream));
    // the pageable copy of `tbl` must complete before the stack frame goes away
    H2_TRY(hipStreamSynchronize(stream));

    QuotientArgs qa {};
    const Fr* const* t = (const Fr* const*)p[4];
    qa.cols[0] = t;
    qa.cols[1] = t + num_fixed;
    qa.cols[2] = t + num_fixed + num_instance;
    qa.constants = (const Fr*)p[2];
    qa.challenges = (const Fr*)p[3];
    qa.rules = (const uint64_t*)p[0];
    qa.vp_rules = (const uint64_t*)p[1];
    qa.num_rules = (uint32_t)num_rules;
    qa.num_vp = (uint32_t)num_vp;
    qa.nq = num_quotient;
    qa.rot_scale = (int64_t)rotation_scale;
    qa.row_start = row_start;
    qa.nrows = nrows;
    memcpy(qa.y.v, (const uint8_t*)expr_constants.ptr + 4 * 32, 32); // expr_constants = [0, 1, -1, 2, y]
    qa.prev = (const Fr*)prev_device;
    qa.out = out_on_device ? (Fr*)out : (Fr*)p[6];
    qa.interm_global = (Fr*)p[5];
    if (use_lds(num_rules)) {
        uint64_t blocks = (nrows + kBlock - 1) / kBlock;
        if (blocks > 4096)
            blocks = 4096;
        quotient_rules_kernel<true><<<(unsigned)blocks, kBlock, 32u * kBlock * (num_rules ? num_rules : 1), stream>>>(
            qa);
    } else {
        quotient_rules_kernel<false><<<kGlobalThreads / kBlock, kBlock, 0, stream>>>(qa);
    }
    H2_LAUNCH_CHECK();
    if (!out_on_device) {
        H2_TRY(hipMemcpyAsync(out, p[6], 32 * nrows, hipMemcpyDeviceToHost, stream));
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
    return workspace_bytes(num_rules, num_vp_rules, num_constants, num_challenges, ncols, row_end - row_start, true,
        ncols * num_quotient);
}

uint64_t _halo2_evaluate_h_max_rows(const uint64_t* rules, uint64_t num_rules, uint64_t num_vp_rules,
    uint64_t num_quotient, uint64_t num_fixed, uint64_t num_instance, uint64_t num_advice, uint64_t num_challenges,
    uint64_t num_constants, uint64_t rotation_scale, uint64_t row_start, uint64_t row_end, uint64_t free_bytes)
{
    (void)rules;
    (void)rotation_scale;
    uint64_t rows = row_end > row_start ? row_end - row_start : 0;
    const uint64_t ncols = num_fixed + num_instance + num_advice;
    while (rows > 0) {
        if (workspace_bytes(num_rules, num_vp_rules, num_constants, num_challenges, ncols, rows, true, 0) < free_bytes)
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
