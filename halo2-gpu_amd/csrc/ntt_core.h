// Fr NTT engine shared by every NTT entry point (single, coset, many, batched).
//
// Algorithm: Stockham-ordered four-step decomposition N = R_1 * ... * R_P
// (P = 1..3 passes, R_p <= 2^12). Pass p reads R_p-point columns with stride
// N/R_p, runs a radix-2 decimation-in-time sub-FFT of the columns in LDS
// (bit-reversed load, natural-order result), multiplies by the inter-pass
// twiddle and writes with the autosort index map, so the output of the last
// pass is in natural order without a bit-reversal pass. The last pass is
// in-place safe (it writes exactly the positions it read), so an NTT needs at
// most one scratch buffer of N elements and none when input and output differ
// and P <= 2.
//
// Twiddles never come from a dense 2^(k-1) table: omega^e = lo[e mod 2^h] *
// hi[e >> h] with two tables of ~2^(k/2) entries that stay resident in L2, and
// each pass radix has a table of R/2 entries.
#pragma once

#include "common.h"
#include "field.h"

namespace h2 {

enum NttTwist : int {
    TWIST_NONE = 0,
    TWIST_ZETA = 1, // x_i *= zeta^(i mod 3)                       (cosetFFT)
    TWIST_POWERS = 2, // x_i *= c^i                                  (CosetFFT_Part)
};
enum NttPost : int {
    POST_NONE = 0,
    POST_SCALE = 1, // y_i *= s                                       (iFFT)
    POST_SCALE_ZETA_INV = 2, // y_i *= s * zeta^-(i mod 3)            (icosetFFT)
};

struct NttPlan {
    int P;
    uint32_t log_r[3];
    uint32_t log_g[3];
};

static constexpr uint32_t NTT_MAX_LOG_R = 12; // 4096 elements = 128 KiB of LDS per work-group
static constexpr uint32_t NTT_MAX_LOG = 28;
static constexpr int NTT_MAX_BATCH = 32; // polynomials per launch (pointers travel as kernel args)

NttPlan ntt_plan(uint32_t log_n);

struct NttTables {
    const Fr* tw_lo = nullptr; // omega^i, i < 2^h
    const Fr* tw_hi = nullptr; // omega^(i 2^h)
    uint32_t h = 0;
    const Fr* twr[3] = { nullptr, nullptr, nullptr }; // omega_{R_p}^x, x < R_p / 2
    const Fr* c_lo = nullptr; // c^i, i < 2^h
    const Fr* c_hi = nullptr; // c^(i 2^h)
};

// Workspace bytes for the tables of one transform size (+ optional c tables).
void ntt_tables_size(WorkspaceSizer& ws, uint32_t log_n, bool powers_twist);

// Builds the tables in `ws`. `d_omega` is a device pointer to omega; `c` is the
// CosetFFT_Part factor (used only with powers_twist).
RustError ntt_tables_build(Workspace& ws, NttTables& t, uint32_t log_n, const Fr* d_omega, bool powers_twist,
    const Fr& c, hipStream_t stream);

struct NttJob {
    const Fr* in;
    Fr* out;
};

// Runs `count` transforms of size 2^log_n that share the tables. `in_len` is
// the number of valid input elements (the rest read as zero, for the coset
// extension). `scratch` holds `scratch_slots` buffers of 2^log_n elements (or
// is null); one slot per concurrently batched job is needed when P == 3, or
// when P == 2 and some job has in == out. Jobs are launched in batches of at
// most NTT_MAX_BATCH (and at most scratch_slots when scratch is needed).
RustError ntt_run(const NttTables& t, uint32_t log_n, const NttJob* jobs, int count, uint64_t in_len,
    Fr* scratch, uint64_t scratch_slots, NttTwist twist, NttPost post, const Fr& scale, hipStream_t stream);

bool ntt_needs_scratch(uint32_t log_n, bool aliased);

// Replaces the tuned tile table: tiles[log_n] = log2 of the elements per
// work-group (0 = default rule, else 6..12) for count <= NTT_MAX_LOG + 1 sizes.
RustError ntt_set_tiles(const uint8_t* tiles, uint32_t count);

// One transform of an uneven batch: its own size and tables (shared between
// jobs of one size), `in_len` valid inputs (the rest read as zero), in != out
// when the transform takes two passes, and a scratch buffer of 2^log_n
// elements when it takes three (log_n > 24).
struct NttBatchJob {
    const NttTables* tables;
    uint32_t log_n;
    uint64_t in_len;
    const Fr* in;
    Fr* out;
    Fr* scratch;
    Fr scale; // POST_SCALE / POST_SCALE_ZETA_INV factor
};

// Device bytes of the per-pass job tables of a batch of `count` transforms.
uint64_t ntt_batch_pass_bytes(int count);

// Runs every job's pass p in one launch (p = 0, 1, 2): work-groups of all
// transforms side by side, so small transforms run concurrently with large
// ones. `d_pass` holds ntt_batch_pass_bytes(count) device bytes. Synchronises
// once, after uploading the job tables.
RustError ntt_run_batch(const NttBatchJob* jobs, int count, NttTwist twist, NttPost post, void* d_pass,
    hipStream_t stream);

} // namespace h2
