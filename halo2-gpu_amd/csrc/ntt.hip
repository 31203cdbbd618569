// NTT entry points of the C ABI (`_halo2_fft_*`, `_halo2_cosetfft`,
// `_halo2_get_fft_split_radix`), mirroring cuda/src/ntt.cu of the reference.
//
// Semantics per NTT kind (ntt.cu:202-303):
//   FFT           out = DFT_omega(in)
//   iFFT          out = divisor * DFT_omega(in)            (callers pass omega^-1, 2^-k)
//   cosetFFT      out = DFT_omega(pad(in)_i * zeta^(i mod 3)) of size 2^extend_log_n
//   icosetFFT     out_i = divisor * DFT_omega(pad(in))_i * zeta^-(i mod 3)
//   CosetFFT_Part out = DFT_omega(in_i * c^i), c = divisor (g_coset * ext_omega^p)
#include <vector>

#include "ntt_core.h"

using namespace h2;

namespace {

bool kind_ok(uint32_t t) { return t == 1 || t == 2 || t == 3 || t == 4 || t == 6; }

// active transform size: coset kinds run at the extended size
uint32_t active_log(uint32_t t, uint32_t log_n, uint32_t ext) { return (t == 3 || t == 4) ? ext : log_n; }

void kind_params(uint32_t t, NttTwist& tw, NttPost& post)
{
    tw = TWIST_NONE;
    post = POST_NONE;
    if (t == 2)
        post = POST_SCALE;
    else if (t == 3)
        tw = TWIST_ZETA;
    else if (t == 4)
        post = POST_SCALE_ZETA_INV;
    else if (t == 6)
        tw = TWIST_POWERS;
}

Fr host_fr(const void* p)
{
    Fr r;
    memcpy(r.v, p, 32);
    return r;
}

// in-slot, out-slot, scratch-slot, omega slot, tables
struct HostLayout {
    uint64_t in_elems, out_elems;
    bool scratch;
};

// Upper bound of what a launcher carves: staging slots, one scratch buffer,
// the omega slot and the tables, plus slack for per-take alignment.
uint64_t sized(uint32_t t, uint32_t log_act, uint64_t in_elems, uint64_t out_elems, bool scratch)
{
    WorkspaceSizer s;
    if (in_elems)
        s.take(32 * in_elems);
    if (out_elems)
        s.take(32 * out_elems);
    if (scratch)
        s.take(32ull << log_act);
    s.take(32);
    ntt_tables_size(s, log_act, t == 6);
    return s.bytes + 4096;
}

RustError check_log(const char* who, uint32_t log_act)
{
    if (log_act < 1 || log_act > NTT_MAX_LOG)
        return fail((int)hipErrorInvalidValue, "%s: transform size 2^%u outside [2^1, 2^%u]", who, log_act, NTT_MAX_LOG);
    return ok();
}

} // namespace

extern "C" {

uint64_t _halo2_fft_normal_workspace_size(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n)
{
    uint32_t la = active_log(ntt_type, log_n, extend_log_n);
    if (la < 1 || la > NTT_MAX_LOG || log_n > la)
        return 0;
    // covers both users of this size: _halo2_fft_normal (data + scratch at
    // 2^log_n) and _halo2_cosetfft (input 2^log_n, output + scratch at 2^ext)
    return sized(ntt_type, la, 1ull << log_n, 1ull << la, true);
}

bool _halo2_fft_normal_check_memory(uint32_t, const void*, uint32_t, uint32_t) { return true; }

RustError _halo2_fft_normal(uint32_t ntt_type, uint32_t log_n, const void* input, void* output, const void* omega,
    const void* divisor, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (ntt_type == 5)
        return fail(-1, "iFFT_cosetFFT not supported");
    if (!kind_ok(ntt_type))
        return fail(-6, "NTT type %u not supported", ntt_type);
    // the reference runs every kind of this entry point at size 2^log_n (ntt.cu:411-415)
    RustError e = check_log("_halo2_fft_normal", log_n);
    if (e.code)
        return e;
    const uint64_t N = 1ull << log_n;
    Workspace ws(scratch, scratch_bytes);
    Fr* data = (Fr*)ws.take(32 * N);
    Fr* scr = (Fr*)ws.take(32 * N);
    Fr* d_omega = (Fr*)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_fft_normal: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    H2_TRY(hipMemcpyAsync(data, input, 32 * N, hipMemcpyHostToDevice, stream));
    H2_TRY(hipMemcpyAsync(d_omega, omega, 32, hipMemcpyHostToDevice, stream));
    NttTwist tw;
    NttPost post;
    kind_params(ntt_type, tw, post);
    Fr div = host_fr(divisor);
    NttTables tabs;
    e = ntt_tables_build(ws, tabs, log_n, d_omega, tw == TWIST_POWERS, div, stream);
    if (e.code)
        return e;
    NttJob job { data, data };
    e = ntt_run(tabs, log_n, &job, 1, N, scr, 1, tw, post, div, stream);
    if (e.code)
        return e;
    H2_TRY(hipMemcpyAsync(output, data, 32 * N, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

uint64_t _halo2_fft_normal_to_device_workspace_size(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n)
{
    (void)extend_log_n;
    if (log_n < 1 || log_n > NTT_MAX_LOG)
        return 0;
    return sized(ntt_type, log_n, 0, 0, true);
}

RustError _halo2_fft_normal_to_device(uint32_t ntt_type, uint32_t log_n, const void* d_input, void* d_output,
    const void* d_omega, const void* h_divisor, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    if (ntt_type != 1 && ntt_type != 2 && ntt_type != 6)
        return fail((int)hipErrorInvalidValue,
            "_halo2_fft_normal_to_device supports only FFT, iFFT, CosetFFT_Part (got %u)", ntt_type);
    RustError e = check_log("_halo2_fft_normal_to_device", log_n);
    if (e.code)
        return e;
    const uint64_t N = 1ull << log_n;
    Workspace ws(scratch, scratch_bytes);
    Fr* scr = (Fr*)ws.take(32 * N);
    (void)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_fft_normal_to_device: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    NttTwist tw;
    NttPost post;
    kind_params(ntt_type, tw, post);
    Fr div = host_fr(h_divisor);
    NttTables tabs;
    e = ntt_tables_build(ws, tabs, log_n, (const Fr*)d_omega, tw == TWIST_POWERS, div, stream);
    if (e.code)
        return e;
    NttJob job { (const Fr*)d_input, (Fr*)d_output };
    return ntt_run(tabs, log_n, &job, 1, N, scr, 1, tw, post, div, stream);
}

RustError _halo2_cosetfft(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n, const void* input, void* output,
    const void* omega, const void* divisor, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    (void)ntt_type; // the reference body always runs cosetFFT here (ntt.cu:623-629)
    (void)divisor;
    hipStream_t stream = (hipStream_t)stream_;
    if (log_n == 0 || extend_log_n == 0 || log_n > extend_log_n)
        return fail((int)hipErrorInvalidValue, "_halo2_cosetfft: need 1 <= log_n <= extend_log_n (got %u, %u)", log_n,
            extend_log_n);
    RustError e = check_log("_halo2_cosetfft", extend_log_n);
    if (e.code)
        return e;
    const uint64_t n = 1ull << log_n, N = 1ull << extend_log_n;
    Workspace ws(scratch, scratch_bytes);
    Fr* din = (Fr*)ws.take(32 * n);
    Fr* dout = (Fr*)ws.take(32 * N);
    Fr* scr = (Fr*)ws.take(32 * N);
    Fr* d_omega = (Fr*)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "_halo2_cosetfft: workspace of %lu bytes is too small",
            (unsigned long)scratch_bytes);
    H2_TRY(hipMemcpyAsync(din, input, 32 * n, hipMemcpyHostToDevice, stream));
    H2_TRY(hipMemcpyAsync(d_omega, omega, 32, hipMemcpyHostToDevice, stream));
    NttTables tabs;
    e = ntt_tables_build(ws, tabs, extend_log_n, d_omega, false, Fr::zero(), stream);
    if (e.code)
        return e;
    NttJob job { din, dout };
    e = ntt_run(tabs, extend_log_n, &job, 1, n, scr, 1, TWIST_ZETA, POST_NONE, Fr::zero(), stream);
    if (e.code)
        return e;
    H2_TRY(hipMemcpyAsync(output, dout, 32 * N, hipMemcpyDeviceToHost, stream));
    H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

uint64_t _halo2_fft_many_workspace_size(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n)
{
    uint32_t la = active_log(ntt_type, log_n, extend_log_n);
    if (la < 1 || la > NTT_MAX_LOG || log_n > la)
        return 0;
    return sized(ntt_type, la, 1ull << log_n, 1ull << la, true);
}

uint64_t _halo2_fft_many_to_device_workspace_size(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n)
{
    uint32_t la = active_log(ntt_type, log_n, extend_log_n);
    if (la < 1 || la > NTT_MAX_LOG || log_n > la)
        return 0;
    return sized(ntt_type, la, 1ull << log_n, 0, true);
}

bool _halo2_fft_many_check_memory(uint32_t, uint32_t, uint32_t) { return true; }

static RustError fft_many_impl(const char* who, uint32_t ntt_type, uint32_t num_many, uint32_t log_n,
    uint32_t extend_log_n, const FFITraitObject* input, const FFITraitObject* output, const void* d_omega,
    const FFITraitObject* divisor, bool input_on_device, bool output_on_device, void* scratch,
    uint64_t scratch_bytes, hipStream_t stream)
{
    if (ntt_type == 4)
        return fail(2, "icosetFFT not supported now");
    if (ntt_type == 5)
        return fail(-1, "iFFT_cosetFFT not supported");
    if (!kind_ok(ntt_type))
        return fail(-6, "NTT type %u not supported", ntt_type);
    if (log_n == 0 || extend_log_n == 0)
        return fail((int)hipErrorInvalidValue, "%s: log_n and extend_log_n must be >= 1", who);
    const uint32_t la = active_log(ntt_type, log_n, extend_log_n);
    RustError e = check_log(who, la);
    if (e.code)
        return e;
    if (log_n > la)
        return fail((int)hipErrorInvalidValue, "%s: log_n %u exceeds the transform size 2^%u", who, log_n, la);
    if (num_many == 0)
        return ok();
    const uint64_t n_in = 1ull << log_n, N = 1ull << la;
    Workspace ws(scratch, scratch_bytes);
    Fr* din = (Fr*)ws.take(32 * n_in);
    Fr* dout = output_on_device ? nullptr : (Fr*)ws.take(32 * N);
    Fr* scr = (Fr*)ws.take(32 * N);
    (void)ws.take(32);
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "%s: workspace of %lu bytes is too small", who,
            (unsigned long)scratch_bytes);
    NttTwist tw;
    NttPost post;
    kind_params(ntt_type, tw, post);
    Fr div = host_fr(divisor->ptr);
    NttTables tabs;
    e = ntt_tables_build(ws, tabs, la, (const Fr*)d_omega, tw == TWIST_POWERS, div, stream);
    if (e.code)
        return e;
    if (input_on_device && output_on_device) {
        // all polynomials device-resident: batch them through the pass kernels
        NttJob jobs[NTT_MAX_BATCH];
        for (uint32_t base = 0; base < num_many; base += NTT_MAX_BATCH) {
            int cnt = 0;
            for (uint32_t i = base; i < num_many && cnt < NTT_MAX_BATCH; ++i, ++cnt)
                jobs[cnt] = NttJob { (const Fr*)input[i].ptr, (Fr*)output[i].ptr };
            e = ntt_run(tabs, la, jobs, cnt, n_in, scr, 1, tw, post, div, stream);
            if (e.code)
                return e;
        }
        return ok();
    }
    for (uint32_t i = 0; i < num_many; ++i) {
        const Fr* src = (const Fr*)input[i].ptr;
        if (!input_on_device) {
            H2_TRY(hipMemcpyAsync(din, input[i].ptr, 32 * n_in, hipMemcpyHostToDevice, stream));
            src = din;
        }
        Fr* dst = output_on_device ? (Fr*)output[i].ptr : dout;
        NttJob job { src, dst };
        e = ntt_run(tabs, la, &job, 1, n_in, scr, 1, tw, post, div, stream);
        if (e.code)
            return e;
        if (!output_on_device)
            H2_TRY(hipMemcpyAsync(output[i].ptr, dout, 32 * N, hipMemcpyDeviceToHost, stream));
    }
    if (!output_on_device)
        H2_TRY(hipStreamSynchronize(stream));
    return ok();
}

RustError _halo2_fft_many(uint32_t ntt_type, uint32_t num_many, uint32_t log_n, uint32_t extend_log_n,
    const FFITraitObject* input, const FFITraitObject* output, const void* d_omega, const FFITraitObject* divisor,
    void* scratch, uint64_t scratch_bytes, void* stream)
{
    return fft_many_impl("_halo2_fft_many", ntt_type, num_many, log_n, extend_log_n, input, output, d_omega, divisor,
        false, false, scratch, scratch_bytes, (hipStream_t)stream);
}

RustError _halo2_fft_many_to_device(uint32_t ntt_type, uint32_t num_many, uint32_t log_n, uint32_t extend_log_n,
    const FFITraitObject* input, const FFITraitObject* output, const void* d_omega, const FFITraitObject* divisor,
    bool input_on_device, void* scratch, uint64_t scratch_bytes, void* stream)
{
    return fft_many_impl("_halo2_fft_many_to_device", ntt_type, num_many, log_n, extend_log_n, input, output, d_omega,
        divisor, input_on_device, true, scratch, scratch_bytes, (hipStream_t)stream);
}

int32_t _halo2_get_fft_split_radix(uint32_t ntt_type, uint32_t log_n, uint32_t extend_log_n, uint64_t free_bytes)
{
    for (uint32_t s = 0; s < log_n; ++s) {
        uint32_t la = active_log(ntt_type, log_n - s, extend_log_n >= s ? extend_log_n - s : 0);
        if (la < 1)
            break;
        if (_halo2_fft_normal_workspace_size(ntt_type, log_n - s, extend_log_n - s) < free_bytes)
            return (int32_t)s;
    }
    return -1;
}

// ----------------------------------------- uneven batches (no reference counterpart)

uint64_t h2amd_fft_batch_to_device_workspace_size(uint32_t ntt_type, uint32_t count, const uint32_t* log_n,
    const uint32_t* extend_log_n)
{
    WorkspaceSizer s;
    for (uint32_t i = 0; i < count; ++i) {
        const uint32_t la = active_log(ntt_type, log_n[i], extend_log_n[i]);
        if (la < 1 || la > NTT_MAX_LOG || log_n[i] > la)
            return 0;
        ntt_tables_size(s, la, ntt_type == 6);
        if (ntt_plan(la).P == 3)
            s.take(32ull << la);
    }
    s.take(ntt_batch_pass_bytes((int)count));
    return s.bytes + 4096;
}

RustError h2amd_fft_batch_to_device(uint32_t ntt_type, uint32_t count, const uint32_t* log_n,
    const uint32_t* extend_log_n, const void* const* d_in, void* const* d_out, const void* const* d_omega,
    const FFITraitObject* divisors, void* scratch, uint64_t scratch_bytes, void* stream_)
{
    const char* who = "h2amd_fft_batch_to_device";
    hipStream_t stream = (hipStream_t)stream_;
    if (ntt_type == 4)
        return fail(2, "icosetFFT not supported now");
    if (ntt_type == 5)
        return fail(-1, "iFFT_cosetFFT not supported");
    if (!kind_ok(ntt_type))
        return fail(-6, "NTT type %u not supported", ntt_type);
    if (count == 0)
        return ok();
    NttTwist tw;
    NttPost post;
    kind_params(ntt_type, tw, post);
    Workspace ws(scratch, scratch_bytes);
    std::vector<NttTables> tabs(count);
    std::vector<uint32_t> owner(count), las(count);
    std::vector<Fr> divs(count);
    std::vector<NttBatchJob> jobs(count);
    for (uint32_t i = 0; i < count; ++i) {
        if (log_n[i] == 0 || extend_log_n[i] == 0)
            return fail((int)hipErrorInvalidValue, "%s: log_n and extend_log_n must be >= 1 (job %u)", who, i);
        const uint32_t la = active_log(ntt_type, log_n[i], extend_log_n[i]);
        RustError e = check_log(who, la);
        if (e.code)
            return e;
        if (log_n[i] > la)
            return fail((int)hipErrorInvalidValue, "%s: job %u: log_n %u exceeds the transform size 2^%u", who, i,
                log_n[i], la);
        las[i] = la;
        divs[i] = host_fr(divisors[i].ptr);
        // jobs of one size, omega (and coset factor) share their tables
        owner[i] = i;
        for (uint32_t k = 0; k < i; ++k) {
            const bool same_c = tw != TWIST_POWERS || !memcmp(divs[k].v, divs[i].v, 32);
            if (las[k] == la && d_omega[k] == d_omega[i] && same_c) {
                owner[i] = owner[k];
                break;
            }
        }
        if (owner[i] == i) {
            e = ntt_tables_build(ws, tabs[i], la, (const Fr*)d_omega[i], tw == TWIST_POWERS, divs[i], stream);
            if (e.code)
                return e;
        }
        Fr* scr = ntt_plan(la).P == 3 ? (Fr*)ws.take(32ull << la) : nullptr;
        jobs[i] = NttBatchJob { &tabs[owner[i]], la, 1ull << log_n[i], (const Fr*)d_in[i], (Fr*)d_out[i], scr,
            divs[i] };
    }
    void* d_pass = ws.take(ntt_batch_pass_bytes((int)count));
    if (ws.overflow)
        return fail((int)hipErrorInvalidValue, "%s: workspace of %lu bytes is too small", who,
            (unsigned long)scratch_bytes);
    return ntt_run_batch(jobs.data(), (int)count, tw, post, d_pass, stream);
}

RustError h2amd_ntt_set_tiles(const uint8_t* tiles, uint32_t count) { return ntt_set_tiles(tiles, count); }

} // extern "C"
