// Per-kernel HIP-event timing registry (see profile.h).
#include "profile.h"

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace h2 {

namespace {

struct Pending {
    std::string name;
    hipEvent_t a, b;
};

struct Registry {
    std::mutex mu;
    bool on = false;
    std::vector<Pending> pending;
    std::map<std::string, std::pair<uint64_t, double>> totals;
};

Registry& reg()
{
    static Registry r;
    return r;
}

void drain_locked(Registry& r)
{
    for (auto& p : r.pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& t = r.totals[p.name];
            t.first += 1;
            t.second += ms;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    r.pending.clear();
}

} // namespace

bool prof_enabled() { return reg().on; }

ProfScope::ProfScope(const char* n, hipStream_t s)
{
    if (!prof_enabled())
        return;
    name = n;
    stream = s;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        a = b = nullptr;
        return;
    }
    (void)hipEventRecord(a, s);
}

ProfScope::~ProfScope()
{
    if (!a)
        return;
    (void)hipEventRecord(b, stream);
    Registry& r = reg();
    std::lock_guard<std::mutex> g(r.mu);
    r.pending.push_back(Pending { name, a, b });
}

// Throughput probe of the instruction the field arithmetic is built on:
// 8 independent v_mad_u64_u32 chains per lane, `iters` x 64 mads per lane.
__global__ void mad_probe_kernel(uint64_t* out, uint32_t iters)
{
    uint32_t x = threadIdx.x + 1, y = blockIdx.x + 3;
    uint64_t a0 = 1, a1 = 2, a2 = 3, a3 = 4, a4 = 5, a5 = 6, a6 = 7, a7 = 8;
    for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            uint64_t c;
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a0), "=s"(c) : "v"(x), "v"(y));
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a1), "=s"(c) : "v"(x), "v"(y));
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a2), "=s"(c) : "v"(x), "v"(y));
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a3), "=s"(c) : "v"(x), "v"(y));
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a4), "=s"(c) : "v"(x), "v"(y));
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a5), "=s"(c) : "v"(x), "v"(y));
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a6), "=s"(c) : "v"(x), "v"(y));
            asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(a7), "=s"(c) : "v"(x), "v"(y));
        }
    }
    out[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

} // namespace h2

extern "C" {

// Measures the device's v_mad_u64_u32 rate (multiply-adds per second) with
// `blocks` x 256 threads; `scratch` must hold blocks * 256 * 8 bytes.
RustError h2amd_mad_peak_probe(void* scratch, uint32_t blocks, uint32_t iters, double* mads_per_s, void* stream_)
{
    hipStream_t stream = (hipStream_t)stream_;
    hipEvent_t a, b;
    H2_TRY(hipEventCreate(&a));
    H2_TRY(hipEventCreate(&b));
    h2::mad_probe_kernel<<<blocks, 256, 0, stream>>>((uint64_t*)scratch, iters); // warm-up
    H2_TRY(hipEventRecord(a, stream));
    h2::mad_probe_kernel<<<blocks, 256, 0, stream>>>((uint64_t*)scratch, iters);
    H2_TRY(hipEventRecord(b, stream));
    H2_TRY(hipEventSynchronize(b));
    float ms = 0.f;
    H2_TRY(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *mads_per_s = (double)blocks * 256.0 * iters * 64.0 / (ms * 1e-3);
    return h2::ok();
}

void h2amd_profile_enable(int on)
{
    auto& r = h2::reg();
    std::lock_guard<std::mutex> g(r.mu);
    r.on = on != 0;
}

void h2amd_profile_reset(void)
{
    auto& r = h2::reg();
    std::lock_guard<std::mutex> g(r.mu);
    h2::drain_locked(r);
    r.totals.clear();
}

int h2amd_profile_read(const char* name, uint64_t* count, double* total_ms)
{
    auto& r = h2::reg();
    std::lock_guard<std::mutex> g(r.mu);
    h2::drain_locked(r);
    auto it = r.totals.find(name);
    if (it == r.totals.end()) {
        *count = 0;
        *total_ms = 0.0;
        return 0;
    }
    *count = it->second.first;
    *total_ms = it->second.second;
    return 1;
}

} // extern "C"
