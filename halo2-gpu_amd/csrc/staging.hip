// Pinned host staging of small per-call uploads (see staging.hip).
#include <mutex>

#include "common.h"

namespace h2 {

// Host staging of small per-call tables (rule programs, value tables, NTT pass
// tables): a ring of pinned buffers, each
// rewritten only after the copy that last read it has run (its event), so a call
// returns without draining the stream (a pageable copy would need the host data to
// outlive it, i.e. a stream synchronisation per call). Serialised by a mutex;
// falls back to a synchronised pageable copy when staging cannot be set up.
struct StageSlot {
    uint8_t* buf = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    int dev = -1;
    bool pending = false;
};
constexpr unsigned kStageSlots = 8;
std::mutex g_stage_mu;
StageSlot g_stage[kStageSlots];
unsigned g_stage_next = 0;

RustError stage_upload(void* d_a, const void* a, size_t na, void* d_b, const void* b, size_t nb, hipStream_t stream)
{
    if (na + nb == 0)
        return ok();
    int dev = 0;
    H2_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_stage_mu);
    StageSlot& sl = g_stage[g_stage_next++ % kStageSlots];
    if (sl.pending) {
        H2_TRY(hipEventSynchronize(sl.ev)); // the copy that last read this buffer has run
        sl.pending = false;
    }
    if (sl.dev != dev || sl.cap < na + nb) {
        if (sl.buf)
            H2_TRY(hipHostFree(sl.buf));
        if (sl.ev && sl.dev != dev) {
            H2_TRY(hipEventDestroy(sl.ev));
            sl.ev = nullptr;
        }
        sl.buf = nullptr;
        sl.cap = 0;
        const size_t cap = (na + nb + 65535) & ~(size_t)65535;
        if (hipHostMalloc((void**)&sl.buf, cap, hipHostMallocDefault) != hipSuccess) {
            sl.buf = nullptr;
            if (na)
                H2_TRY(hipMemcpyAsync(d_a, a, na, hipMemcpyHostToDevice, stream));
            if (nb)
                H2_TRY(hipMemcpyAsync(d_b, b, nb, hipMemcpyHostToDevice, stream));
            H2_TRY(hipStreamSynchronize(stream)); // pageable sources must outlive the copies
            return ok();
        }
        sl.cap = cap;
        if (!sl.ev)
            H2_TRY(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
        sl.dev = dev;
    }
    if (na)
        memcpy(sl.buf, a, na);
    if (nb)
        memcpy(sl.buf + na, b, nb);
    if (na)
        H2_TRY(hipMemcpyAsync(d_a, sl.buf, na, hipMemcpyHostToDevice, stream));
    if (nb)
        H2_TRY(hipMemcpyAsync(d_b, sl.buf + na, nb, hipMemcpyHostToDevice, stream));
    H2_TRY(hipEventRecord(sl.ev, stream));
    sl.pending = true;
    return ok();
}

} // namespace h2
