// Optional per-kernel timing with HIP events recorded on the launch stream.
// Off by default; when enabled, launchers bracket their hot kernels with an
// event pair and the host reads back count and total milliseconds per name.
#pragma once

#include <hip/hip_runtime.h>

namespace h2 {

bool prof_enabled();

struct ProfScope {
    hipEvent_t a = nullptr, b = nullptr;
    const char* name = nullptr;
    hipStream_t stream = nullptr;
    ProfScope(const char* n, hipStream_t s);
    ~ProfScope();
};

} // namespace h2
