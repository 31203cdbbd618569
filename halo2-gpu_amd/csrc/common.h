// Boundary plumbing shared by every launcher: the by-value error struct, the
// pointer carrier, the caller-owned workspace carver and HIP error checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "halo2_gpu.h"

namespace h2 {

static inline RustError ok() { return RustError { 0, nullptr }; }

static inline RustError fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static inline RustError fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    return RustError { code, strdup(buf) };
}

#define H2_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::h2::fail((int)e_, "%s@%s:%d failed: %s", #expr, __FILE__, __LINE__,      \
                hipGetErrorString(e_));                                                       \
    } while (0)

#define H2_LAUNCH_CHECK()                                                                     \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess)                                                                 \
            return ::h2::fail((int)e_, "kernel launch @%s:%d failed: %s", __FILE__, __LINE__, \
                hipGetErrorString(e_));                                                       \
    } while (0)

#define H2_CHECK_ARG(cond, ...)                                                               \
    do {                                                                                      \
        if (!(cond))                                                                          \
            return ::h2::fail((int)hipErrorInvalidValue, __VA_ARGS__);                        \
    } while (0)

static inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + (a - 1)) & ~(a - 1); }

// Caller-owned scratch carved front to back. `take` returns nullptr once the
// span is exhausted; launchers turn that into an error before any launch, so a
// short workspace never reaches a kernel.
struct Workspace {
    uint8_t* ptr;
    uint64_t left;
    bool overflow = false;

    Workspace(void* p, uint64_t bytes)
        : ptr((uint8_t*)p)
        , left(bytes)
    {
    }

    void* take(uint64_t n, uint64_t align = 256)
    {
        uintptr_t cur = (uintptr_t)ptr;
        uintptr_t al = (cur + (align - 1)) & ~(uintptr_t)(align - 1);
        uint64_t pad = al - cur;
        if (overflow || pad > left || n > left - pad) {
            overflow = true;
            return nullptr;
        }
        ptr = (uint8_t*)(al + n);
        left -= pad + n;
        return (void*)al;
    }
};

// Size accumulator mirroring Workspace::take (the caller's buffer is assumed
// 256-byte aligned, as hipMalloc guarantees).
struct WorkspaceSizer {
    uint64_t bytes = 0;
    void take(uint64_t n, uint64_t align = 256) { bytes = align_up(bytes, align) + n; }
};

// Copies a (na bytes) to d_a and b (nb bytes) to d_b on `stream` through a ring of
// pinned host buffers, so the caller's host arrays may go away on return without a
// stream synchronisation (staging.hip).
RustError stage_upload(void* d_a, const void* a, size_t na, void* d_b, const void* b, size_t nb,
    hipStream_t stream);

static inline unsigned grid_for(uint64_t n, unsigned block, unsigned max_blocks = 65535u * 16u)
{
    uint64_t g = (n + block - 1) / block;
    if (g == 0)
        g = 1;
    if (g > max_blocks)
        g = max_blocks;
    return (unsigned)g;
}

} // namespace h2
