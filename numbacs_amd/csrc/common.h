// Shared host/device helpers of libnumbacs_amd: error reporting, launch
// checks and a per-device grow-only workspace.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/numbacs_amd.h"

namespace nbx {

void set_error(const std::string& msg);

// Returns -1 with a message naming the failing HIP call.
#define NBX_HIP(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            ::nbx::set_error(std::string(#call) + ": " + hipGetErrorString(e_));        \
            return -1;                                                                 \
        }                                                                              \
    } while (0)

#define NBX_REQUIRE(cond, msg)                                                         \
    do {                                                                               \
        if (!(cond)) {                                                                 \
            ::nbx::set_error(msg);                                                     \
            return -2;                                                                 \
        }                                                                              \
    } while (0)

// Check the launch that was just enqueued.
#define NBX_LAUNCHED(name)                                                             \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess) {                                                        \
            ::nbx::set_error(std::string("launch of ") + name + ": " + hipGetErrorString(e_)); \
            return -1;                                                                 \
        }                                                                              \
    } while (0)

inline hipStream_t as_stream(nbx_stream s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned blocks_for(int64_t n, int threads) {
    return static_cast<unsigned>((n + threads - 1) / threads);
}

// Grow-only device scratch owned by the library, one per device and slot.
// Calls that use the same slot on one device must not overlap in time.
void* workspace(int slot, size_t bytes, int* err);

}  // namespace nbx
