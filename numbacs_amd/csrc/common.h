// Shared host/device helpers of libnumbacs_amd: error reporting, launch
// checks and the grow-only device workspaces.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
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

// Grow-only device scratch owned by the library: one buffer per (device,
// stream, slot), held until the process ends.  A Workspace keeps its buffer
// locked from construction to destruction, i.e. for the whole entry-point
// call that holds it: calls on different streams use different buffers and
// run concurrently, calls on one stream from several host threads take turns
// (and their device work is ordered by the stream).  Growing frees the old
// buffer with hipFree, which waits for the device work that used it.
class Workspace {
   public:
    Workspace(hipStream_t stream, int slot, size_t bytes);
    Workspace(const Workspace&) = delete;
    Workspace& operator=(const Workspace&) = delete;
    char* get() const { return ptr_; }
    int error() const { return err_; }

   private:
    std::unique_lock<std::mutex> lock_;
    char* ptr_ = nullptr;
    int err_ = 0;
};

}  // namespace nbx
