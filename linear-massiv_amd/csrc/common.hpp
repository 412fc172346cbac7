// Internal runtime pieces shared by the kernels and the C ABI: error
// handling, the per-device context, stream-ordered scratch memory and the
// optional per-kernel-family event timing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <utility>
#include <vector>
#include <mutex>
#include <stdexcept>
#include <string>

#include "lmx.h"

namespace lmx {

// An error raised inside the library; converted to a status code at the C
// ABI boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char *what, const char *file, int line);
[[noreturn]] inline void fail_arg(const std::string &m) { throw Error(LMX_ERR_ARG, m); }

#define LMX_HIP(call)                                                        \
    do {                                                                     \
        hipError_t lmx_e_ = (call);                                          \
        if (lmx_e_ != hipSuccess) ::lmx::throw_hip(lmx_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define LMX_LAUNCH_CHECK() LMX_HIP(hipGetLastError())

void set_last_error(const std::string &m);

// Runs f() and maps exceptions to status codes; f returns the status.
template <class F>
int guarded(F &&f) {
    try {
        return f();
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return LMX_ERR_NOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return LMX_ERR_INTERNAL;
    }
}

// Per-device state.
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;  // used by the host-pointer calls
    std::mutex host_mu;            // serialises host-pointer calls (they share the staging buffers)
    void *staging[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t staging_bytes[4] = {0, 0, 0, 0};
    int cu_count = 0;
    // Scratch's memory pool: it keeps freed memory for reuse (release
    // threshold at its maximum), so an allocation after a host
    // synchronisation does not map memory again
    hipMemPool_t pool = nullptr;
    // Freed scratch blocks by the stream they were last used on, for reuse
    // on that stream with no HIP call (stream order makes the reuse safe),
    // at most kScratchCacheBytes in all; and idle lookahead streams.
    struct Block {
        void *p;
        size_t bytes;
    };
    std::mutex scratch_mu;
    std::vector<std::pair<hipStream_t, std::vector<Block>>> scratch_free;
    size_t scratch_cached = 0;
    struct Aux {
        hipStream_t aux;
        hipEvent_t ready, done;
    };
    std::vector<Aux> aux_free;
};

// Context of the current device (initialised on first use).
Context &ctx();
// Context of device `dev`, or nullptr when it is not initialised.
Context *context_of(int dev);
hipStream_t resolve_stream(void *s);
// Grow-only device staging buffer `slot` of at least `bytes` (host calls only,
// under host_mu).
void *staging(Context &c, int slot, size_t bytes);

// Stream-ordered scratch on `s`: a cached block last used on `s` when one
// fits, else a new one from the current device's pool; returned to the cache
// when the object goes out of scope.
struct Scratch {
    void *p = nullptr;
    hipStream_t s = nullptr;
    size_t bytes = 0;
    int device = -1;
    Scratch(size_t bytes, hipStream_t st);
    ~Scratch();
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

// A second stream and two events for a factorisation's lookahead: the
// caller's stream records `ready` when a block's update may start, the
// auxiliary stream waits for it, runs the update and records `done`.
// Taken from the device's idle set (created on first need) and returned to
// it when the factorisation returns; the next user's work follows in the
// stream's order.
struct AuxStream {
    hipStream_t aux = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    int device = -1;
    AuxStream();
    ~AuxStream();
    AuxStream(const AuxStream &) = delete;
    AuxStream &operator=(const AuxStream &) = delete;
};

// Optional per-family kernel timing (lmx_kernel_timing).
bool timing_enabled();
// `work` is the algorithmic work of the region (flops, or bytes for
// memory-bound families), summed per family.
struct TimedRegion {
    int slot = -1;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double work = 0;
    TimedRegion(const char *family, hipStream_t st, double work = 0);
    ~TimedRegion();
};

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline bool is_trans(char t) {
    if (t == 'N' || t == 'n') return false;
    if (t == 'T' || t == 't' || t == 'C' || t == 'c') return true;
    fail_arg(std::string("trans flag must be N or T, got '") + t + "'");
}

}  // namespace lmx
