// Shared host/device plumbing for the pynear_amd HIP library: status codes, the thread-local
// error string behind pn_last_error(), HIP error checks, a small device-buffer type and the
// library stream.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../include/pynear_amd.h"

namespace pn {

// Error carrying a pn_status; the C-ABI entry points catch it and return the status.
struct Error : std::runtime_error {
    int32_t status;
    Error(int32_t s, const std::string &m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string &msg);

[[noreturn]] inline void fail(int32_t status, const std::string &msg) { throw Error(status, msg); }

#define PN_HIP(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            ::pn::fail(e_ == hipErrorOutOfMemory ? PN_ENOMEM : PN_EDEVICE,                            \
                       std::string(#expr) + ": " + hipGetErrorString(e_));                            \
    } while (0)

#define PN_REQUIRE(cond, status, msg)                                                                 \
    do {                                                                                              \
        if (!(cond)) ::pn::fail((status), (msg));                                                     \
    } while (0)

// Runs f(), converting exceptions into a status and the thread-local error string.
template <typename F> int32_t guard(F &&f) {
    try {
        f();
        return PN_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return PN_ENOMEM;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return PN_ERUNTIME;
    }
}

// Owned device allocation that only grows (used for per-index workspaces).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void *ensure(size_t need) {
        if (need <= bytes && p) return p;
        release();
        size_t n = need ? need : 256;
        PN_HIP(hipMalloc(&p, n));
        bytes = n;
        return p;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

// Library stream for the host-pointer entry points (default priority, non-blocking).
hipStream_t lib_stream();

// Device selected for the calling thread while the guard lives.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        PN_HIP(hipGetDevice(&prev));
        if (prev != dev) PN_HIP(hipSetDevice(dev));
    }
    ~DeviceScope() {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
    }
};

inline int current_device() {
    int d = 0;
    PN_HIP(hipGetDevice(&d));
    return d;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Launch-shape tuning table: named integer parameters with defaults, overridable through the
// PN_TUNE environment variable ("name=value,name=value") and pn_tune_set().
int64_t tune_get(const char *name);

// Optional per-stage timing with HIP events recorded on the launch stream (off by default).
enum Stage { STAGE_SCAN = 0, STAGE_SELECT = 1, STAGE_COARSE = 2, STAGE_RERANK = 3, STAGE_OTHER = 4, STAGE_COUNT = 5 };
void stage_mark(hipStream_t st, int stage, bool begin);
struct StageScope {
    hipStream_t st;
    int stage;
    StageScope(hipStream_t s, int g) : st(s), stage(g) { stage_mark(st, stage, true); }
    ~StageScope() { stage_mark(st, stage, false); }
};

} // namespace pn
