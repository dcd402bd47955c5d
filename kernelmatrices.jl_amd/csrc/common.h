// Shared host-side infrastructure of libkmhodlr: errors, the context (device,
// stream, descriptor arena, stream-ordered scratch), device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "kmhodlr.h"

namespace kmh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define KMH_HIP(x)                                                                       \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            throw ::kmh::Error(e_ == hipErrorOutOfMemory ? KMH_E_OOM : KMH_E_DEVICE,     \
                               std::string(#x) + ": " + hipGetErrorString(e_));          \
    } while (0)

#define KMH_CHECK(cond, code, msg)                                                       \
    do {                                                                                 \
        if (!(cond)) throw ::kmh::Error((code), (msg));                                  \
    } while (0)

void set_last_error(const std::string& m);
const char* last_error_cstr();

// Device allocation owned by a handle (hipMalloc; counted against the budget).
struct Ctx;
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    Ctx* ctx = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
    DevBuf& operator=(DevBuf&& o) noexcept;
    ~DevBuf();
    void alloc(Ctx* c, size_t nbytes);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

struct Ctx {
    int device = 0;
    uint64_t budget = 0;       // 0 = unlimited
    uint64_t in_use = 0;       // bytes held by handles
    hipStream_t stream = nullptr;
    // Descriptor arena: kernels read their problem lists from here.  Uploads
    // are stream-ordered, so a region may be rewritten by the next launch.
    char* arena = nullptr;
    char* staging = nullptr;   // pinned host mirror of the arena
    size_t arena_size = 0;
    size_t arena_off = 0;
    // Device flag for Cholesky failures (non-positive pivots).
    int* d_info = nullptr;
    // GEMM profiling: events bracketing every grouped-GEMM launch on this
    // stream and the algorithmic flops (2mnk) of its problems.
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<double> ev_flops;
    hipEvent_t prof_event();

    void set_device() const { KMH_HIP(hipSetDevice(device)); }
    // Copy `bytes` of host data into the arena; returns the device address.
    void* upload(const void* host, size_t bytes);
    template <class T> T* upload_vec(const std::vector<T>& v) {
        return static_cast<T*>(upload(v.data(), v.size() * sizeof(T)));
    }
    void arena_reset() { arena_off = 0; }
    // Stream-ordered scratch (freed with scratch_free on the same stream).
    void* scratch(size_t bytes);
    void scratch_free(void* p);
    void sync() { KMH_HIP(hipStreamSynchronize(stream)); }
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace kmh

// Opaque C-ABI context handle.
struct kmh_ctx {
    kmh::Ctx c;
};
