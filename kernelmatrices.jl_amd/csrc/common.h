// Shared host-side infrastructure of libkmhodlr: errors, the context (device,
// stream, descriptor arena, launch queue, scratch pool), device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
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
    // land in a pinned host mirror at the same offsets; the range written
    // since the last flush goes to the device in one copy at the flush.
    char* arena = nullptr;
    char* staging = nullptr;
    size_t arena_size = 0;
    size_t arena_off = 0;
    size_t dirty_lo = 0, dirty_hi = 0;
    static constexpr int ARENA_SEGS = 4;
    struct Fence {
        hipEvent_t ev = nullptr;
        bool pending = false;
    };
    Fence seg_fence[ARENA_SEGS];
    int arena_seg = 0;   // segment arena_off lies in (arena_off may equal its end)
    // Launch queue: kernels are recorded and issued, in order and after the
    // descriptor copy, at the next flush.  Every direct operation on the
    // stream goes through q(), which flushes first, so stream order is the
    // program order.  After a synchronization the stream is idle, so the
    // queue is flushed early, at 1, 2, 4, ... pending launches (up to
    // FLUSH_MAX): the device starts on a stage while the host still plans
    // the rest of it.
    std::vector<std::function<void()>> pending;
    static constexpr size_t FLUSH_MAX = 64;
    size_t flush_at = 1;
    // Device flag for Cholesky failures (non-positive pivots).
    int* d_info = nullptr;
    // Profiling spans (measurement hooks): one kernel launch each, timed by
    // event markers recorded around it on the stream, with a tag,
    // the algorithmic amount (bytes) for that tag and the launch's flops
    // (GEMM launches; negative for the others, which stay out of the GEMM
    // totals).
    struct ProfSpan {
        int tag;
        double amount;
        double flops;
        hipEvent_t a, b;
        unsigned grid;   // workgroups of the launch
    };
    bool prof = false;
    uint32_t prof_mask = ~0u;   // tags whose launches are timed
    int prof_stride = 1;        // time every prof_stride-th launch of a tag
    uint64_t prof_seen[KMH_PROF_TAGS] = {};
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<ProfSpan> spans;
    double prof_sum[KMH_PROF_TAGS][3] = {};   // per tag: amount, ms, spans
    hipEvent_t prof_event();
    // Scratch pool: device chunks carved first-fit on the host.  One in-order
    // stream runs every kernel of the context, so a block freed here may be
    // handed to a later kernel at once.
    std::vector<std::pair<char*, size_t>> chunks;
    std::map<char*, std::pair<size_t, int>> free_blocks;   // start -> (bytes, chunk)
    struct LiveBlock {
        size_t bytes;
        int chunk;
        uint64_t seq;   // allocation number (scratch_sweep)
    };
    std::unordered_map<void*, LiveBlock> live;
    uint64_t alloc_seq = 0;
    size_t pool_bytes = 0;
    // Arenas replaced by a grow: a caller may still hold descriptor
    // addresses in them for a launch it has not queued yet, so they are
    // freed with the context, not at the grow.
    std::vector<char*> retired_arenas;
    // Test hook: scratch() fails (KMH_E_OOM, no device call) once this many
    // more blocks have been handed out; negative = never.
    int64_t fail_scratch_after = -1;
    // Fixed scratch region (the allocation-free entry points): while set,
    // scratch() hands out blocks of [fixed_base, fixed_base + fixed_size) as
    // a stack (a block is reused once it and every block above it are
    // freed), never touches the pool, and fails when the region is full;
    // upload() fails instead of growing the arena.
    char* fixed_base = nullptr;
    size_t fixed_size = 0;
    struct FixedBlock {
        size_t off, bytes;
        bool freed;
    };
    std::vector<FixedBlock> fixed_stack;
    // Graph capture of a call's launches: while set, upload() copies the
    // descriptors into cap_host and hands out addresses in the device block
    // cap_dev (cap_size bytes), which the caller fills from cap_host once
    // the capture ends; the arena is not touched.
    char* cap_dev = nullptr;
    size_t cap_size = 0;
    std::vector<char>* cap_host = nullptr;
    // GEMM kernel-class table of the autotuner (la.hip): shape class ->
    // kernel class; gemm_force routes every eligible problem to one class
    // (the autotuner's own timing runs); 0 = the built-in rules
    std::unordered_map<uint32_t, uint8_t> gemm_tune;
    int gemm_force = 0;
    size_t live_bytes() const {
        size_t b = 0;
        for (const auto& kv : live) b += kv.second.bytes;
        return b;
    }

    void set_device() const { KMH_HIP(hipSetDevice(device)); }
    // Copy `bytes` of host data into the arena; returns the device address.
    void* upload(const void* host, size_t bytes);
    template <class T> T* upload_vec(const std::vector<T>& v) {
        return static_cast<T*>(upload(v.data(), v.size() * sizeof(T)));
    }
    bool queued = false;   // listed among the calling thread's contexts with pending launches
    template <class F> void launch(F&& f) {
        pending.emplace_back(std::forward<F>(f));
        if (!queued) note_queued();
        if (pending.size() >= flush_at) {
            flush();
            flush_at = std::min(flush_at * 2, FLUSH_MAX);
        }
    }
    // Queue kernel<<<g, b, shm>>>(args...): the arguments are converted to the
    // kernel's parameter types now and held by value until the launch.
    template <class... P, class... A>
    void launch_kernel(void (*kernel)(P...), dim3 g, dim3 b, unsigned shm, A&&... a) {
        static_assert(sizeof...(P) == sizeof...(A) && sizeof...(P) > 0, "kernel argument count");
        launch([kernel, g, b, shm, s = stream, args = std::tuple<P...>(static_cast<P>(a)...)]() mutable {
            std::apply([&](auto&... x) {
                void* argv[] = {static_cast<void*>(&x)...};
                KMH_HIP(hipLaunchKernel(reinterpret_cast<const void*>(kernel), g, b, argv, shm, s));
            }, args);
        });
    }
    // launch_kernel, timed while profiling is on (see ProfSpan): event
    // markers on the stream right before and right after the launch, so the
    // span starts when the previous work has finished (a start event carried
    // on the dispatch itself would be taken when the dispatch is processed,
    // possibly before its predecessor ends).  A marker pair idles the queue
    // for several microseconds: time a sample (prof_mask, prof_stride).
    template <class... P, class... A>
    void launch_timed(int tag, double amount, double flops, void (*kernel)(P...), dim3 g, dim3 b, unsigned shm,
                      A&&... a) {
        if (!prof || !((prof_mask >> tag) & 1u) || prof_seen[tag]++ % (uint64_t)prof_stride != 0) {
            launch_kernel(kernel, g, b, shm, std::forward<A>(a)...);
            return;
        }
        static_assert(sizeof...(P) == sizeof...(A) && sizeof...(P) > 0, "kernel argument count");
        hipEvent_t e0 = prof_event(), e1 = prof_event();
        spans.push_back({tag, amount, flops, e0, e1, g.x});
        launch([kernel, g, b, shm, s = stream, e0, e1, args = std::tuple<P...>(static_cast<P>(a)...)]() mutable {
            std::apply([&](auto&... x) {
                void* argv[] = {static_cast<void*>(&x)...};
                KMH_HIP(hipEventRecord(e0, s));
                KMH_HIP(hipLaunchKernel(reinterpret_cast<const void*>(kernel), g, b, argv, shm, s));
                KMH_HIP(hipEventRecord(e1, s));
            }, args);
        });
    }
    void note_queued();
    void flush();
    hipStream_t q() { flush(); return stream; }
    void* scratch(size_t bytes);
    void scratch_free(void* p);
    // Return every block handed out since allocation number `since` (after a
    // failed stage: its stage-local blocks; the handles it built are freed
    // by their owners first).
    void scratch_sweep(uint64_t since);
    void sync() {
        flush();
        KMH_HIP(hipStreamSynchronize(stream));
        flush_at = 1;
    }
    // flush + synchronize without throwing (destructors, teardown)
    void drain() noexcept;
    void release_pool() noexcept;
};

// Issue the launches every context queued in this thread (end of a C-ABI
// call), or drop them (a call that failed part-way: what it queued and did
// not issue yet does not run).
void flush_queued();
void discard_queued() noexcept;
void forget_queued(Ctx* c) noexcept;

// Body of a C-ABI entry point: status code from exceptions; the launches the
// call queued are issued before it returns (dropped if it fails).
#define KMH_API_TRY(...)                                                    \
    try {                                                                   \
        __VA_ARGS__;                                                        \
        ::kmh::flush_queued();                                              \
        return KMH_OK;                                                      \
    } catch (const ::kmh::Error& e) {                                       \
        ::kmh::discard_queued();                                            \
        ::kmh::set_last_error(e.what());                                    \
        return e.code;                                                      \
    } catch (const std::bad_alloc&) {                                       \
        ::kmh::discard_queued();                                            \
        ::kmh::set_last_error("host allocation failed");                    \
        return KMH_E_OOM;                                                   \
    } catch (const std::exception& e) {                                     \
        ::kmh::discard_queued();                                            \
        ::kmh::set_last_error(e.what());                                    \
        return KMH_E_DEVICE;                                                \
    }

// Record a kernel launch on the context's queue (see Ctx::launch).
#define KMH_LAUNCH(ctx, kernel, grid, block, ...) (ctx).launch_kernel(kernel, grid, block, 0u, __VA_ARGS__)
// The same, timed under the profiling hooks with a tag, bytes and flops.
#define KMH_LAUNCH_TIMED(ctx, tag, amount, flops, kernel, grid, block, ...) \
    (ctx).launch_timed(tag, amount, flops, kernel, grid, block, 0u, __VA_ARGS__)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Grouped launches of uneven problems: workgroup b works on problem map[b]
// (problems own consecutive workgroups), so a kernel finds its problem with
// one load instead of a search over the descriptors.
inline void map_append(std::vector<int>& map, int problem, int64_t count) {
    map.insert(map.end(), (size_t)count, problem);
}

}  // namespace kmh

// Opaque C-ABI context handle.  HODLR handles and servers hold their
// context: kmh_ctx_destroy with holders left only marks it closing, and the
// last holder's destroy completes the teardown (kmh_ctx_release).
struct kmh_ctx {
    kmh::Ctx c;
    int holders = 0;
    bool closing = false;
};
namespace kmh {
kmh_ctx* ctx_create(int device, uint64_t budget);   // a new context (stream, arena, flags)
void ctx_teardown(kmh_ctx* ctx) noexcept;           // frees the context's device state and ctx itself
}
inline void kmh_ctx_hold(kmh_ctx* ctx) { ++ctx->holders; }
inline void kmh_ctx_release(kmh_ctx* ctx) noexcept {
    if (--ctx->holders == 0 && ctx->closing) kmh::ctx_teardown(ctx);
}
