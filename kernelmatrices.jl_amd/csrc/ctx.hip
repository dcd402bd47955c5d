// Context and device-buffer management for libkmhodlr.
#include <algorithm>

#include "common.h"

namespace kmh {

static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
const char* last_error_cstr() { return g_last_error.c_str(); }

DevBuf& DevBuf::operator=(DevBuf&& o) noexcept {
    if (this != &o) {
        release();
        p = o.p; bytes = o.bytes; ctx = o.ctx;
        o.p = nullptr; o.bytes = 0; o.ctx = nullptr;
    }
    return *this;
}

DevBuf::~DevBuf() { release(); }

void DevBuf::alloc(Ctx* c, size_t nbytes) {
    release();
    if (nbytes == 0) return;
    KMH_CHECK(c->budget == 0 || c->in_use + nbytes <= c->budget, KMH_E_BUDGET,
              "HBM budget exceeded: need " + std::to_string(nbytes) + " more bytes, " +
                  std::to_string(c->budget - std::min<uint64_t>(c->budget, c->in_use)) + " available");
    c->set_device();
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, nbytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(KMH_E_OOM, "hipMalloc of " + std::to_string(nbytes) + " bytes failed");
    }
    p = q; bytes = nbytes; ctx = c;
    c->in_use += nbytes;
}

void DevBuf::release() {
    if (p) {
        // frees are ordered after all work queued on the context's stream
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(p);
        ctx->in_use -= bytes;
    }
    p = nullptr; bytes = 0; ctx = nullptr;
}

// The descriptor arena is mirrored by a pinned host staging buffer at the
// same offsets.  Both wrap together; wrapping waits for the stream so that
// neither the pending host->device copies nor the kernels reading older
// descriptors can be overtaken.
void* Ctx::upload(const void* host, size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (need > arena_size / 4) {
        // grow (rare): drain the stream, reallocate both sides
        KMH_HIP(hipStreamSynchronize(stream));
        size_t ns = std::max<size_t>(arena_size * 2, need * 8);
        if (arena) KMH_HIP(hipFree(arena));
        if (staging) KMH_HIP(hipHostFree(staging));
        KMH_HIP(hipMalloc(&arena, ns));
        KMH_HIP(hipHostMalloc(&staging, ns, hipHostMallocDefault));
        arena_size = ns;
        arena_off = 0;
    }
    if (arena_off + need > arena_size) {
        KMH_HIP(hipStreamSynchronize(stream));
        arena_off = 0;
    }
    char* h = staging + arena_off;
    char* d = arena + arena_off;
    if (bytes) std::memcpy(h, host, bytes);
    if (bytes) KMH_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
    arena_off += need;
    return d;
}

hipEvent_t Ctx::prof_event() {
    if (ev_used == ev_pool.size()) {
        hipEvent_t e;
        KMH_HIP(hipEventCreate(&e));
        ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
}

void* Ctx::scratch(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMallocAsync(&p, bytes, stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(KMH_E_OOM, "scratch allocation of " + std::to_string(bytes) + " bytes failed");
    }
    return p;
}

void Ctx::scratch_free(void* p) {
    if (p) KMH_HIP(hipFreeAsync(p, stream));
}

}  // namespace kmh
