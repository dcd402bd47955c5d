// Context and device-buffer management for libkmhodlr.
#include <algorithm>

#include "common.h"
#include "serve.h"

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

// Handle buffers come from the context's pool: on the context's one
// in-order stream a block can be reused by later work as soon as it is
// released, so neither allocation nor release touches the device.
void DevBuf::alloc(Ctx* c, size_t nbytes) {
    release();
    if (nbytes == 0) return;
    KMH_CHECK(c->budget == 0 || c->in_use + nbytes <= c->budget, KMH_E_BUDGET,
              "HBM budget exceeded: need " + std::to_string(nbytes) + " more bytes, " +
                  std::to_string(c->budget - std::min<uint64_t>(c->budget, c->in_use)) + " available");
    p = c->scratch(nbytes);
    bytes = nbytes;
    ctx = c;
    c->in_use += nbytes;
}

void DevBuf::release() {
    if (p) {
        try {
            ctx->scratch_free(p);
        } catch (...) {   // a block unknown to the pool is left alone
        }
        ctx->in_use -= bytes;
    }
    p = nullptr; bytes = 0; ctx = nullptr;
}

// The staging mirror and the arena are a ring of ARENA_SEGS segments.
// Leaving a segment issues every queued launch and records an event after
// them; entering a segment again waits for its event, i.e. for the copies
// and kernels that read the segment's previous contents (issued at least a
// whole ring of uploads earlier, so the wait rarely blocks).
void* Ctx::upload(const void* host, size_t bytes) {
    const size_t need = (bytes + 255) & ~size_t(255);
    if (cap_dev) {
        const size_t off = cap_host->size();
        KMH_CHECK(off + need <= cap_size, KMH_E_BAD_ARG, "graph capture: descriptor block full");
        cap_host->resize(off + need);
        if (bytes) std::memcpy(cap_host->data() + off, host, bytes);
        return cap_dev + off;
    }
    const size_t seg = arena_size / ARENA_SEGS;
    if (need > seg) {
        KMH_CHECK(!fixed_base, KMH_E_BAD_ARG,
                  "descriptor upload of " + std::to_string(bytes) + " bytes needs a larger arena (allocation-free call)");
        // grow (rare): drain the stream, reallocate both sides.  The old
        // arena stays allocated until teardown: descriptors uploaded into it
        // by the caller for a launch not queued yet are still read from it.
        sync();
        size_t ns = std::max<size_t>(arena_size * 2, need * ARENA_SEGS * 2);
        ns = (ns + 255 * ARENA_SEGS) / (256 * ARENA_SEGS) * (256 * ARENA_SEGS);
        if (arena) retired_arenas.push_back(arena);
        if (staging) KMH_HIP(hipHostFree(staging));
        arena = nullptr;
        staging = nullptr;
        KMH_HIP(hipMalloc(&arena, ns));
        KMH_HIP(hipHostMalloc(&staging, ns, hipHostMallocDefault));
        arena_size = ns;
        arena_off = dirty_lo = dirty_hi = 0;
        arena_seg = 0;
        for (auto& f : seg_fence) f.pending = false;
        return upload(host, bytes);
    }
    const int cur = arena_seg;
    if (arena_off + need > (size_t)(cur + 1) * seg) {
        flush();
        Fence& leave = seg_fence[cur];
        if (!leave.ev) KMH_HIP(hipEventCreateWithFlags(&leave.ev, hipEventDisableTiming));
        KMH_HIP(hipEventRecord(leave.ev, stream));
        leave.pending = true;
        const int nxt = (cur + 1) % ARENA_SEGS;
        Fence& enter = seg_fence[nxt];
        if (enter.pending) {
            KMH_HIP(hipEventSynchronize(enter.ev));
            enter.pending = false;
        }
        arena_seg = nxt;
        arena_off = dirty_lo = dirty_hi = (size_t)nxt * seg;
    }
    if (bytes) std::memcpy(staging + arena_off, host, bytes);
    if (dirty_hi == dirty_lo) dirty_lo = arena_off;
    dirty_hi = arena_off + need;
    char* d = arena + arena_off;
    arena_off += need;
    return d;
}

// The descriptor copy of a flush: the dirty range of the pinned staging
// mirror (device-accessible host memory) into the arena, read over the host
// link by a kernel on the context's stream.  System-scope loads, so that no
// device cache serves a line of an earlier use of the same staging bytes.
// A kernel rather than a runtime host-to-device copy: such copies issued
// from two host threads at once (server lanes) took ~64 us of host time
// each, against ~3 us from one thread and ~11 us for a kernel launch.
__global__ __launch_bounds__(256) void k_fetch_desc(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                    size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = __hip_atomic_load(const_cast<uint64_t*>(src + i), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

void Ctx::flush() {
    if (dirty_hi > dirty_lo) {
        // dirty_lo, dirty_hi and the arena are multiples of 256 bytes
        const size_t n = (dirty_hi - dirty_lo) / sizeof(uint64_t);
        const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
        hipLaunchKernelGGL(k_fetch_desc, dim3(g), dim3(256), 0, stream,
                           reinterpret_cast<const uint64_t*>(staging + dirty_lo),
                           reinterpret_cast<uint64_t*>(arena + dirty_lo), n);
        KMH_HIP(hipGetLastError());
        dirty_lo = dirty_hi = arena_off;
    }
    if (pending.empty()) return;
    std::vector<std::function<void()>> run;
    run.swap(pending);
    for (auto& f : run) {
        f();
        KMH_HIP(hipGetLastError());
    }
}

static thread_local std::vector<Ctx*> g_queued;

void Ctx::note_queued() {
    queued = true;
    g_queued.push_back(this);
}

void flush_queued() {
    while (!g_queued.empty()) {
        Ctx* c = g_queued.back();
        g_queued.pop_back();
        c->queued = false;
        c->flush();
    }
}

void discard_queued() noexcept {
    for (Ctx* c : g_queued) {
        c->queued = false;
        c->pending.clear();
    }
    g_queued.clear();
}

void forget_queued(Ctx* c) noexcept {
    for (size_t i = 0; i < g_queued.size(); ++i)
        if (g_queued[i] == c) {
            g_queued.erase(g_queued.begin() + (long)i);
            break;
        }
    c->queued = false;
}

void Ctx::drain() noexcept {
    try {
        flush();
    } catch (...) {
        pending.clear();
    }
    (void)hipStreamSynchronize(stream);
    flush_at = 1;
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
    if (fail_scratch_after >= 0 && fail_scratch_after-- == 0)
        throw Error(KMH_E_OOM, "scratch allocation of " + std::to_string(bytes) + " bytes failed (injected)");
    bytes = std::max<size_t>(256, (bytes + 255) & ~size_t(255));
    if (fixed_base) {
        const size_t off = fixed_stack.empty() ? 0 : fixed_stack.back().off + fixed_stack.back().bytes;
        KMH_CHECK(off + bytes <= fixed_size, KMH_E_BAD_ARG,
                  "the fixed scratch of " + std::to_string(fixed_size) + " bytes is too small for this call");
        fixed_stack.push_back({off, bytes, false});
        return fixed_base + off;
    }
    auto it = free_blocks.begin();
    for (; it != free_blocks.end(); ++it)
        if (it->second.first >= bytes) break;
    if (it == free_blocks.end()) {
        // new chunk: at least double the pool, at least 64 MiB
        const size_t sz = std::max<size_t>({bytes, pool_bytes, size_t(64) << 20});
        void* q = nullptr;
        set_device();
        hipError_t e = hipMalloc(&q, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(KMH_E_OOM, "scratch allocation of " + std::to_string(bytes) + " bytes failed");
        }
        chunks.push_back({static_cast<char*>(q), sz});
        pool_bytes += sz;
        it = free_blocks.emplace(static_cast<char*>(q), std::make_pair(sz, (int)chunks.size() - 1)).first;
    }
    char* p = it->first;
    const size_t avail = it->second.first;
    const int chunk = it->second.second;
    free_blocks.erase(it);
    if (avail > bytes) free_blocks.emplace(p + bytes, std::make_pair(avail - bytes, chunk));
    live.emplace(p, LiveBlock{bytes, chunk, alloc_seq++});
    return p;
}

void ctx_reserve(Ctx& c, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    for (const auto& fb : c.free_blocks)
        if (fb.second.first >= bytes) return;
    void* q = nullptr;
    c.set_device();
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(KMH_E_OOM, "reserving " + std::to_string(bytes) + " bytes of device memory failed");
    }
    c.chunks.push_back({static_cast<char*>(q), bytes});
    c.pool_bytes += bytes;
    c.free_blocks.emplace(static_cast<char*>(q), std::make_pair(bytes, (int)c.chunks.size() - 1));
}

void Ctx::scratch_free(void* vp) {
    if (!vp) return;
    char* cp = static_cast<char*>(vp);
    if (fixed_base && cp >= fixed_base && cp < fixed_base + fixed_size) {
        for (auto& b : fixed_stack)
            if (fixed_base + b.off == cp) b.freed = true;
        while (!fixed_stack.empty() && fixed_stack.back().freed) fixed_stack.pop_back();
        return;
    }
    auto lt = live.find(vp);
    KMH_CHECK(lt != live.end(), KMH_E_BAD_ARG, "scratch_free of an unknown block");
    char* p = static_cast<char*>(vp);
    size_t bytes = lt->second.bytes;
    const int chunk = lt->second.chunk;
    live.erase(lt);
    // coalesce with the free neighbours of the same chunk
    auto nx = free_blocks.lower_bound(p);
    if (nx != free_blocks.end() && nx->first == p + bytes && nx->second.second == chunk) {
        bytes += nx->second.first;
        nx = free_blocks.erase(nx);
    }
    if (nx != free_blocks.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second.first == p && pv->second.second == chunk) {
            pv->second.first += bytes;
            return;
        }
    }
    free_blocks.emplace(p, std::make_pair(bytes, chunk));
}

void Ctx::scratch_sweep(uint64_t since) {
    std::vector<void*> stale;
    for (const auto& kv : live)
        if (kv.second.seq >= since) stale.push_back(kv.first);
    for (void* p : stale) scratch_free(p);
}

void Ctx::release_pool() noexcept {
    drain();
    for (auto& ch : chunks) (void)hipFree(ch.first);
    chunks.clear();
    free_blocks.clear();
    live.clear();
    pool_bytes = 0;
}

void ctx_teardown(kmh_ctx* ctx) noexcept {
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    c.release_pool();
    forget_queued(&c);
    for (auto e : c.ev_pool) (void)hipEventDestroy(e);
    for (auto& f : c.seg_fence)
        if (f.ev) (void)hipEventDestroy(f.ev);
    if (c.arena) (void)hipFree(c.arena);
    for (char* a : c.retired_arenas) (void)hipFree(a);
    if (c.staging) (void)hipHostFree(c.staging);
    if (c.d_info) (void)hipFree(c.d_info);
    (void)hipStreamDestroy(c.stream);
    delete ctx;
}

}  // namespace kmh
