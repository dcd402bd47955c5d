// Test hooks (include/kmhodlr_testing.h): one-matrix wrappers around the
// batched primitives, host buffers in and out.
#include "kmhodlr_testing.h"
#include "la.h"


namespace {

struct Dev {
    kmh::Ctx& c;
    double* p = nullptr;
    size_t bytes = 0;
    Dev(kmh::Ctx& cc, const double* host, size_t count) : c(cc) {
        bytes = std::max<size_t>(8, count * sizeof(double));
        p = static_cast<double*>(c.scratch(bytes));
        if (host && count) KMH_HIP(hipMemcpyAsync(p, host, count * sizeof(double), hipMemcpyHostToDevice, c.q()));
    }
    void to_host(double* host, size_t count) {
        KMH_HIP(hipMemcpyAsync(host, p, count * sizeof(double), hipMemcpyDeviceToHost, c.q()));
    }
    ~Dev() { c.scratch_free(p); }
};

// one workgroup checks the pattern word i = seed * 1000003 + i of its upload
__global__ __launch_bounds__(256) void k_check_upload(const unsigned long long* __restrict__ d, int words,
                                                      unsigned long long seed, unsigned long long* __restrict__ bad) {
    int wrong = 0;
    for (int i = threadIdx.x; i < words; i += 256) wrong |= d[i] != seed * 1000003ull + (unsigned long long)i;
    wrong = __syncthreads_or(wrong);
    if (threadIdx.x == 0 && wrong) atomicAdd(bad, 1ull);
}

}  // namespace

#define TRY_(...) KMH_API_TRY(__VA_ARGS__)

extern "C" {

int kmh_test_gemm(kmh_ctx* ctx, int32_t ta, int32_t tb, int32_t m, int32_t n, int32_t k, const double* A, int32_t lda,
                  const double* B, int32_t ldb, double* Cm, int32_t ldc, double alpha, double beta, int32_t lower) {
    TRY_({
        kmh::Ctx& c = ctx->c;
        c.set_device();
        const size_t na = (size_t)lda * (ta ? m : k), nb = (size_t)ldb * (tb ? k : n), nc = (size_t)ldc * n;
        Dev a(c, A, na), b(c, B, nb), cc(c, Cm, nc);
        kmh::gemm(c, ta != 0, tb != 0, {{m, n, k, a.p, lda, b.p, ldb, cc.p, ldc, alpha, beta, lower}});
        cc.to_host(Cm, nc);
        c.sync();
    })
}

int kmh_test_capacitance(kmh_ctx* ctx, int32_t r, const double* R1, const double* R2, int32_t ldr, double* X,
                         double* G, int32_t ld, double* logsum) {
    TRY_({
        kmh::Ctx& c = ctx->c;
        c.set_device();
        KMH_CHECK(r >= 0 && r <= kmh::CAP_MAX && ldr >= r && ld >= 2 * r && logsum, KMH_E_BAD_ARG,
                  "kmh_test_capacitance: invalid shape");
        KMH_HIP(hipMemsetAsync(c.d_info, 0, sizeof(int), c.q()));
        const size_t nr = (size_t)ldr * r, nx = (size_t)ld * 2 * r;
        Dev a(c, R1, nr), b(c, R2, nr), x(c, X, nx), g(c, G, nx), part(c, nullptr, 1);
        kmh::capacitance(c, {{a.p, b.p, ldr, r, x.p, g.p, ld, part.p, c.d_info}});
        x.to_host(X, nx);
        g.to_host(G, nx);
        part.to_host(logsum, 1);
        int info = 0;
        KMH_HIP(hipMemcpyAsync(&info, c.d_info, sizeof(int), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        if (info) {
            KMH_HIP(hipMemsetAsync(c.d_info, 0, sizeof(int), c.q()));
            c.sync();
            throw kmh::Error(KMH_E_NOT_PD, "PosDefException: capacitance block is not positive definite");
        }
    })
}

int kmh_test_potrf(kmh_ctx* ctx, int32_t n, double* A, int32_t lda) {
    TRY_({
        kmh::Ctx& c = ctx->c;
        c.set_device();
        KMH_HIP(hipMemsetAsync(c.d_info, 0, sizeof(int), c.q()));
        Dev a(c, A, (size_t)lda * n);
        kmh::potrf(c, {{a.p, n, lda}});
        a.to_host(A, (size_t)lda * n);
        int info = 0;
        KMH_HIP(hipMemcpyAsync(&info, c.d_info, sizeof(int), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        if (info) {
            KMH_HIP(hipMemsetAsync(c.d_info, 0, sizeof(int), c.q()));
            c.sync();
            throw kmh::Error(KMH_E_NOT_PD, "PosDefException: matrix is not positive definite");
        }
    })
}

int kmh_test_trsm(kmh_ctx* ctx, int32_t trans, int32_t n, const double* Lm, int32_t ldl, int32_t ncols, double* B,
                  int32_t ldb) {
    TRY_({
        kmh::Ctx& c = ctx->c;
        c.set_device();
        Dev l(c, Lm, (size_t)ldl * n), b(c, B, (size_t)ldb * ncols);
        kmh::trsm_left(c, trans != 0, {{l.p, n, ldl, b.p, ncols, ldb}});
        b.to_host(B, (size_t)ldb * ncols);
        c.sync();
    })
}

int kmh_test_qr(kmh_ctx* ctx, int32_t m, int32_t r, double* A, int32_t lda, double* R, int32_t ldr) {
    TRY_({
        kmh::Ctx& c = ctx->c;
        c.set_device();
        Dev a(c, A, (size_t)lda * r), rr(c, nullptr, (size_t)ldr * r);
        kmh::qr_panels(c, {{a.p, m, r, lda, rr.p, ldr}});
        a.to_host(A, (size_t)lda * r);
        rr.to_host(R, (size_t)ldr * r);
        c.sync();
    })
}

int kmh_test_arena(kmh_ctx* ctx, int32_t count, int32_t words, int64_t* bad) {
    TRY_({
        KMH_CHECK(ctx && bad && count >= 0 && words >= 1, KMH_E_BAD_ARG, "invalid arena test arguments");
        kmh::Ctx& c = ctx->c;
        c.set_device();
        unsigned long long* d_bad = static_cast<unsigned long long*>(c.scratch(sizeof(unsigned long long)));
        KMH_HIP(hipMemsetAsync(d_bad, 0, sizeof(unsigned long long), c.q()));
        std::vector<unsigned long long> h(words);
        for (int q = 0; q < count; ++q) {
            for (int i = 0; i < words; ++i) h[i] = (unsigned long long)q * 1000003ull + (unsigned long long)i;
            const unsigned long long* d = c.upload_vec(h);
            KMH_LAUNCH(c, k_check_upload, dim3(1), dim3(256), d, words, (unsigned long long)q, d_bad);
        }
        unsigned long long hb = 0;
        KMH_HIP(hipMemcpyAsync(&hb, d_bad, sizeof(hb), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        c.scratch_free(d_bad);
        *bad = (int64_t)hb;
    })
}

int kmh_test_profile_spans(kmh_ctx* ctx, int32_t cap, int32_t* tags, double* amounts, double* flops, double* ms,
                           int32_t* grids, int32_t* count) {
    TRY_({
        KMH_CHECK(ctx && count && cap >= 0 && (cap == 0 || (tags && amounts && flops && ms && grids)),
                  KMH_E_BAD_ARG, "null argument");
        kmh::Ctx& c = ctx->c;
        KMH_CHECK(c.prof, KMH_E_STATE, "no open profiling window");
        c.set_device();
        c.sync();
        *count = (int32_t)c.spans.size();
        for (int32_t i = 0; i < cap && i < (int32_t)c.spans.size(); ++i) {
            const auto& sp = c.spans[(size_t)i];
            float t = 0.f;
            KMH_HIP(hipEventElapsedTime(&t, sp.a, sp.b));
            tags[i] = sp.tag;
            amounts[i] = sp.amount;
            flops[i] = sp.flops;
            ms[i] = t;
            grids[i] = (int32_t)sp.grid;
        }
    })
}

int kmh_test_fail_scratch(kmh_ctx* ctx, int64_t after) {
    TRY_({
        KMH_CHECK(ctx, KMH_E_BAD_ARG, "null context");
        ctx->c.fail_scratch_after = after < 0 ? -1 : after;
    })
}

int kmh_test_scratch_live(kmh_ctx* ctx, uint64_t* bytes) {
    TRY_({
        KMH_CHECK(ctx && bytes, KMH_E_BAD_ARG, "null argument");
        *bytes = (uint64_t)ctx->c.live_bytes();
    })
}

int kmh_test_pool_allocs(kmh_ctx* ctx, uint64_t* count) {
    TRY_({
        KMH_CHECK(ctx && count, KMH_E_BAD_ARG, "null argument");
        *count = ctx->c.alloc_seq;
    })
}

int kmh_test_arena_shrink(kmh_ctx* ctx, uint64_t bytes) {
    TRY_({
        KMH_CHECK(ctx, KMH_E_BAD_ARG, "null context");
        kmh::Ctx& c = ctx->c;
        const size_t segb = 256 * kmh::Ctx::ARENA_SEGS;
        KMH_CHECK(bytes >= segb && bytes % segb == 0, KMH_E_BAD_ARG, "arena size must be a multiple of 1 KiB");
        c.set_device();
        c.sync();
        char* na = nullptr;
        char* ns = nullptr;
        KMH_HIP(hipMalloc(&na, bytes));
        KMH_HIP(hipHostMalloc(&ns, bytes, hipHostMallocDefault));
        if (c.arena) c.retired_arenas.push_back(c.arena);
        if (c.staging) KMH_HIP(hipHostFree(c.staging));
        c.arena = na;
        c.staging = ns;
        c.arena_size = bytes;
        c.arena_off = c.dirty_lo = c.dirty_hi = 0;
        c.arena_seg = 0;
        for (auto& f : c.seg_fence) f.pending = false;
    })
}

}  // extern "C"
