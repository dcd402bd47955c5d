// C ABI entry points of libkmhodlr (declared in include/kmhodlr.h).
#include <memory>

#include "deriv.h"
#include "hodlr.h"
#include "la.h"
#include "serve.h"

using kmh::Error;

#define KMH_TRY(...) KMH_API_TRY(__VA_ARGS__)

static kmh::Hodlr& H_of(kmh_hodlr* h) {
    KMH_CHECK(h && h->h, KMH_E_BAD_ARG, "null HODLR handle");
    h->h->ctx->set_device();
    return *h->h;
}

// Host wrappers: stage x into device memory, run, copy back.
template <class F>
static void with_device_rhs(kmh::Hodlr& H, const double* x, double* y, int64_t nrhs, bool two, F f) {
    kmh::Ctx& c = *H.ctx;
    KMH_CHECK(x && y, KMH_E_BAD_ARG, "null vector");
    KMH_CHECK(nrhs >= 1 && nrhs < (int64_t(1) << 20), KMH_E_BAD_ARG, "invalid number of right-hand sides");
    const size_t bytes = (size_t)H.n * nrhs * sizeof(double);
    double* dx = static_cast<double*>(c.scratch(bytes));
    double* dy = two ? static_cast<double*>(c.scratch(bytes)) : dx;
    KMH_HIP(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, c.q()));
    f(dx, dy);
    KMH_HIP(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, c.q()));
    if (two) c.scratch_free(dy);
    c.scratch_free(dx);
    c.sync();
}

namespace kmh {
kmh_ctx* ctx_create(int device, uint64_t budget) {
    int n = 0;
    KMH_HIP(hipGetDeviceCount(&n));
    KMH_CHECK(device >= 0 && device < n, KMH_E_BAD_ARG, "device index out of range");
    std::unique_ptr<kmh_ctx> c(new kmh_ctx());
    c->c.device = device;
    c->c.budget = budget;
    KMH_HIP(hipSetDevice(device));
    KMH_HIP(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
    try {
        c->c.arena_size = size_t(64) << 20;
        KMH_HIP(hipMalloc(&c->c.arena, c->c.arena_size));
        KMH_HIP(hipHostMalloc(&c->c.staging, c->c.arena_size, hipHostMallocDefault));
        KMH_HIP(hipMalloc(&c->c.d_info, 256));
        KMH_HIP(hipMemset(c->c.d_info, 0, 256));
    } catch (...) {
        ctx_teardown(c.release());
        throw;
    }
    (void)hipGetLastError();
    return c.release();
}
}  // namespace kmh

extern "C" {

const char* kmh_last_error(void) { return kmh::last_error_cstr(); }

int kmh_device_count(int* count) {
    KMH_TRY({
        KMH_CHECK(count, KMH_E_BAD_ARG, "null output");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            n = 0;
        }
        *count = n;
    })
}


int kmh_ctx_create(int device, uint64_t hbm_budget_bytes, kmh_ctx** out) {
    KMH_TRY({
        KMH_CHECK(out, KMH_E_BAD_ARG, "null output");
        *out = kmh::ctx_create(device, hbm_budget_bytes);
    })
}

int kmh_ctx_destroy(kmh_ctx* ctx) {
    KMH_TRY({
        if (!ctx) return KMH_OK;
        KMH_CHECK(!ctx->closing, KMH_E_STATE, "context already destroyed");
        ctx->closing = true;
        if (ctx->holders == 0) kmh::ctx_teardown(ctx);
    })
}

int kmh_ctx_synchronize(kmh_ctx* ctx) {
    KMH_TRY({
        KMH_CHECK(ctx, KMH_E_BAD_ARG, "null context");
        ctx->c.set_device();
        ctx->c.sync();
    })
}

int kmh_ctx_bytes_in_use(kmh_ctx* ctx, uint64_t* out) {
    KMH_TRY({
        KMH_CHECK(ctx && out, KMH_E_BAD_ARG, "null argument");
        *out = ctx->c.in_use;
    })
}

int kmh_ctx_reserve(kmh_ctx* ctx, uint64_t bytes) {
    KMH_TRY({
        KMH_CHECK(ctx, KMH_E_BAD_ARG, "null context");
        kmh::ctx_reserve(ctx->c, (size_t)bytes);
    })
}

int kmh_ctx_profile_begin(kmh_ctx* ctx) {
    KMH_TRY({
        KMH_CHECK(ctx, KMH_E_BAD_ARG, "null context");
        kmh::Ctx& c = ctx->c;
        c.set_device();
        c.sync();
        c.prof = true;
        c.ev_used = 0;
        c.spans.clear();
        for (auto& t : c.prof_sum) t[0] = t[1] = t[2] = 0.0;
        for (auto& k : c.prof_seen) k = 0;
    })
}

int kmh_ctx_profile_config(kmh_ctx* ctx, uint32_t tag_mask, int32_t stride) {
    KMH_TRY({
        KMH_CHECK(ctx, KMH_E_BAD_ARG, "null context");
        KMH_CHECK(stride >= 1, KMH_E_BAD_ARG, "profile stride must be >= 1");
        KMH_CHECK(tag_mask != 0 && tag_mask < (1u << KMH_PROF_TAGS), KMH_E_BAD_ARG, "profile tag mask out of range");
        KMH_CHECK(!ctx->c.prof, KMH_E_STATE, "profile_config inside a begin..end window");
        ctx->c.prof_mask = tag_mask;
        ctx->c.prof_stride = stride;
    })
}

int kmh_ctx_profile_end(kmh_ctx* ctx, double* gemm_flops, double* gemm_ms, int64_t* gemm_launches) {
    KMH_TRY({
        KMH_CHECK(ctx && gemm_flops && gemm_ms && gemm_launches, KMH_E_BAD_ARG, "null argument");
        kmh::Ctx& c = ctx->c;
        c.set_device();
        c.sync();
        for (auto& t : c.prof_sum) t[0] = t[1] = t[2] = 0.0;
        for (const auto& sp : c.spans) {
            float t = 0.f;
            KMH_HIP(hipEventElapsedTime(&t, sp.a, sp.b));
            if (sp.flops >= 0.0) {
                c.prof_sum[KMH_PROF_GEMM][0] += sp.flops;
                c.prof_sum[KMH_PROF_GEMM][1] += t;
                c.prof_sum[KMH_PROF_GEMM][2] += 1.0;
            }
            if (sp.tag != KMH_PROF_GEMM) {
                c.prof_sum[sp.tag][0] += sp.amount;
                c.prof_sum[sp.tag][1] += t;
                c.prof_sum[sp.tag][2] += 1.0;
            }
        }
        *gemm_flops = c.prof_sum[KMH_PROF_GEMM][0];
        *gemm_ms = c.prof_sum[KMH_PROF_GEMM][1];
        *gemm_launches = (int64_t)c.prof_sum[KMH_PROF_GEMM][2];
        c.prof = false;
        c.ev_used = 0;
        c.spans.clear();
    })
}

int kmh_ctx_profile_query(kmh_ctx* ctx, int32_t tag, double* amount, double* ms, int64_t* spans) {
    KMH_TRY({
        KMH_CHECK(ctx && amount && ms && spans, KMH_E_BAD_ARG, "null argument");
        KMH_CHECK(tag >= 0 && tag < KMH_PROF_TAGS, KMH_E_BAD_ARG, "unknown profiling tag");
        const kmh::Ctx& c = ctx->c;
        *amount = c.prof_sum[tag][0];
        *ms = c.prof_sum[tag][1];
        *spans = (int64_t)c.prof_sum[tag][2];
    })
}

static int build_common(kmh_ctx* ctx, const double* pts, bool dev, int64_t n, int32_t d, const kmh_kernel* kernel,
                        const kmh_build_opts* opts, kmh_hodlr** out) {
    KMH_TRY({
        KMH_CHECK(ctx && kernel && opts && out, KMH_E_BAD_ARG, "null argument");
        std::unique_ptr<kmh_hodlr> h(new kmh_hodlr(ctx));
        h->h.reset(kmh::hodlr_build(ctx->c, pts, dev, n, d, *kernel, *opts));
        *out = h.release();
    })
}

int kmh_hodlr_build(kmh_ctx* ctx, const double* pts, int64_t n, int32_t d, const kmh_kernel* kernel,
                    const kmh_build_opts* opts, kmh_hodlr** out) {
    return build_common(ctx, pts, false, n, d, kernel, opts, out);
}

int kmh_hodlr_build_dev(kmh_ctx* ctx, const double* pts_dev, int64_t n, int32_t d, const kmh_kernel* kernel,
                        const kmh_build_opts* opts, kmh_hodlr** out) {
    return build_common(ctx, pts_dev, true, n, d, kernel, opts, out);
}

int kmh_hodlr_destroy(kmh_hodlr* h) {
    KMH_TRY({
        if (!h) return KMH_OK;
        KMH_CHECK(!h->borrowed, KMH_E_STATE,
                  "this handle is a served request's resident factor: release it with kmh_server_release");
        if (h->h) h->h->ctx->set_device();
        delete h;
    })
}

int kmh_hodlr_synchronize(kmh_hodlr* h) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        H.ctx->sync();
    })
}

int kmh_hodlr_info(kmh_hodlr* h, kmh_info* out) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(out, KMH_E_BAD_ARG, "null output");
        out->n = H.n;
        out->levels = H.L;
        out->nleaves = H.nleaves();
        out->factorized = H.factorized ? 1 : 0;
        out->method = H.opts.method;
        out->device_bytes = (int64_t)H.device_bytes();
    })
}

int kmh_hodlr_leaf_offsets(kmh_hodlr* h, int64_t* offsets) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(offsets, KMH_E_BAD_ARG, "null output");
        for (size_t i = 0; i < H.leaf_off.size(); ++i) offsets[i] = H.leaf_off[i];
    })
}

int kmh_hodlr_ranks(kmh_hodlr* h, int32_t* ranks) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(ranks, KMH_E_BAD_ARG, "null output");
        size_t q = 0;
        for (int l = 0; l < H.L; ++l)
            for (int r : H.ranks[l]) ranks[q++] = r;
    })
}

int kmh_hodlr_full(kmh_hodlr* h, double* out) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(out, KMH_E_BAD_ARG, "null output");
        kmh::hodlr_full_host(H, out);
    })
}

int kmh_hodlr_block_uv(kmh_hodlr* h, int32_t level, int32_t block, double* U, double* V) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(U && V, KMH_E_BAD_ARG, "null output");
        kmh::hodlr_block_uv_host(H, level, block, U, V);
    })
}

int kmh_hodlr_matvec(kmh_hodlr* h, const double* x, double* y, int64_t nrhs) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        with_device_rhs(H, x, y, nrhs, true, [&](double* dx, double* dy) { kmh::hodlr_matvec(H, dx, dy, (int)nrhs); });
    })
}

int kmh_hodlr_matvec_dev(kmh_hodlr* h, const double* x, double* y, int64_t nrhs) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(x && y && nrhs >= 1, KMH_E_BAD_ARG, "invalid vectors");
        kmh::hodlr_matvec(H, x, y, (int)nrhs);
    })
}

int kmh_hodlr_factorize(kmh_hodlr* h) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        kmh::hodlr_factorize(H);
    })
}

int kmh_hodlr_factor_apply(kmh_hodlr* h, int32_t op, const double* x, double* y, int64_t nrhs) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        with_device_rhs(H, x, y, nrhs, false, [&](double* dx, double*) { kmh::hodlr_factor_apply(H, op, dx, (int)nrhs); });
    })
}

int kmh_hodlr_factor_apply_dev(kmh_hodlr* h, int32_t op, const double* x, double* y, int64_t nrhs) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        KMH_CHECK(x && y && nrhs >= 1, KMH_E_BAD_ARG, "invalid vectors");
        kmh::Ctx& c = *H.ctx;
        if (x != y) kmh::copy2d(c, {{x, (int)H.n, y, (int)H.n, (int)H.n, (int)nrhs, 1.0}});
        kmh::hodlr_factor_apply(H, op, y, (int)nrhs);
    })
}

int kmh_hodlr_solve(kmh_hodlr* h, const double* b, double* x, int64_t nrhs) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        with_device_rhs(H, b, x, nrhs, false, [&](double* dx, double*) { kmh::hodlr_solve(H, dx, (int)nrhs); });
    })
}

int kmh_hodlr_solve_dev(kmh_hodlr* h, const double* b, double* x, int64_t nrhs) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        KMH_CHECK(b && x && nrhs >= 1, KMH_E_BAD_ARG, "invalid vectors");
        kmh::Ctx& c = *H.ctx;
        if (!kmh::solve_dev_graph(H, b, x, nrhs)) {
            if (b != x) kmh::copy2d(c, {{b, (int)H.n, x, (int)H.n, (int)H.n, (int)nrhs, 1.0}});
            kmh::hodlr_solve(H, x, (int)nrhs);
        }
    })
}

int kmh_hodlr_refinement(kmh_hodlr* h, int32_t* iters, double* relres) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(iters && relres, KMH_E_BAD_ARG, "null output");
        *iters = H.ir_iters;
        *relres = H.ir_relres;
    })
}

int kmh_hodlr_logdet(kmh_hodlr* h, double* out) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(out, KMH_E_BAD_ARG, "null output");
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No logdet without factorization.");
        *out = H.logdet;
    })
}

int kmh_hodlr_build_batch(kmh_ctx* ctx, int32_t count, const kmh_build_req* reqs, kmh_hodlr** out,
                          int32_t* status) {
    KMH_TRY({
        KMH_CHECK(ctx && reqs && out && status && count >= 0, KMH_E_BAD_ARG, "null argument");
        std::vector<kmh::BuildReq> rq;
        for (int i = 0; i < count; ++i)
            rq.push_back({reqs[i].pts, reqs[i].on_device != 0, reqs[i].n, reqs[i].d, reqs[i].kernel, reqs[i].opts});
        std::vector<kmh::Status> st;
        auto hs = kmh::hodlr_build_batch(ctx->c, rq, st);
        int first = KMH_OK;
        for (int i = 0; i < count; ++i) {
            status[i] = st[i].code;
            out[i] = nullptr;
            if (hs[i]) {
                out[i] = new kmh_hodlr(ctx);
                out[i]->h.reset(hs[i]);
            } else if (first == KMH_OK) {
                first = st[i].code;
                kmh::set_last_error(st[i].msg);
            }
        }
        if (first != KMH_OK) return first;
    })
}

int kmh_hodlr_factorize_batch(kmh_ctx* ctx, int32_t count, kmh_hodlr* const* hs, int32_t* status) {
    KMH_TRY({
        KMH_CHECK(ctx && hs && status && count >= 0, KMH_E_BAD_ARG, "null argument");
        std::vector<kmh::Hodlr*> v;
        for (int i = 0; i < count; ++i) {
            KMH_CHECK(hs[i] && hs[i]->h && hs[i]->h->ctx == &ctx->c, KMH_E_BAD_ARG, "handle from another context");
            v.push_back(hs[i]->h.get());
        }
        std::vector<kmh::Status> st;
        kmh::hodlr_factorize_batch(ctx->c, v, st);
        int first = KMH_OK;
        for (int i = 0; i < count; ++i) {
            status[i] = st[i].code;
            if (st[i].code != KMH_OK && first == KMH_OK) {
                first = st[i].code;
                kmh::set_last_error(st[i].msg);
            }
        }
        if (first != KMH_OK) return first;
    })
}

int kmh_hodlr_solve_batch_dev(kmh_ctx* ctx, int32_t count, kmh_hodlr* const* hs, const double* const* b,
                              double* const* x, const int64_t* nrhs) {
    KMH_TRY({
        KMH_CHECK(ctx && hs && b && x && nrhs && count >= 0, KMH_E_BAD_ARG, "null argument");
        kmh::Ctx& c = ctx->c;
        c.set_device();
        std::vector<kmh::Hodlr*> v;
        std::vector<double*> T;
        std::vector<int> k;
        std::vector<kmh::Copy2D> cp;
        for (int i = 0; i < count; ++i) {
            KMH_CHECK(hs[i] && hs[i]->h && hs[i]->h->ctx == &c, KMH_E_BAD_ARG, "handle from another context");
            kmh::Hodlr& H = *hs[i]->h;
            KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
            KMH_CHECK(nrhs[i] >= 1 && b[i] && x[i], KMH_E_BAD_ARG, "invalid right-hand side");
            if (b[i] != x[i]) cp.push_back({b[i], (int)H.n, x[i], (int)H.n, (int)H.n, (int)nrhs[i], 1.0});
            v.push_back(&H);
            T.push_back(x[i]);
            k.push_back((int)nrhs[i]);
        }
        kmh::copy2d(c, cp);
        kmh::hodlr_solve_batch(c, v, T, k);
    })
}

int kmh_stoch_gradient(kmh_hodlr* h, const kmh_kernel* dks, int32_t p, const double* y, const double* saa,
                       int64_t nv, double* grad) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized for this to work.");
        KMH_CHECK(dks && y && grad && p >= 1, KMH_E_BAD_ARG, "null argument");
        KMH_CHECK(nv >= 1 && saa, KMH_E_BAD_ARG, "need at least one SAA vector");
        kmh::hodlr_stoch_gradient(H, dks, p, y, saa, nv, grad);
    })
}

int kmh_stoch_gradient_dev(kmh_hodlr* h, const kmh_kernel* dks, int32_t p, const double* y, const double* saa,
                           int64_t nv, double* grad) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized for this to work.");
        KMH_CHECK(dks && y && grad && p >= 1, KMH_E_BAD_ARG, "null argument");
        KMH_CHECK(nv >= 1 && saa, KMH_E_BAD_ARG, "need at least one SAA vector");
        kmh::hodlr_stoch_gradient(H, dks, p, y, saa, nv, grad, false, true);
    })
}

int kmh_stoch_profile_gradient(kmh_hodlr* h, const kmh_kernel* dks, int32_t p, const double* y, const double* saa,
                               int64_t nv, double* grad) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized for this to work.");
        KMH_CHECK(dks && y && grad && p >= 1, KMH_E_BAD_ARG, "null argument");
        KMH_CHECK(nv >= 1 && saa, KMH_E_BAD_ARG, "need at least one SAA vector");
        kmh::hodlr_stoch_gradient(H, dks, p, y, saa, nv, grad, true);
    })
}

int kmh_stoch_fisher(kmh_hodlr* h, const kmh_kernel* dks, int32_t p, const double* saa, int64_t nv, double* F) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized.");
        KMH_CHECK(dks && saa && F && p >= 1 && nv >= 1, KMH_E_BAD_ARG, "null argument");
        kmh::hodlr_stoch_fisher(H, dks, p, saa, nv, F);
    })
}

int kmh_stoch_hessian(kmh_hodlr* h, const kmh_kernel* d1, int32_t p, const kmh_kernel* d2, const double* y,
                      const double* saa, int64_t nv, int32_t exact_second, double* Hout) {
    KMH_TRY({
        kmh::Hodlr& H = H_of(h);
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "The matrix needs to be factorized.");
        KMH_CHECK(d1 && d2 && y && saa && Hout && p >= 1 && nv >= 1, KMH_E_BAD_ARG, "null argument");
        kmh::hodlr_stoch_hessian(H, d1, p, d2, y, saa, nv, exact_second != 0, Hout);
    })
}

}  // extern "C"
