// Serving loop for GP likelihood requests with continuous batching.
//
// A request is one nll_objective evaluation (maxlikfunctions.jl:29-50):
// build the Nystrom (or ACA) HODLR, symmetricfactorize!, negloglik and,
// when derivative kernels are given, stoch_gradient.  Requests wait in a
// queue and are admitted while their HBM footprint fits the budget and a
// slot is free.  Each step advances every in-flight request by one stage
// (build, leaf stage, one tree level, finish); the stages of all in-flight
// requests run together in the grouped launches of the batched primitives,
// so requests that arrive while others are on the GPU join at the next step
// and their work overlaps.  A request's result is the same alone and in the
// mix: every reduction has a fixed order independent of the other requests.
//
// Lanes: the in-flight requests are spread over lanes, each with its own
// context (HIP stream, descriptor arena, scratch pool); lane 0 is the
// caller's context.  A step advances the lanes on parallel host threads, so
// one lane's latency-bound phases (small tree levels, host synchronizations,
// launch planning) overlap another lane's bandwidth-bound launches on the
// device.  Lanes change which launches a request shares, never its
// arithmetic.
#include <algorithm>
#include <chrono>
#include <deque>
#include <exception>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

#include "hodlr.h"
#include "la.h"
#include "serve.h"

namespace kmh {

static double now_us() {
    using namespace std::chrono;
    return (double)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() * 1e-3;
}

enum Stage { ST_WAIT = 0, ST_BUILD, ST_LEAF, ST_LEVEL, ST_FINISH, ST_DONE };

struct Req {
    int64_t id = 0;
    int stage = ST_WAIT;
    int lv = 0;
    int lane = 0;
    std::vector<double> pts, y, saa;
    const double* dpts = nullptr;   // device-resident inputs (not owned)
    const double* dy = nullptr;
    const double* dsaa = nullptr;
    int64_t n = 0;
    int nv = 0;
    kmh_kernel kern{};
    kmh_build_opts opts{};
    std::vector<kmh_kernel> dks;
    std::unique_ptr<Hodlr> H;
    uint64_t footprint = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    Status st;
    double nll = 0, logdet = 0;
    std::vector<double> grad;
    bool keep = false;               // keep the factor resident after completion
    bool read = false;               // result read (kmh_server_result)
    std::unique_ptr<kmh_hodlr> handle;   // the resident factor, lent out by kmh_server_factor
};

struct Lane {
    kmh_ctx* own = nullptr;   // private context (lanes after the first)
    kmh_ctx* kctx = nullptr;  // the C-ABI context of the lane (the caller's for lane 0)
    Ctx* ctx = nullptr;
    std::vector<Req*> inflight;
};

struct Server {
    Ctx* ctx = nullptr;       // the caller's context (lane 0)
    std::vector<Lane> lanes;
    int max_inflight = 1;
    uint64_t budget = 0;
    uint64_t reserved = 0;
    int64_t next_id = 1;
    std::deque<Req*> waiting;
    std::vector<Req*> inflight;
    std::map<int64_t, std::unique_ptr<Req>> all;
};

// Device bytes a request holds while in flight (FP64): points, leaves,
// Nystrom panels, level slabs, per-block factors and the transient cascade
// panels.  ACA ranks are not known before the build: the estimate takes the
// largest rank the build's first panels hold (256, or maxrank, or the block
// size if smaller) on every level, and counts the build's transient panels
// (one n x capacity panel per level beside the compacted level slabs).
// After the build the reservation is set from the actual ranks
// (resident_bytes + the factorization's transients), and a request that no
// longer fits the budget then fails alone.
static int levels_of(int64_t n, const kmh_build_opts& o) {
    const int fl = [&] { int l = -1; int64_t m = n; while (m) { m >>= 1; ++l; } return l; }();
    return o.level_kind == KMH_FIXED_LEVEL ? std::min(std::max(0, o.level), fl)
                                           : std::min(std::max(0, fl - o.level), fl);
}

uint64_t request_footprint(int64_t n, const kmh_build_opts& o, int nv) {
    const int L = levels_of(n, o);
    const int64_t nleaf = int64_t(1) << L;
    const int64_t m = (n + nleaf - 1) / nleaf;
    uint64_t r = (uint64_t)o.maxrank;
    uint64_t slabs = 0, xg = 0;
    if (o.method == KMH_ACA) {
        for (int l = 0; l < L; ++l) {
            const uint64_t h = (uint64_t)m << l;   // block size at level l (1-based l+1)
            uint64_t rl = std::min<uint64_t>(h, o.maxrank > 0 ? (uint64_t)o.maxrank : h);
            rl = std::min<uint64_t>(rl, 256);
            slabs += 8 * (uint64_t)n * rl * 2;                   // slab + the build's panel
            xg += 8 * (uint64_t)(nleaf >> (l + 1)) * 2 * (4 * rl * rl);
            r = std::max(r, rl);
        }
    } else {
        slabs = 8 * (uint64_t)n * r * (uint64_t)L;
        xg = 8 * (uint64_t)(nleaf - 1) * 2 * (4 * r * r);
    }
    uint64_t b = 16 * (uint64_t)n + 8 * (uint64_t)n * m + slabs + xg;
    if (o.method == KMH_NYSTROM) b += 8 * (uint64_t)n * (r * 4 + 3) + 24 * r * r;   // + carried y columns, W^-1 y
    b += 8 * (uint64_t)n * (2 + nv) * 2;                          // right-hand sides
    b += 8 * (uint64_t)n * 16;                                    // the leaf stage's diagonal-block inverses
    if (nv > 0) b += 8 * (uint64_t)n * (m + r + 2);               // derivative leaves and panel
    return b;
}

// Device bytes of a request from its built handle: the handle's buffers,
// the factorization's transients (cascade panels, per-block factors not
// allocated yet) and the right-hand sides.
static uint64_t built_footprint(const Hodlr& H, int nv) {
    uint64_t xg = 0, rmax = 0;
    for (const auto& lv : H.ranks)
        for (int r : lv) {
            xg += 8 * 2 * (4 * (uint64_t)r * r);
            rmax = std::max<uint64_t>(rmax, (uint64_t)r);
        }
    const uint64_t n = (uint64_t)H.n;
    uint64_t b = H.device_bytes() + xg + 8 * n * (2 + nv) * 2 + 8 * n * 16;   // + leaf-stage block inverses
    if (H.nystrom()) b += 8 * n * (uint64_t)H.Rtot + 8 * n * (2 * rmax + 3);   // level slabs, cascade panels, W^-1 y
    if (nv > 0) {
        int64_t m = 0;
        for (int i = 0; i < H.nleaves(); ++i) m = std::max<int64_t>(m, H.leaf_size(i));
        b += 8 * n * ((uint64_t)m + rmax + 2);
    }
    return b;
}

static void finish_batch(Ctx& c, std::vector<Req*>& rs) {
    std::vector<Hodlr*> hs;
    for (Req* q : rs) hs.push_back(q->H.get());
    std::vector<Status> st;
    factor_end_batch(c, hs, st);
    std::vector<Hodlr*> ok;
    std::vector<Req*> okr;
    for (size_t i = 0; i < rs.size(); ++i) {
        if (st[i].code != KMH_OK) rs[i]->st = st[i];
        else { ok.push_back(hs[i]); okr.push_back(rs[i]); }
    }
    // negloglik = 0.5 logdet + 0.5 y' K^{-1} y  (maxlikfunctions.jl:2-9),
    // with y' K^{-1} y = |W^{-1} y|^2 for the symmetric factor K = W W^T:
    // one application of W^{-1} per request, and W^{-T} on top only for the
    // requests whose gradient needs K^{-1} y
    if (!ok.empty()) {
        size_t tot = 0;
        for (Hodlr* h : ok) tot += 2 * (size_t)h->n;
        double* buf = static_cast<double*>(c.scratch((tot + ok.size()) * sizeof(double)));
        double* dots = buf + tot;
        std::vector<double*> T;
        std::vector<int> k;
        std::vector<Copy2D> cp;
        size_t off = 0;
        for (size_t i = 0; i < ok.size(); ++i) {
            const int64_t n = ok[i]->n;
            if (okr[i]->dy) {
                cp.push_back({okr[i]->dy, (int)n, buf + off, (int)n, (int)n, 1, 1.0});
            } else {
                KMH_HIP(hipMemcpyAsync(buf + off, okr[i]->y.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                       c.q()));
                cp.push_back({buf + off, (int)n, buf + off + n, (int)n, (int)n, 1, 1.0});
            }
            T.push_back(buf + off + n);
            k.push_back(1);
            off += 2 * n;
        }
        // y into the right-hand side (and, from device inputs, into buf too)
        std::vector<Copy2D> cp2;
        off = 0;
        for (size_t i = 0; i < ok.size(); ++i) {
            const int64_t n = ok[i]->n;
            if (okr[i]->dy) cp2.push_back({okr[i]->dy, (int)n, buf + off + n, (int)n, (int)n, 1, 1.0});
            off += 2 * n;
        }
        copy2d(c, cp);
        copy2d(c, cp2);
        // FP64 form: z = W^{-1} y, y' K^{-1} y = z' z.  Mixed form: x = K^{-1} y
        // refined against the FP64 operator, y' K^{-1} y = y' x.
        // (W^{-1} y of a factorization that carried y is already there)
        std::vector<Hodlr*> h64, h32;
        std::vector<double*> t64, t32;
        std::vector<DotJob> dj;
        std::vector<Copy2D> wcp;
        off = 0;
        for (size_t i = 0; i < ok.size(); ++i) {
            const int64_t n = ok[i]->n;
            if (ok[i]->p32()) {
                h32.push_back(ok[i]);
                t32.push_back(T[i]);
                dj.push_back({T[i], buf + off, n, dots + i});
            } else {
                if (ok[i]->has_winvy) {
                    wcp.push_back({ok[i]->winvy.as<double>(), (int)n, T[i], (int)n, (int)n, 1, 1.0});
                } else {
                    h64.push_back(ok[i]);
                    t64.push_back(T[i]);
                }
                dj.push_back({T[i], T[i], n, dots + i});
            }
            off += 2 * n;
        }
        copy2d(c, wcp);
        hodlr_factor_apply_batch(c, KMH_WINV, h64, t64, std::vector<int>(h64.size(), 1));
        if (!h32.empty()) hodlr_solve_batch(c, h32, t32, std::vector<int>(h32.size(), 1));
        dots_batch(c, dj);
        std::vector<Hodlr*> gh;
        std::vector<double*> gT;
        std::vector<int> gk;
        for (size_t i = 0; i < ok.size(); ++i)
            if (!okr[i]->dks.empty() && !ok[i]->p32()) {
                gh.push_back(ok[i]);
                gT.push_back(T[i]);
                gk.push_back(1);
            }
        if (!gh.empty()) hodlr_factor_apply_batch(c, KMH_WTINV, gh, gT, gk);   // K^{-1} y = W^{-T} z
        std::vector<double> hd(ok.size());
        KMH_HIP(hipMemcpyAsync(hd.data(), dots, ok.size() * sizeof(double), hipMemcpyDeviceToHost, c.q()));
        c.sync();
        for (size_t i = 0; i < ok.size(); ++i) {
            okr[i]->logdet = ok[i]->logdet;
            okr[i]->nll = 0.5 * ok[i]->logdet + 0.5 * hd[i];
        }
        // stochastic gradient (stochasticgradient.jl:30-45) of every request
        // that asks for one, together; y and K^-1 y are still in buf
        std::vector<GradJob> gj;
        off = 0;
        for (size_t i = 0; i < ok.size(); ++i) {
            Req* q = okr[i];
            const int64_t n = ok[i]->n;
            if (!q->dks.empty()) {
                q->grad.assign(q->dks.size(), 0.0);
                GradJob jb{ok[i], q->dks.data(), (int)q->dks.size(), nullptr, buf + off, buf + off + n,
                           q->dsaa ? nullptr : q->saa.data(), q->dsaa, q->nv, q->grad.data(), false};
                try {
                    check_grad_job(jb);
                    gj.push_back(jb);
                } catch (const Error& e) {
                    q->st.code = e.code;
                    q->st.msg = e.what();
                }
            }
            off += 2 * n;
        }
        if (!gj.empty()) {
            try {
                hodlr_stoch_gradient_batch(c, gj);
            } catch (const Error& e) {
                for (size_t i = 0; i < okr.size(); ++i)
                    if (!okr[i]->dks.empty() && okr[i]->st.code == KMH_OK) {
                        okr[i]->st.code = e.code;
                        okr[i]->st.msg = e.what();
                    }
            }
        }
        c.scratch_free(buf);
    }
    for (Req* q : rs) q->stage = ST_DONE;
}

// One stage of every in-flight request of a lane, on the lane's context.
// adjust: called (on the lane's thread) with the requests just built.
static void lane_step(Lane& ln, const std::function<void(std::vector<Req*>&)>& adjust) {
    Ctx& c = *ln.ctx;
    c.set_device();
    // group the lane's requests by stage (each advances one stage this step)
    std::vector<Req*> build, leaf, level, fin;
    for (Req* q : ln.inflight) {
        switch (q->stage) {
            case ST_BUILD: build.push_back(q); break;
            case ST_LEAF: leaf.push_back(q); break;
            case ST_LEVEL: level.push_back(q); break;
            case ST_FINISH: fin.push_back(q); break;
            default: break;
        }
    }
    if (!level.empty()) {
        std::vector<Hodlr*> hs;
        std::vector<int> lvs;
        for (Req* q : level) { hs.push_back(q->H.get()); lvs.push_back(q->lv); }
        factor_level_batch(c, hs, lvs);
        for (Req* q : level) {
            if (++q->lv >= q->H->L) q->stage = ST_FINISH;
        }
    }
    if (!leaf.empty()) {
        std::vector<Hodlr*> hs;
        for (Req* q : leaf) {
            // y rides through the factorization (W^{-1} y for the quadratic form)
            q->H->carry_dev = q->dy;
            q->H->carry_host = q->dy ? nullptr : q->y.data();
            hs.push_back(q->H.get());
        }
        factor_begin_batch(c, hs);
        for (Req* q : leaf) {
            q->lv = 0;
            q->stage = q->H->L > 0 ? ST_LEVEL : ST_FINISH;
        }
    }
    if (!build.empty()) {
        std::vector<BuildReq> br;
        for (Req* q : build) {
            br.push_back({q->dpts ? q->dpts : q->pts.data(), q->dpts != nullptr, q->n, 2, q->kern, q->opts});
            br.back().defer_checks = true;   // the leaf stage follows on the same stream
        }
        std::vector<Status> st;
        auto hs = hodlr_build_batch(c, br, st);
        for (size_t i = 0; i < build.size(); ++i) {
            if (hs[i]) {
                build[i]->H.reset(hs[i]);
                build[i]->stage = ST_LEAF;
            } else {
                build[i]->st = st[i];
                build[i]->stage = ST_DONE;
            }
        }
        if (adjust) adjust(build);
    }
    if (!fin.empty()) finish_batch(c, fin);
}

int server_step(Server& S) {
    S.ctx->set_device();
    // admission by footprint and slots; a new request joins the lane with
    // the fewest in flight
    while (!S.waiting.empty() && (int)S.inflight.size() < S.max_inflight) {
        Req* q = S.waiting.front();
        if (S.budget && S.reserved + q->footprint > S.budget) {
            if (S.inflight.empty()) {
                // can never fit: fail it instead of blocking the queue
                S.waiting.pop_front();
                q->st.code = KMH_E_BUDGET;
                q->st.msg = "request footprint exceeds the HBM budget";
                q->stage = ST_DONE;
                q->t_done = now_us();
                continue;
            }
            break;
        }
        S.waiting.pop_front();
        S.reserved += q->footprint;
        q->t_admit = now_us();
        q->stage = ST_BUILD;
        size_t best = 0;
        for (size_t l = 1; l < S.lanes.size(); ++l)
            if (S.lanes[l].inflight.size() < S.lanes[best].inflight.size()) best = l;
        q->lane = (int)best;
        S.lanes[best].inflight.push_back(q);
        S.inflight.push_back(q);
    }
    // the lanes route their products by the server context's GEMM table, so
    // that a request's result does not depend on its lane
    for (auto& ln : S.lanes)
        if (ln.ctx != S.ctx && ln.ctx->gemm_tune != S.ctx->gemm_tune) ln.ctx->gemm_tune = S.ctx->gemm_tune;
    // advance the lanes: the first active one on this thread, the others on
    // their own threads (each flushes its own launch queue)
    std::vector<Lane*> act;
    for (auto& ln : S.lanes)
        if (!ln.inflight.empty()) act.push_back(&ln);
    std::vector<std::exception_ptr> err(act.size());
    // after the build the reservation follows the actual ranks; a request
    // that no longer fits the budget fails alone (lanes run in parallel:
    // the server's counters are updated under a lock)
    std::mutex mu;
    auto adjust = [&](std::vector<Req*>& built) {
        std::lock_guard<std::mutex> lk(mu);
        for (Req* q : built) {
            if (!q->H) continue;
            const uint64_t actual = built_footprint(*q->H, q->nv);
            if (actual > q->footprint && S.budget && S.reserved - q->footprint + actual > S.budget) {
                q->st.code = KMH_E_BUDGET;
                q->st.msg = "request footprint exceeds the HBM budget after compression (" +
                            std::to_string(actual) + " bytes)";
                q->stage = ST_DONE;
                q->H.reset();
                continue;
            }
            S.reserved = S.reserved - q->footprint + actual;
            q->footprint = actual;
        }
    };
    std::vector<uint64_t> mark(act.size());
    for (size_t i = 0; i < act.size(); ++i) mark[i] = act[i]->ctx->alloc_seq;
    std::vector<std::thread> th;
    for (size_t i = 1; i < act.size(); ++i)
        th.emplace_back([&, i]() {
            try {
                lane_step(*act[i], adjust);
                flush_queued();
            } catch (...) {
                discard_queued();
                err[i] = std::current_exception();
            }
        });
    if (!act.empty()) {
        try {
            lane_step(*act[0], adjust);
            flush_queued();
        } catch (...) {
            discard_queued();
            err[0] = std::current_exception();
        }
    }
    for (auto& t : th) t.join();
    // A lane whose step failed (allocation failure, budget, HIP error) leaves
    // its requests part-way through a stage whose in-place updates are not
    // repeatable: every request of that lane fails with the error, and the
    // other lanes and later steps go on.
    for (size_t i = 0; i < act.size(); ++i) {
        if (!err[i]) continue;
        Status fs;
        try {
            std::rethrow_exception(err[i]);
        } catch (const Error& e) {
            fs.code = e.code;
            fs.msg = e.what();
        } catch (const std::bad_alloc&) {
            fs.code = KMH_E_OOM;
            fs.msg = "host allocation failed";
        } catch (const std::exception& e) {
            fs.code = KMH_E_DEVICE;
            fs.msg = e.what();
        }
        Ctx& lc = *act[i]->ctx;
        lc.drain();
        for (Req* q : act[i]->inflight) {
            q->H.reset();
            if (q->stage == ST_DONE && q->st.code != KMH_OK) continue;
            q->st = fs;
            q->stage = ST_DONE;
        }
        lc.scratch_sweep(mark[i]);   // stage-local blocks of the failed step
    }
    // retire finished requests
    int done = 0;
    for (auto& ln : S.lanes) {
        std::vector<Req*> keep;
        for (Req* q : ln.inflight) {
            if (q->stage == ST_DONE) {
                q->t_done = now_us();
                if (q->keep && q->st.code == KMH_OK && q->H) {
                    // resident factor: the reservation becomes the factor's bytes
                    const uint64_t res = q->H->device_bytes();
                    S.reserved = S.reserved - q->footprint + res;
                    q->footprint = res;
                    q->handle.reset(new kmh_hodlr(ln.kctx));
                    q->handle->borrowed = true;
                    q->handle->h = std::move(q->H);
                } else {
                    q->H.reset();
                    S.reserved -= q->footprint;
                    q->footprint = 0;
                }
                ++done;
            } else {
                keep.push_back(q);
            }
        }
        ln.inflight.swap(keep);
    }
    std::vector<Req*> keep;
    for (Req* q : S.inflight)
        if (q->stage != ST_DONE) keep.push_back(q);
    S.inflight.swap(keep);
    return done;
}

}  // namespace kmh

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct kmh_server {
    kmh::Server s;
    kmh_ctx* owner = nullptr;
    explicit kmh_server(kmh_ctx* c) : owner(c) { kmh_ctx_hold(c); }
    ~kmh_server() {
        s.all.clear();   // request handles back to their lanes' pools first
        s.inflight.clear();
        s.waiting.clear();
        drop_lanes();
        kmh_ctx_release(owner);
    }
    void drop_lanes() noexcept {
        for (auto& ln : s.lanes)
            if (ln.own) {
                ln.own->closing = true;   // torn down with its last resident factor
                if (ln.own->holders == 0) kmh::ctx_teardown(ln.own);
            }
        s.lanes.clear();
    }
    // lane 0 runs on the caller's context, the others on private ones
    void set_lanes(int n) {
        drop_lanes();
        s.lanes.resize(1);
        s.lanes[0].ctx = s.ctx;
        s.lanes[0].kctx = owner;
        for (int l = 1; l < n; ++l) {
            kmh::Lane ln;
            ln.own = kmh::ctx_create(s.ctx->device, 0);
            ln.kctx = ln.own;
            ln.ctx = &ln.own->c;
            s.lanes.push_back(ln);
        }
    }
};

#define SRV_TRY(...) KMH_API_TRY(__VA_ARGS__)

extern "C" {

int kmh_server_create(kmh_ctx* ctx, int32_t max_inflight, uint64_t hbm_budget_bytes, kmh_server** out) {
    SRV_TRY({
        KMH_CHECK(ctx && out && max_inflight >= 1, KMH_E_BAD_ARG, "invalid server arguments");
        std::unique_ptr<kmh_server> s(new kmh_server(ctx));
        s->s.ctx = &ctx->c;
        s->s.max_inflight = max_inflight;
        s->s.budget = hbm_budget_bytes;
        s->set_lanes(ctx->c.budget == 0 && max_inflight >= 2 ? 2 : 1);
        *out = s.release();
    })
}

int kmh_server_set_lanes(kmh_server* srv, int32_t lanes) {
    SRV_TRY({
        KMH_CHECK(srv, KMH_E_BAD_ARG, "null server");
        KMH_CHECK(lanes >= 1 && lanes <= KMH_SERVE_MAX_LANES, KMH_E_BAD_ARG, "lane count out of range");
        KMH_CHECK(srv->s.inflight.empty() && srv->s.waiting.empty(), KMH_E_STATE,
                  "lanes can change only while no request is queued or running");
        for (const auto& kv : srv->s.all)
            KMH_CHECK(!kv.second->handle, KMH_E_STATE, "lanes can change only while no factor is resident");
        srv->s.ctx->set_device();
        srv->set_lanes(lanes);
    })
}

int kmh_server_reserve(kmh_server* srv, int64_t n, const kmh_build_opts* opts, int32_t nv, int32_t requests) {
    SRV_TRY({
        KMH_CHECK(srv && opts && n > 0 && nv >= 0 && requests >= 0, KMH_E_BAD_ARG, "invalid reserve arguments");
        const int nl = (int)srv->s.lanes.size();
        const uint64_t per = kmh::request_footprint(n, *opts, nv);
        const uint64_t each = per * (uint64_t)((requests + nl - 1) / nl);
        if (each == 0) return KMH_OK;
        for (auto& ln : srv->s.lanes) kmh::ctx_reserve(*ln.ctx, (size_t)each);
    })
}

int kmh_server_destroy(kmh_server* srv) {
    SRV_TRY({
        if (srv) {
            for (auto& ln : srv->s.lanes) ln.ctx->sync();
            delete srv;
        }
    })
}

int kmh_server_submit(kmh_server* srv, const kmh_serve_req* rq, int64_t* id) {
    SRV_TRY({
        KMH_CHECK(srv && rq && id && rq->pts && rq->y, KMH_E_BAD_ARG, "null argument");
        KMH_CHECK(rq->n >= 2 && rq->d == 2, KMH_E_BAD_ARG, "invalid request size");
        KMH_CHECK(rq->ndk >= 0 && rq->ndk <= KMH_SERVE_MAX_PARAMS, KMH_E_BAD_ARG, "too many derivative kernels");
        KMH_CHECK(rq->ndk == 0 || (rq->saa && rq->nv >= 1), KMH_E_BAD_ARG, "gradient requests need SAA vectors");
        kmh::check_kernel(rq->kernel, false);
        std::unique_ptr<kmh::Req> q(new kmh::Req());
        q->id = srv->s.next_id++;
        q->n = rq->n;
        if (rq->on_device) {
            q->dpts = rq->pts;
            q->dy = rq->y;
        } else {
            q->pts.assign(rq->pts, rq->pts + 2 * rq->n);
            q->y.assign(rq->y, rq->y + rq->n);
        }
        q->kern = rq->kernel;
        q->opts = rq->opts;
        q->keep = rq->keep_factor != 0;
        q->nv = rq->ndk ? (int)rq->nv : 0;
        if (rq->ndk) {
            if (rq->on_device) q->dsaa = rq->saa;
            else q->saa.assign(rq->saa, rq->saa + (size_t)rq->n * rq->nv);
            for (int j = 0; j < rq->ndk; ++j) {
                kmh::check_kernel(rq->dks[j], true);
                q->dks.push_back(rq->dks[j]);
            }
        }
        q->footprint = kmh::request_footprint(rq->n, rq->opts, q->nv);
        q->t_submit = kmh::now_us();
        *id = q->id;
        srv->s.waiting.push_back(q.get());
        srv->s.all[q->id] = std::move(q);
    })
}

int kmh_server_step(kmh_server* srv, int32_t* completed) {
    SRV_TRY({
        KMH_CHECK(srv, KMH_E_BAD_ARG, "null server");
        const int d = kmh::server_step(srv->s);
        if (completed) *completed = d;
    })
}

int kmh_server_pending(kmh_server* srv, int32_t* waiting, int32_t* inflight, uint64_t* reserved_bytes) {
    SRV_TRY({
        KMH_CHECK(srv, KMH_E_BAD_ARG, "null server");
        if (waiting) *waiting = (int32_t)srv->s.waiting.size();
        if (inflight) *inflight = (int32_t)srv->s.inflight.size();
        if (reserved_bytes) *reserved_bytes = srv->s.reserved;
    })
}

int kmh_server_result(kmh_server* srv, int64_t id, kmh_serve_result* out) {
    SRV_TRY({
        KMH_CHECK(srv && out, KMH_E_BAD_ARG, "null argument");
        auto it = srv->s.all.find(id);
        KMH_CHECK(it != srv->s.all.end(), KMH_E_BAD_ARG, "unknown request id");
        kmh::Req& q = *it->second;
        KMH_CHECK(q.stage == kmh::ST_DONE, KMH_E_STATE, "request not finished");
        out->status = q.st.code;
        out->nll = q.nll;
        out->logdet = q.logdet;
        out->ngrad = (int32_t)q.grad.size();
        for (int j = 0; j < KMH_SERVE_MAX_PARAMS; ++j) out->grad[j] = j < (int)q.grad.size() ? q.grad[j] : 0.0;
        out->queue_us = q.t_admit > 0 ? q.t_admit - q.t_submit : 0.0;
        out->latency_us = q.t_done - q.t_submit;
        out->resident = q.handle ? 1 : 0;
        out->factor_bytes = q.handle ? (int64_t)q.footprint : 0;
        if (q.st.code != KMH_OK) kmh::set_last_error(q.st.msg);
        q.read = true;
        if (!q.handle) srv->s.all.erase(it);
    })
}

int kmh_server_factor(kmh_server* srv, int64_t id, kmh_hodlr** out) {
    SRV_TRY({
        KMH_CHECK(srv && out, KMH_E_BAD_ARG, "null argument");
        auto it = srv->s.all.find(id);
        KMH_CHECK(it != srv->s.all.end(), KMH_E_BAD_ARG, "unknown request id");
        kmh::Req& q = *it->second;
        KMH_CHECK(q.stage == kmh::ST_DONE, KMH_E_STATE, "request not finished");
        KMH_CHECK(q.handle, KMH_E_STATE, "the request keeps no resident factor");
        *out = q.handle.get();
    })
}

int kmh_server_release(kmh_server* srv, int64_t id) {
    SRV_TRY({
        KMH_CHECK(srv, KMH_E_BAD_ARG, "null server");
        auto it = srv->s.all.find(id);
        KMH_CHECK(it != srv->s.all.end(), KMH_E_BAD_ARG, "unknown request id");
        kmh::Req& q = *it->second;
        KMH_CHECK(q.stage == kmh::ST_DONE, KMH_E_STATE, "request not finished");
        if (q.handle) {
            q.handle->owner->c.set_device();
            srv->s.reserved -= q.footprint;
        }
        srv->s.all.erase(it);   // the factor's buffers go back to its lane's pool
    })
}

}  // extern "C"
