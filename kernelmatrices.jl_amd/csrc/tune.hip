// GEMM autotuner: for each shape class of the HODLR path's products, time
// every kernel class that can take it (generic 64x64-tile MFMA kernel, tall
// TN kernel, row-per-thread NN kernel, wide NN kernel) on a batch of
// problems of that shape and keep the fastest.  The table maps a shape class
// (transposes, ceil(log2) of m, n, k) to a kernel class; gemm() consults it
// before its built-in rules.  The choice depends on the problem's shape only,
// so a problem's result stays the same alone and in any batch under one
// table.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "la.h"

namespace kmh {
uint32_t gemm_class_key(bool ta, bool tb, int m, int n, int k);
bool gemm_class_eligible(int cls, bool ta, bool tb, const GemmProb& p);
}

namespace {
const char* const kClassName[] = {"rule", "generic", "tall", "rows", "wide"};
}

extern "C" int kmh_gemm_autotune(kmh_ctx* ctx, int32_t count, const int32_t* shapes, int32_t reps, const char* path) {
    KMH_API_TRY({
        KMH_CHECK(ctx && shapes && count >= 0 && reps >= 1, KMH_E_BAD_ARG, "invalid autotune arguments");
        kmh::Ctx& c = ctx->c;
        c.set_device();
        KMH_CHECK(!c.prof, KMH_E_STATE, "autotune inside a profiling window");
        std::ostringstream out;
        out << "# kmhodlr GEMM kernel-class table (gfx950): gemm ta tb ceil_log2(m) ceil_log2(n) ceil_log2(k) class\n";
        hipEvent_t e0, e1;
        KMH_HIP(hipEventCreate(&e0));
        KMH_HIP(hipEventCreate(&e1));
        struct EvGuard {
            hipEvent_t a, b;
            ~EvGuard() {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } evg{e0, e1};
        for (int s = 0; s < count; ++s) {
            const int* sh = shapes + 6 * s;
            const bool ta = sh[0] != 0, tb = sh[1] != 0;
            const int m = sh[2], n = sh[3], k = sh[4], P = sh[5];
            KMH_CHECK(m > 0 && n > 0 && k > 0 && P > 0, KMH_E_BAD_ARG, "invalid autotune shape");
            const size_t per = (size_t)m * k + (size_t)k * n + (size_t)m * n;
            double* buf = static_cast<double*>(c.scratch(per * P * sizeof(double)));
            KMH_HIP(hipMemsetAsync(buf, 0, per * P * sizeof(double), c.q()));
            std::vector<kmh::GemmProb> probs;
            for (int q = 0; q < P; ++q) {
                double* A = buf + per * q;
                double* B = A + (size_t)m * k;
                double* C = B + (size_t)k * n;
                probs.push_back({m, n, k, A, ta ? k : m, B, tb ? n : k, C, m, 1.0, 1.0, 0});
            }
            double best = 1e300;
            int best_cls = 0;
            std::ostringstream cand;
            for (int cls = 1; cls <= 4; ++cls) {
                if (!kmh::gemm_class_eligible(cls, ta, tb, probs[0])) continue;
                c.gemm_force = cls;
                try {
                    for (int w = 0; w < 2; ++w) kmh::gemm(c, ta, tb, probs);
                    std::vector<float> t;
                    for (int r = 0; r < reps; ++r) {
                        c.flush();
                        KMH_HIP(hipEventRecord(e0, c.stream));
                        kmh::gemm(c, ta, tb, probs);
                        c.flush();
                        KMH_HIP(hipEventRecord(e1, c.stream));
                        KMH_HIP(hipEventSynchronize(e1));
                        float ms = 0.f;
                        KMH_HIP(hipEventElapsedTime(&ms, e0, e1));
                        t.push_back(ms);
                    }
                    std::sort(t.begin(), t.end());
                    const double med = t[t.size() / 2] * 1e3;
                    cand << " " << kClassName[cls] << "=" << med << "us";
                    if (med < best) {
                        best = med;
                        best_cls = cls;
                    }
                } catch (...) {
                    c.gemm_force = 0;
                    c.scratch_free(buf);
                    throw;
                }
                c.gemm_force = 0;
            }
            c.sync();
            c.scratch_free(buf);
            if (best_cls) {
                const uint32_t key = kmh::gemm_class_key(ta, tb, m, n, k);
                c.gemm_tune[key] = (uint8_t)best_cls;
                char line[160];
                std::snprintf(line, sizeof line, "gemm %d %d %u %u %u %s", (int)ta, (int)tb, (key >> 20) & 1023u,
                              (key >> 10) & 1023u, key & 1023u, kClassName[best_cls]);
                out << line << "    # m=" << m << " n=" << n << " k=" << k << " problems=" << P << ":" << cand.str()
                    << "\n";
            }
        }
        if (path) {
            std::ofstream f(path);
            KMH_CHECK(f.good(), KMH_E_BAD_ARG, std::string("cannot write ") + path);
            f << out.str();
        }
    })
}

extern "C" int kmh_gemm_tuning_load(kmh_ctx* ctx, const char* path) {
    KMH_API_TRY({
        KMH_CHECK(ctx, KMH_E_BAD_ARG, "null context");
        kmh::Ctx& c = ctx->c;
        c.gemm_tune.clear();
        if (!path) return KMH_OK;
        std::ifstream f(path);
        KMH_CHECK(f.good(), KMH_E_BAD_ARG, std::string("cannot read ") + path);
        std::string line;
        while (std::getline(f, line)) {
            std::istringstream is(line);
            std::string tag, cls;
            int ta, tb;
            unsigned lm, ln, lk;
            if (!(is >> tag) || tag != "gemm") continue;
            KMH_CHECK(bool(is >> ta >> tb >> lm >> ln >> lk >> cls) && lm < 1024 && ln < 1024 && lk < 1024,
                      KMH_E_BAD_ARG, "malformed tuning line: " + line);
            int id = 0;
            for (int q = 1; q <= 4; ++q)
                if (cls == kClassName[q]) id = q;
            KMH_CHECK(id, KMH_E_BAD_ARG, "unknown kernel class in tuning line: " + line);
            const uint32_t key = (uint32_t(ta != 0) << 31) | (uint32_t(tb != 0) << 30) | (lm << 20) | (ln << 10) | lk;
            c.gemm_tune[key] = (uint8_t)id;
        }
    })
}

extern "C" int kmh_gemm_tuning_size(kmh_ctx* ctx, int32_t* entries) {
    KMH_API_TRY({
        KMH_CHECK(ctx && entries, KMH_E_BAD_ARG, "null argument");
        *entries = (int32_t)ctx->c.gemm_tune.size();
    })
}
