// Solves of a factorized HODLR matrix captured as hipGraphs.
//
// A follow-up solve on a resident factor (the same handle, the same device
// buffers b and x, the same number of right-hand sides) issues the same
// launches with the same arguments every time: the leaf triangular solves,
// then per tree level the Woodbury products of W^-1 and of W^-T.  The
// second solve with a given (b, x, nrhs) records them into a graph by
// stream capture, and later solves replay it with one hipGraphLaunch, which
// removes the host's planning of ~60 grouped launches and the gaps between
// them.  Descriptors and temporaries of the captured launches cannot live in
// the context's descriptor ring or scratch pool (both are reused by later
// work), so each graph owns two blocks of the pool: the descriptors, copied
// in once the capture ends, and the temporaries (a fixed scratch scope of the
// solve's bound).  The solve path clears no memory with memset nodes.
// Graphs are used for FP64 handles only: the mixed form's refinement reads
// residual norms back to the host between its steps.
#include <algorithm>

#include "hodlr.h"
#include "la.h"

namespace kmh {

static constexpr size_t GRAPH_DESC_BYTES = size_t(8) << 20;
static constexpr size_t GRAPHS_PER_HANDLE = 8;

void Hodlr::release_graphs() noexcept {
    if (graphs.empty()) return;
    if (ctx) ctx->drain();   // no replay in flight
    for (auto& g : graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        try {
            if (ctx) {
                ctx->scratch_free(g.desc);
                ctx->scratch_free(g.tmp);
            }
        } catch (...) {
        }
    }
    graphs.clear();
    graph_seen.clear();
}

// the eager solve K^-1 b -> x (copy, then the in-place solve)
static void solve_eager(Hodlr& H, const double* b, double* x, int64_t nrhs) {
    Ctx& c = *H.ctx;
    if (b != x) copy2d(c, {{b, (int)H.n, x, (int)H.n, (int)H.n, (int)nrhs, 1.0}});
    hodlr_solve(H, x, (int)nrhs);
}

bool solve_dev_graph(Hodlr& H, const double* b, double* x, int64_t nrhs) {
    Ctx& c = *H.ctx;
    if (!H.graphs_on || H.p32() || c.prof || c.fixed_base || c.cap_dev) return false;
    for (auto& g : H.graphs)
        if (g.b == b && g.x == x && g.nrhs == nrhs) {
            c.flush();
            KMH_HIP(hipGraphLaunch(g.exec, c.stream));
            ++H.graph_replays;
            return true;
        }
    bool seen = false;
    for (auto& g : H.graph_seen)
        if (g.b == b && g.x == x && g.nrhs == nrhs) seen = true;
    if (!seen || H.graphs.size() >= GRAPHS_PER_HANDLE) {
        if (!seen && H.graph_seen.size() < 4 * GRAPHS_PER_HANDLE) H.graph_seen.push_back({b, x, nrhs});
        return false;
    }
    // capture
    SolveGraph g{b, x, nrhs};
    const size_t tmp_bytes = std::max<size_t>(256, solve_scratch_bound(H, nrhs));
    g.desc = c.scratch(GRAPH_DESC_BYTES);
    g.tmp = c.scratch(tmp_bytes);
    std::vector<char> host;
    host.reserve(size_t(1) << 20);
    c.flush();
    hipGraph_t graph = nullptr;
    KMH_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    c.cap_dev = static_cast<char*>(g.desc);
    c.cap_size = GRAPH_DESC_BYTES;
    c.cap_host = &host;
    bool ok = true;
    std::string err;
    try {
        FixedScope fs(c, static_cast<char*>(g.tmp), tmp_bytes);
        solve_eager(H, b, x, nrhs);
        c.flush();
    } catch (const std::exception& e) {
        ok = false;
        err = e.what();
        c.pending.clear();
    }
    c.cap_dev = nullptr;
    c.cap_size = 0;
    c.cap_host = nullptr;
    const hipError_t ec = hipStreamEndCapture(c.stream, &graph);
    if (!ok || ec != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        c.scratch_free(g.tmp);
        c.scratch_free(g.desc);
        (void)hipGetLastError();
        throw Error(KMH_E_DEVICE, "solve graph capture failed: " + (ok ? std::string(hipGetErrorString(ec)) : err));
    }
    // the descriptors the captured kernels read, before the first replay
    KMH_HIP(hipMemcpyAsync(g.desc, host.data(), host.size(), hipMemcpyHostToDevice, c.stream));
    hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) {
        c.sync();
        c.scratch_free(g.tmp);
        c.scratch_free(g.desc);
        throw Error(KMH_E_DEVICE, std::string("solve graph instantiation failed: ") + hipGetErrorString(ei));
    }
    KMH_HIP(hipGraphLaunch(g.exec, c.stream));
    c.sync();   // host staging of the descriptors goes out of scope
    H.graphs.push_back(g);
    ++H.graph_captures;
    return true;
}

}  // namespace kmh

extern "C" int kmh_hodlr_solve_graphs(kmh_hodlr* h, int32_t enable, int64_t* captures, int64_t* replays) {
    KMH_API_TRY({
        KMH_CHECK(h && h->h, KMH_E_BAD_ARG, "null HODLR handle");
        kmh::Hodlr& H = *h->h;
        H.ctx->set_device();
        if (enable == 0 || enable == 1) {
            H.graphs_on = enable != 0;
            if (!H.graphs_on) H.release_graphs();
        }
        if (captures) *captures = H.graph_captures;
        if (replays) *replays = H.graph_replays;
    })
}
