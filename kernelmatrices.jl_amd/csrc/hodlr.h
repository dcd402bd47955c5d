// HODLR matrices held on one device: tree, leaves, off-diagonal panels,
// factor.  Every stage runs over a list of handles, so several independent
// problems (of different sizes, kernels and ranks) share each grouped launch.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "common.h"

namespace kmh {

// Off-diagonal block of one level: rows a0..b1 (second child) x cols c0..d1
// (first child), 0-based half-open (indices.jl:40-51).
struct Block {
    int64_t a0, b1, c0, d1;
};

// A solve captured as a hipGraph (graph.hip): the launches of K^-1 b -> x
// for one (b, x, nrhs), with their descriptors and temporaries in blocks of
// the context's pool owned by the graph.
struct SolveGraph {
    const double* b = nullptr;
    double* x = nullptr;
    int64_t nrhs = 0;
    hipGraphExec_t exec = nullptr;
    void* desc = nullptr;
    void* tmp = nullptr;
};

struct Hodlr {
    Ctx* ctx = nullptr;
    int64_t n = 0;
    int L = 0;                                   // number of off-diagonal levels
    std::vector<int64_t> leaf_off;               // nleaves + 1 boundaries
    std::vector<std::vector<Block>> blocks;      // [level][block]
    std::vector<std::vector<int>> ranks;         // [level][block]
    std::vector<int> rpad;                       // panel width per level
    std::vector<int64_t> coff;                   // first column of each level in P (L+1 entries)
    int64_t Rtot = 0;
    kmh_kernel kern{};
    kmh_build_opts opts{};

    DevBuf pts;                                  // x[0..n), y[0..n)
    DevBuf info;                                 // device failure flags: [0] Cholesky of the
                                                 // factorization, [1] landmark Cholesky of a
                                                 // build that deferred its check
    bool landmark_check = false;                 // info[1] not read yet (read with the
                                                 // factorization's flags, factor_end_batch)
    DevBuf leaves;                               // leaf i at leaf_mat_off[i], ld = size
    bool leaves_ready = false;                   // leaves hold K's leaves (assembled on first
                                                 // use, ensure_leaves; a factorization of
                                                 // leaves of <= 256 rows evaluates them itself)
    std::vector<int64_t> leaf_mat_off;
    DevBuf P;                                    // n x Rtot, ld n: ACA U/V, then Q per level
                                                 // (FP32 elements in the mixed form)
    DevBuf UV;                                   // mixed-form ACA: U/V kept beside the Q slabs (FP32)

    // Nystrom state (kept for the derivative HODLR)
    int nys_r = 0;
    std::vector<int64_t> nyind;
    DevBuf lm;                                   // landmark points, x[0..r), y[0..r)
    DevBuf kpp;                                  // chol(K(P,P)) lower, r x r
    DevBuf kinv;                                 // K(P,P)^{-1}, r x r
    DevBuf kxp;                                  // K(x, P), n x r, ld n
    DevBuf ynys;                                 // K(x, P) K(P,P)^{-1}, n x r, ld n

    // factor
    bool factorized = false;
    bool factor_started = false;                 // between factor_begin_batch and factor_end_batch
    std::vector<std::vector<int64_t>> coff_blk;  // offset of block (2r x 2r) in X / G
    DevBuf X;                                    // chol(C) - I per block
    DevBuf G;                                    // I - chol(C)^{-1} per block
    double logdet = 0.0;
    // iterative refinement of the last solve (mixed form)
    int ir_iters = 0;
    double ir_relres = 0.0;

    // W^{-1} y carried through the factorization (FP64 Nystrom handles with
    // levels): with carry_dev (device) or carry_host (host) set before
    // factor_begin_batch, y rides in the cascade panels as one more column
    // beside the K(x,P) half and one beside the other half (the level
    // products that need Q^T y read it there), the leaf solve and every
    // W_l^{-1} act on it, and winvy holds W^{-1} y after the last level
    const double* carry_dev = nullptr;
    const double* carry_host = nullptr;
    bool carry = false;                          // this factorization carries y
    bool has_winvy = false;
    DevBuf winvy;
    // cascade panel columns: [y | K(x,P) half | other half | y] with a
    // carry, else the two halves
    int ab_a() const { return carry ? 1 : 0; }
    int ab_b() const { return ab_a() + nys_r; }
    int ab_w() const { return 2 * nys_r + (carry ? 2 : 0); }

    // transient state of a factorization in progress (scratch blocks of
    // ctx's pool, released by factor_end_batch, by a new factorization or
    // with the handle)
    double* AB = nullptr;                        // Nystrom cascade panels (n x ab_w())
    double* parts = nullptr;                     // log-diagonal partial sums
    int nparts = 0;

    // captured solves (graph.hip)
    bool graphs_on = true;
    int64_t graph_captures = 0, graph_replays = 0;
    std::vector<SolveGraph> graphs;
    std::vector<SolveGraph> graph_seen;   // solved once eagerly, not captured yet
    void release_graphs() noexcept;

    Hodlr() = default;
    Hodlr(const Hodlr&) = delete;
    Hodlr& operator=(const Hodlr&) = delete;
    ~Hodlr() {
        release_graphs();
        release_transient();
    }
    void release_transient() noexcept {
        if (!ctx) return;
        for (double** p : {&AB, &parts}) {
            if (*p) {
                try {
                    ctx->scratch_free(*p);
                } catch (...) {
                }
            }
            *p = nullptr;
        }
        nparts = 0;
    }

    bool nystrom() const { return opts.method == KMH_NYSTROM; }
    // mixed forms: the Q slabs (and ACA panels) stored in FP32; the Nystrom
    // cascade panels too in KMH_MIXED
    bool p32() const { return opts.precision != KMH_FP64; }
    bool ab32() const { return opts.precision == KMH_MIXED; }
    // element i of P (FP64 or FP32 storage) as a descriptor pointer
    double* Pat(size_t i) const { return p32() ? reinterpret_cast<double*>(P.as<float>() + i) : P.as<double>() + i; }
    // U (rows a0..b1) and V (rows c0..d1) of block k at level l, leading
    // dimension n.  Nystrom blocks all share the two panels K(x,P) and
    // K(x,P)K(P,P)^{-1} (src/factorizations.jl:86-94; FP64 in both forms);
    // ACA blocks live in P before factorization (in UV after it, mixed form
    // only), FP32 when uv32().
    bool uv32() const { return p32() && !nystrom(); }
    const double* uvbase(size_t i) const {
        const DevBuf& s = (UV.p && (factorized || factor_started)) ? UV : P;
        return p32() ? reinterpret_cast<const double*>(s.as<float>() + i) : s.as<double>() + i;
    }
    const double* Uptr(int l, int k) const {
        const Block& b = blocks[l][k];
        return nystrom() ? kxp.as<double>() + b.a0 : uvbase((size_t)coff[l] * n + b.a0);
    }
    const double* Vptr(int l, int k) const {
        const Block& b = blocks[l][k];
        return nystrom() ? ynys.as<double>() + b.c0 : uvbase((size_t)coff[l] * n + b.c0);
    }
    // U/V available for the operator's matvec (before factorization, or
    // kept for the refinement of a mixed-form solve)
    bool has_uv() const { return !factorized || nystrom() || UV.p; }
    int64_t leaf_size(int i) const { return leaf_off[i + 1] - leaf_off[i]; }
    int nleaves() const { return (int)leaf_off.size() - 1; }
    int nblocks() const {
        int s = 0;
        for (const auto& b : blocks) s += (int)b.size();
        return s;
    }
    double* leaf_ptr(int i) const { return leaves.as<double>() + leaf_mat_off[i]; }
    size_t device_bytes() const {
        return pts.bytes + info.bytes + leaves.bytes + P.bytes + UV.bytes + lm.bytes + kpp.bytes + kinv.bytes +
               kxp.bytes + ynys.bytes + X.bytes + G.bytes + winvy.bytes;
    }
};

// The temporaries of one solve of H with nrhs columns (an upper bound of the
// scratch it holds at once), and a scope that puts the context's scratch on
// a fixed block for one call (noalloc.hip).
size_t solve_scratch_bound(const Hodlr& H, int64_t nrhs);
struct FixedScope {
    Ctx& c;
    FixedScope(Ctx& cx, char* base, size_t size) : c(cx) {
        c.fixed_base = base;
        c.fixed_size = size;
        c.fixed_stack.clear();
    }
    ~FixedScope() {
        c.fixed_base = nullptr;
        c.fixed_size = 0;
        c.fixed_stack.clear();
    }
};

// K^-1 b -> x on the device through a captured graph: replays the graph of
// (b, x, nrhs) if there is one, captures it on the second solve with the same
// buffers; false when the solve is not graph-eligible (mixed form, profiling,
// graphs off) or this is the first solve with these buffers.
bool solve_dev_graph(Hodlr& H, const double* b, double* x, int64_t nrhs);

struct BuildReq {
    const double* pts;
    bool on_device;
    int64_t n;
    int d;
    kmh_kernel kern;
    kmh_build_opts opts;
    // Leave the landmark Cholesky's failure flag to the factorization's
    // final check instead of synchronizing on it here (the server's build
    // stage, whose requests always go on to be factorized): with every
    // request's points on the device the build then returns without waiting
    // for the device.
    bool defer_checks = false;
};

// Per-request outcome of a batched stage.
struct Status {
    int code = KMH_OK;
    std::string msg;
};

// Build every request (one grouped launch per step across requests).
// Requests that fail get a status and a null handle; the others proceed.
std::vector<Hodlr*> hodlr_build_batch(Ctx& c, const std::vector<BuildReq>& reqs, std::vector<Status>& st);
// Resumable factorization stages (continuous batching): begin (leaf stage),
// one level per handle (levels may differ), end (logdet, flags).
void factor_begin_batch(Ctx& c, const std::vector<Hodlr*>& hs);
void factor_level_batch(Ctx& c, const std::vector<Hodlr*>& hs, const std::vector<int>& lvs);
void factor_end_batch(Ctx& c, const std::vector<Hodlr*>& hs, std::vector<Status>& st);
// symmetricfactorize! of every handle, level by level across handles.
void hodlr_factorize_batch(Ctx& c, const std::vector<Hodlr*>& hs, std::vector<Status>& st);
// K^{-1} T in place for each handle's device block T (n x nrhs, ld n)
void hodlr_solve_batch(Ctx& c, const std::vector<Hodlr*>& hs, const std::vector<double*>& T,
                       const std::vector<int>& nrhs);

Hodlr* hodlr_build(Ctx& c, const double* pts, bool pts_on_device, int64_t n, int d, const kmh_kernel& k,
                   const kmh_build_opts& o);
void hodlr_factorize(Hodlr& H);
// y = K x (unfactorized: leaves + U V^T blocks; factorized: W W^T x)
void hodlr_matvec(Hodlr& H, const double* x, double* y, int nrhs);
// in-place on device data T (n x nrhs, ld n)
void hodlr_factor_apply(Hodlr& H, int op, double* T, int nrhs);
void hodlr_solve(Hodlr& H, double* T, int nrhs);
void hodlr_full_host(Hodlr& H, double* out);
void hodlr_block_uv_host(Hodlr& H, int level, int block, double* U, double* V);
void hodlr_stoch_gradient(Hodlr& H, const kmh_kernel* dks, int p, const double* y, const double* saa, int64_t nv,
                          double* grad, bool profile = false, bool on_device = false);
// W^-1 (KMH_WINV) or W^-T (KMH_WTINV) of each handle's device block T in place.
void hodlr_factor_apply_batch(Ctx& c, int op, const std::vector<Hodlr*>& hs, const std::vector<double*>& T,
                              const std::vector<int>& nrhs);
// One handle's stochastic gradient in a batch: y and the SAA vectors
// (n x nv) on the host or the device, K^-1 y optionally already computed on
// the device (it is then not solved again).
struct GradJob {
    Hodlr* H;
    const kmh_kernel* dks;
    int p;
    const double* y_host;
    const double* y_dev;
    const double* kinv_y_dev;
    const double* saa_host;
    const double* saa_dev;
    int64_t nv;
    double* grad;
    bool profile;
};
void check_grad_job(const GradJob& jb);   // the checks of the batch, for one job
void hodlr_stoch_gradient_batch(Ctx& c, const std::vector<GradJob>& jobs);
void check_kernel(const kmh_kernel& k, bool allow_deriv);

}  // namespace kmh

// Opaque C-ABI HODLR handle: the matrix and the context it lives on (held
// until the handle is destroyed).  A handle lent out by a server for a
// resident factor (kmh_server_factor) is owned by the server.
struct kmh_hodlr {
    std::unique_ptr<kmh::Hodlr> h;
    kmh_ctx* owner = nullptr;
    bool borrowed = false;
    explicit kmh_hodlr(kmh_ctx* c) : owner(c) { kmh_ctx_hold(c); }
    ~kmh_hodlr() {
        h.reset();   // buffers back to the context's pool first
        kmh_ctx_release(owner);
    }
};
