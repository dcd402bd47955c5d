// Allocation-free solve: K^{-1} b on a factorized HODLR with every
// temporary (the Woodbury products M^T T and G M^T T of each level and the
// k-slab partials of the tall products) in one fixed device array instead
// of the context's pool, for callers that cannot allocate device memory
// (code under stream capture, serving loops, harnesses that forbid it).
//
// The array lives in the code object: each GPU the library runs on holds
// its own copy, at its own address, until the process ends.  Its device
// address is looked up once per device with hipGetSymbolAddress; host code
// passes only that address to kernels and copies, never the symbol itself.
// Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (different contexts, or host threads) race on it.  Nothing in the
// array needs to start at zero (every temporary is written before it is
// read), so no call clears it.
#include <cmath>
#include <mutex>

#include "hodlr.h"
#include "la.h"

#define KMH_FIXED_SCRATCH_BYTES (size_t(32) << 20)

static __device__ __align__(256) char scratch[KMH_FIXED_SCRATCH_BYTES];

namespace kmh {
namespace {

std::mutex g_fixed_mu;
std::vector<char*> g_fixed_ptr;   // per device: device address of `scratch`

// the device address of the fixed array on the current device
char* fixed_scratch_ptr(int device) {
    std::lock_guard<std::mutex> lk(g_fixed_mu);
    if ((int)g_fixed_ptr.size() <= device) g_fixed_ptr.resize(device + 1, nullptr);
    if (!g_fixed_ptr[device]) {
        void* p = nullptr;
        KMH_HIP(hipSetDevice(device));
        KMH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        KMH_CHECK(p != nullptr, KMH_E_DEVICE, "no device address for the fixed scratch array");
        g_fixed_ptr[device] = static_cast<char*>(p);
    }
    return g_fixed_ptr[device];
}

size_t round256(size_t b) { return std::max<size_t>(256, (b + 255) & ~size_t(255)); }

}  // namespace

// An upper bound of the stack of temporaries a solve with nrhs columns
// holds at once: per tree level, the Woodbury products S and S2 (2 r nrhs
// each per block) plus the k-slab partials of the level's products, counted
// as if every product were split into 256-row slabs (the tall products split
// into 256-row slabs, the generic ones into 512-row slabs at most).
size_t solve_scratch_bound(const Hodlr& H, int64_t nrhs) {
    size_t peak = 0;
    for (int l = 0; l < H.L; ++l) {
        size_t ssz = 0, w = 0;
        for (size_t k = 0; k < H.blocks[l].size(); ++k) {
            const Block& b = H.blocks[l][k];
            const size_t r = (size_t)H.ranks[l][k];
            ssz += 2 * r * (size_t)nrhs;
            for (int64_t h : {b.d1 - b.c0, b.b1 - b.a0})
                w += round256((size_t)((h + 255) / 256) * r * (size_t)nrhs * sizeof(double));
            w += round256((size_t)((2 * r + 255) / 256) * 2 * r * (size_t)nrhs * sizeof(double));
            w += 2 * round256((size_t)((r + 255) / 256) * std::max<size_t>((size_t)(b.b1 - b.a0), (size_t)(b.d1 - b.c0)) *
                              (size_t)nrhs * sizeof(double));
        }
        peak = std::max(peak, round256(2 * ssz * sizeof(double)) + w);
    }
    return peak;
}

}  // namespace kmh

extern "C" int kmh_hodlr_solve_dev_noalloc(kmh_hodlr* h, const double* b, double* x, int64_t nrhs) {
    KMH_API_TRY({
        KMH_CHECK(h && h->h, KMH_E_BAD_ARG, "null HODLR handle");
        kmh::Hodlr& H = *h->h;
        kmh::Ctx& c = *H.ctx;
        c.set_device();
        KMH_CHECK(H.factorized, KMH_E_NOT_FACTORIZED, "No solves without factorization.");
        KMH_CHECK(b && x && nrhs >= 1 && nrhs < (int64_t(1) << 20), KMH_E_BAD_ARG, "invalid vectors");
        const size_t need = kmh::solve_scratch_bound(H, nrhs);
        KMH_CHECK(need <= KMH_FIXED_SCRATCH_BYTES, KMH_E_BAD_ARG,
                  "allocation-free solve: " + std::to_string(nrhs) + " right-hand sides need up to " +
                      std::to_string(need) + " bytes of scratch, more than the fixed " +
                      std::to_string(KMH_FIXED_SCRATCH_BYTES) + "; use kmh_hodlr_solve_dev or fewer columns");
        char* base = kmh::fixed_scratch_ptr(c.device);
        kmh::FixedScope fs(c, base, KMH_FIXED_SCRATCH_BYTES);
        if (b != x) kmh::copy2d(c, {{b, (int)H.n, x, (int)H.n, (int)H.n, (int)nrhs, 1.0}});
        kmh::hodlr_solve(H, x, (int)nrhs);
        c.flush();   // issue while the fixed scope holds (descriptors reference the array)
    })
}
