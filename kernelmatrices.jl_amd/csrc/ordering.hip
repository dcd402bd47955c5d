// KD-tree point ordering on the host (maxlikdata(sortmethod=:KDTree),
// HODLR/src/constructor.jl:171-181, which takes NearestNeighbors.KDTree(pts).data).
//
// Points of one KD cell become contiguous, which is what keeps the HODLR
// off-diagonal blocks low-rank.  Each cell of more than `leafsize` points is
// split at its midpoint index after a stable sort along the coordinate of
// largest spread of the cell's points (first such coordinate on ties).
// NearestNeighbors' own median selection is an implementation detail of a
// dependency that is not part of the reference, so orderings are close to,
// not equal to, the reference's.
#include <algorithm>
#include <numeric>
#include <vector>

#include "common.h"

namespace kmh {

void kd_order(const double* pts, int64_t n, int d, int leafsize, int64_t* perm) {
    std::iota(perm, perm + n, int64_t(0));
    std::vector<std::pair<int64_t, int64_t>> stack{{0, n}};
    std::vector<int64_t> tmp;
    while (!stack.empty()) {
        const auto [lo, hi] = stack.back();
        stack.pop_back();
        if (hi - lo <= leafsize) continue;
        int dim = 0;
        double best = -1.0;
        for (int k = 0; k < d; ++k) {
            double mn = pts[perm[lo] * d + k], mx = mn;
            for (int64_t i = lo + 1; i < hi; ++i) {
                const double v = pts[perm[i] * d + k];
                mn = std::min(mn, v);
                mx = std::max(mx, v);
            }
            if (mx - mn > best) {
                best = mx - mn;
                dim = k;
            }
        }
        std::stable_sort(perm + lo, perm + hi,
                         [&](int64_t a, int64_t b) { return pts[a * d + dim] < pts[b * d + dim]; });
        const int64_t mid = lo + (hi - lo) / 2;
        stack.push_back({lo, mid});
        stack.push_back({mid, hi});
    }
}

}  // namespace kmh

extern "C" int kmh_kd_order(const double* pts, int64_t n, int32_t d, int32_t leafsize, int64_t* perm) {
    KMH_API_TRY({
        KMH_CHECK(pts && perm, KMH_E_BAD_ARG, "null argument");
        KMH_CHECK(n >= 0 && d >= 1 && leafsize >= 1, KMH_E_BAD_ARG, "invalid sizes");
        kmh::kd_order(pts, n, d, leafsize, perm);
    })
}
