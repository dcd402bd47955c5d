// Sums across the lanes of one 64-wide wavefront on gfx950, in a fixed
// order (the same bits on every run): lane-pair additions through DPP and
// the gfx950 lane swaps, and column sums of per-lane values staged in LDS.
#pragma once
#include <hip/hip_runtime.h>

namespace kmh {

// s + (s of lane l ^ M) for M = 8, 16, 32 without the LDS crossbar: a DPP
// rotation by 8 within each 16-lane row (equal to xor 8 there), and the
// gfx950 lane swaps v_permlane16_swap / v_permlane32_swap, which exchange
// odd with even 16-lane rows and the upper with the lower 32 lanes of two
// registers; swapping a value with itself and adding the two results gives
// every lane s_l + s_(l^M) (the same two terms in one addition, so the same
// bits as a shuffle-and-add).
template <int M>
__device__ __forceinline__ double xor_add(double s) {
    const unsigned lo = (unsigned)__double2loint(s), hi = (unsigned)__double2hiint(s);
    if constexpr (M == 32) {
        const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        return __hiloint2double((int)b[0], (int)a[0]) + __hiloint2double((int)b[1], (int)a[1]);
    } else if constexpr (M == 16) {
        const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        return __hiloint2double((int)b[0], (int)a[0]) + __hiloint2double((int)b[1], (int)a[1]);
    } else {
        static_assert(M == 8, "xor_add: M = 8, 16 or 32");
        const unsigned l2 = __builtin_amdgcn_update_dpp(0u, lo, 0x128, 0xF, 0xF, false);   // row_ror:8
        const unsigned h2 = __builtin_amdgcn_update_dpp(0u, hi, 0x128, 0xF, 0xF, false);
        return s + __hiloint2double((int)h2, (int)l2);
    }
}

// Sums over the wave of per-lane values p[k], k < H: lane l ends with the
// sum of p[l % H] over all 64 lanes.  The lanes write their values to LDS,
// lane l sums a segment of column l % H, and the segments meet by
// xor-shuffles.  LDS operations of one wave complete in issue order, so no
// barrier is needed.  Fixed summation order.  LDS: 64 x (H + 1) doubles.
template <int H>
__device__ __forceinline__ double lds_colsum_read(double* pw, int lane) {
    static_assert(H >= 8 && H <= 32 && 64 % H == 0, "column sums of 8, 16 or 32 columns");
    constexpr int SEGS = 64 / H, RPS = 64 / SEGS;
    __builtin_amdgcn_wave_barrier();
    const int k = lane % H, seg = lane / H;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < RPS; i += 2) {
        s0 += pw[(seg * RPS + i) * (H + 1) + k];
        s1 += pw[(seg * RPS + i + 1) * (H + 1) + k];
    }
    double s = s0 + s1;
    if constexpr (H <= 8) s = xor_add<8>(s);
    if constexpr (H <= 16) s = xor_add<16>(s);
    s = xor_add<32>(s);
    __builtin_amdgcn_wave_barrier();
    return s;
}
template <int H>
__device__ __forceinline__ double lds_colsum(const double (&p)[H], double* pw, int lane) {
#pragma unroll
    for (int k = 0; k < H; ++k) pw[lane * (H + 1) + k] = p[k];
    return lds_colsum_read<H>(pw, lane);
}

}  // namespace kmh
