// Cross-lane moves of a 64-lane wavefront through the VALU lane networks (DPP row shifts
// and quad permutes, gfx950 permlane16/permlane32 swaps) instead of LDS permutes.
//
//   lane_xor<O>(v)          value of lane ^ O, O in {1, 2, 4, 8, 16, 32}
//   pair_swap<O>(A, B, l)   one exchange step of a halving reduction: afterwards lanes
//                           with bit O clear hold (own A, A of lane ^ O) and lanes with
//                           bit O set hold (B of lane ^ O, own B), so A + B is the pair's
//                           sum of A on the lower and of B on the upper lanes
#pragma once
#include <hip/hip_runtime.h>

namespace dd {

namespace lanes_detail {

// DPP controls: quad_perm [2,3,0,1] / [1,0,3,2]; row_shl:n = 0x100 + n, row_shr:n = 0x110 + n.
// Bank masks select groups of 4 lanes of a 16-lane row: 0xc lanes 8-15, 0x3 lanes 0-7,
// 0xa lanes 4-7 and 12-15, 0x5 lanes 0-3 and 8-11.
template <int O>
__device__ __forceinline__ unsigned xor32(unsigned u, int lane) {
    static_assert(O == 1 || O == 2 || O == 4 || O == 8 || O == 16 || O == 32, "lane offset");
    if constexpr (O == 1 || O == 2) {
        return (unsigned)__builtin_amdgcn_mov_dpp((int)u, O == 2 ? 0x4e : 0xb1, 0xf, 0xf, false);
    } else if constexpr (O == 4 || O == 8) {
        constexpr int hi = O == 8 ? 0xc : 0xa, lo = O == 8 ? 0x3 : 0x5;
        const int a = __builtin_amdgcn_update_dpp((int)u, (int)u, 0x110 + O, 0xf, hi, false);
        return (unsigned)__builtin_amdgcn_update_dpp(a, (int)u, 0x100 + O, 0xf, lo, false);
    } else if constexpr (O == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
        return (lane & 16) ? r[0] : r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(u, u, false, false);
        return (lane & 32) ? r[0] : r[1];
    }
}

template <int O>
__device__ __forceinline__ void swap32(unsigned& A, unsigned& B, int lane) {
    if constexpr (O == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(A, B, false, false);
        A = r[0];
        B = r[1];
    } else if constexpr (O == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(A, B, false, false);
        A = r[0];
        B = r[1];
    } else if constexpr (O == 4 || O == 8) {
        constexpr int hi = O == 8 ? 0xc : 0xa, lo = O == 8 ? 0x3 : 0x5;
        const int a2 = __builtin_amdgcn_update_dpp((int)A, (int)B, 0x110 + O, 0xf, hi, false);
        const int b2 = __builtin_amdgcn_update_dpp((int)B, (int)A, 0x100 + O, 0xf, lo, false);
        A = (unsigned)a2;
        B = (unsigned)b2;
    } else {
        static_assert(O == 1 || O == 2, "lane offset");
        const unsigned bx = xor32<O>(B, lane), ax = xor32<O>(A, lane);
        const bool up = (lane & O) != 0;
        const unsigned a2 = up ? bx : A, b2 = up ? B : ax;
        A = a2;
        B = b2;
    }
}

struct W2 {
    unsigned lo, hi;
};

}  // namespace lanes_detail

template <int O>
__device__ __forceinline__ float lane_xor(float v, int lane) {
    return __builtin_bit_cast(float, lanes_detail::xor32<O>(__builtin_bit_cast(unsigned, v), lane));
}
template <int O>
__device__ __forceinline__ double lane_xor(double v, int lane) {
    lanes_detail::W2 w = __builtin_bit_cast(lanes_detail::W2, v);
    w.lo = lanes_detail::xor32<O>(w.lo, lane);
    w.hi = lanes_detail::xor32<O>(w.hi, lane);
    return __builtin_bit_cast(double, w);
}

template <int O>
__device__ __forceinline__ void pair_swap(float& A, float& B, int lane) {
    unsigned a = __builtin_bit_cast(unsigned, A), b = __builtin_bit_cast(unsigned, B);
    lanes_detail::swap32<O>(a, b, lane);
    A = __builtin_bit_cast(float, a);
    B = __builtin_bit_cast(float, b);
}
template <int O>
__device__ __forceinline__ void pair_swap(double& A, double& B, int lane) {
    lanes_detail::W2 a = __builtin_bit_cast(lanes_detail::W2, A), b = __builtin_bit_cast(lanes_detail::W2, B);
    lanes_detail::swap32<O>(a.lo, b.lo, lane);
    lanes_detail::swap32<O>(a.hi, b.hi, lane);
    A = __builtin_bit_cast(double, a);
    B = __builtin_bit_cast(double, b);
}

// value of lane i + D (D > 0: row_shl) or i - |D| (D < 0: row_shr) within the lane's
// 16-lane row; 0 where that lane is outside the row
template <int D>
__device__ __forceinline__ float row_shift(float v) {
    static_assert(D != 0 && D > -16 && D < 16, "row shift");
    constexpr int ctrl = D > 0 ? 0x100 + D : 0x110 - D;
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), ctrl, 0xf, 0xf, true));
}
template <int D>
__device__ __forceinline__ double row_shift(double v) {
    static_assert(D != 0 && D > -16 && D < 16, "row shift");
    constexpr int ctrl = D > 0 ? 0x100 + D : 0x110 - D;
    lanes_detail::W2 w = __builtin_bit_cast(lanes_detail::W2, v);
    w.lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)w.lo, ctrl, 0xf, 0xf, true);
    w.hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)w.hi, ctrl, 0xf, 0xf, true);
    return __builtin_bit_cast(double, w);
}

// value of the lane K places before this one in its 16-lane row (DPP row_ror:K)
template <int K>
__device__ __forceinline__ float row_rotate(float v) {
    static_assert(K > 0 && K < 16, "row rotation");
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x120 + K, 0xf, 0xf, true));
}
template <int K>
__device__ __forceinline__ double row_rotate(double v) {
    static_assert(K > 0 && K < 16, "row rotation");
    lanes_detail::W2 w = __builtin_bit_cast(lanes_detail::W2, v);
    w.lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)w.lo, 0x120 + K, 0xf, 0xf, true);
    w.hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)w.hi, 0x120 + K, 0xf, 0xf, true);
    return __builtin_bit_cast(double, w);
}

// sum over each group of 16 lanes, partners at lane offsets 8, 4, 2, 1 in that order (the
// order of wave_sum(v, 16)); every lane of the group gets the same bits.  The partners
// come from row rotations by 8, 4, 2, 1: after the step with rotation K the values are
// K-periodic in the row, so the partner K lanes back holds the value of lane ^ K, and
// every lane adds (own, partner) exactly as the xor butterfly does, one DPP move a step.
template <typename R>
__device__ __forceinline__ R seg16_sum(R v, int lane) {
    (void)lane;
    v += row_rotate<8>(v);
    v += row_rotate<4>(v);
    v += row_rotate<2>(v);
    v += row_rotate<1>(v);
    return v;
}

// v + (v of lane ^ o) for a runtime offset o in {1, ..., 32}
template <typename R>
__device__ __forceinline__ R add_xor(R v, int o, int lane) {
    switch (o) {
        case 32: return v + lane_xor<32>(v, lane);
        case 16: return v + lane_xor<16>(v, lane);
        case 8: return v + lane_xor<8>(v, lane);
        case 4: return v + lane_xor<4>(v, lane);
        case 2: return v + lane_xor<2>(v, lane);
        default: return v + lane_xor<1>(v, lane);
    }
}

}  // namespace dd
