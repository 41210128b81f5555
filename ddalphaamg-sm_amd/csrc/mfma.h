// Complex block products on the gfx950 matrix cores (16x16x4 MFMA, FP32 or FP64).
//
// A coarse-level hop  out_j += L v_j  (L a DOF x DOF complex link, v_j the neighbour's
// DOF-vector of system j, 16 systems per tile) is computed as one real product
//   Out (2DOF x 16) += A (2DOF x 2DOF) B (2DOF x 16)
// with real rows and columns interleaving re/im:  m = 2r + p, k = 2c + q,
//   A[m][k] = [[Lr, -Li], [Li, Lr]][p][q],   B[k][j] = (q ? Im : Re) v_j[c].
// Row tile t covers m in [16t, 16t+16) (complex rows 8t..8t+7), K-step s covers
// k in [4s, 4s+4) (complex columns 2s, 2s+1); MT = DOF/8 tiles and KS = DOF/2 steps.
//
// Lane layout of v_mfma_*_16x16x4 (one wave): lane l supplies A[l&15][l>>4] and
// B[l>>4][l&15].  The accumulator element e (0..3) of lane l is D[4(l>>4) + e][l&15]
// for FP32 and D[(l>>4) + 4e][l&15] for FP64 (both measured on gfx950).  The MFMA row
// index i is mapped to the real row m of the tile so that, in both precisions, lane l
// ends up holding for system j = l&15 the complex rows r = 8t + 2(l>>4) (elements
// 0, 1 = re, im) and r + 1 (elements 2, 3):
//   FP32: m - 16t = i                      (r = 8t + i/2,             p = i&1)
//   FP64: r = 8t + 2(i&3) + (i>>3),  p = (i>>2)&1.
#pragma once
#include "common.h"

namespace dd {

template <typename R>
struct Mfma;

template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

// Per-lane addressing of the A operand inside one column-major complex link matrix
// (viewed as 2*DOF*DOF reals): element (s, t) sits at aoff + (2s*DOF + 8t)*2 and is
// negated when aneg (the -Li quadrant).
struct LinkLane {
    int aoff;
    bool aneg;
    template <typename R, int DOF>
    __device__ __forceinline__ static LinkLane make(int lane) {
        const int i = lane & 15, g = lane >> 4, q = g & 1;
        const int rr = sizeof(R) == 4 ? (i >> 1) : 2 * (i & 3) + (i >> 3);   // complex row within the tile
        const int p = sizeof(R) == 4 ? (i & 1) : ((i >> 2) & 1);
        LinkLane ll;
        ll.aoff = ((g >> 1) * DOF + rr) * 2 + (p ^ q);
        ll.aneg = (p == 0) && (q == 1);
        return ll;
    }
};

template <int DOF>
__device__ __forceinline__ constexpr int link_step(int s, int t) {
    return (2 * s * DOF + 8 * t) * 2;
}

// acc[t] += hop contribution of one complex link matrix Lh (reals of type T: R, or
// binary16 widened to R on load) on a B operand given per K-step in bv[s] (already
// multiplied by the hop's boundary sign).  The link operands are loaded CH K-steps at a
// time (all MT tiles) ahead of their products.
template <typename R, int DOF, int CH = DOF / 2, typename T = R>
__device__ __forceinline__ void mfma_hop(typename Mfma<R>::acc_t (&acc)[DOF / 8], const T* __restrict__ Lh,
                                         const LinkLane& ll, const R (&bv)[DOF / 2]) {
    constexpr int MT = DOF / 8, KS = DOF / 2;
    static_assert(KS % CH == 0, "chunk must divide the K-steps");
#pragma unroll
    for (int s0 = 0; s0 < KS; s0 += CH) {
        R a[CH][MT];
#pragma unroll
        for (int s = 0; s < CH; s++)
#pragma unroll
            for (int t = 0; t < MT; t++) a[s][t] = (R)Lh[ll.aoff + link_step<DOF>(s0 + s, t)];
#pragma unroll
        for (int s = 0; s < CH; s++)
#pragma unroll
            for (int t = 0; t < MT; t++)
                acc[t] = Mfma<R>::mac(ll.aneg ? -a[s][t] : a[s][t], bv[s0 + s], acc[t]);
    }
}

// acc[k][t] += the same hop for NT tiles of 16 systems: each link operand is loaded once
// and used for the NT tiles' products.
template <typename R, int DOF, int NT>
__device__ __forceinline__ void mfma_hop_nt(typename Mfma<R>::acc_t (&acc)[NT][DOF / 8], const R* __restrict__ Lh,
                                            const LinkLane& ll, const R (&bv)[NT][DOF / 2]) {
    constexpr int MT = DOF / 8, KS = DOF / 2;
#pragma unroll
    for (int s = 0; s < KS; s++) {
        R a[MT];
#pragma unroll
        for (int t = 0; t < MT; t++) {
            a[t] = Lh[ll.aoff + link_step<DOF>(s, t)];
            a[t] = ll.aneg ? -a[t] : a[t];
        }
#pragma unroll
        for (int k = 0; k < NT; k++)
#pragma unroll
            for (int t = 0; t < MT; t++) acc[k][t] = Mfma<R>::mac(a[t], bv[k][s], acc[k][t]);
    }
}

// Store of a stencil result: out[i] = D v at row i (v = the input's own value there).
template <typename R>
struct StoreEpi {
    __device__ __forceinline__ void operator()(cplx<R>* __restrict__ out, size_t i, cplx<R> v, cplx<R> Dv) const {
        (void)v;
        out[i] = Dv;
    }
};

// One site of the generic level stencil for NT tiles of 16 systems (one wave):
//   out = (m0+2) v - A v - sum_mu (B_mu s+ v(x+mu) + C_mu s- v(x-mu))      (RES: b - that)
// for site n and systems 16*(tile0 + k) + (lane&15), k < NT, that are in `mask` and below
// nrhs; links [V][5][DOF][DOF] column-major, vectors [system][V][DOF] with stride ld.
// Every load is unconditional (indices of systems that do not take part are replaced by
// system 0) and masked afterwards, so no branch separates the loads of a hop.  Without
// RES the result goes through epi(out, row index, own input value, result).
template <typename R, int DOF, bool RES, int NT = 1, typename Epi = StoreEpi<R>>
__device__ __forceinline__ void coarse_site(const cplx<R>* __restrict__ L, const cplx<R>* __restrict__ in,
                                            cplx<R>* __restrict__ out, const cplx<R>* __restrict__ b, int Nx, int Nt,
                                            R m0, uint64_t mask, size_t ld, int nrhs, int n, int tile0, int lane,
                                            const Epi& epi = Epi()) {
    using acc_t = typename Mfma<R>::acc_t;
    constexpr int MT = DOF / 8, KS = DOF / 2;
    const int g = lane >> 4;
    int jj[NT];
    bool on[NT];
    const R* vin[NT];
#pragma unroll
    for (int k = 0; k < NT; k++) {
        jj[k] = (tile0 + k) * 16 + (lane & 15);
        on[k] = jj[k] < nrhs && ((mask >> jj[k]) & 1ull);
        vin[k] = reinterpret_cast<const R*>(in + (size_t)(on[k] ? jj[k] : 0) * ld);
    }
    const int x = n / Nt, t = n - x * Nt;
    // neighbour of hop h (0: self, 1: t+1, 2: x+1, 3: t-1, 4: x-1) and its boundary sign,
    // computed per hop (no private array indexed by the rolled hop loop)
    auto nbr = [&](int h) {
        return h == 0 ? n
             : h == 1 ? ((t == Nt - 1) ? n - (Nt - 1) : n + 1)
             : h == 2 ? ((x == Nx - 1) ? t : n + Nt)
             : h == 3 ? ((t == 0) ? n + (Nt - 1) : n - 1)
                      : ((x == 0) ? (Nx - 1) * Nt + t : n - Nt);
    };
    auto sgn = [&](int h) { return (h == 1 && t == Nt - 1) || (h == 3 && t == 0) ? R(-1) : R(1); };
    const LinkLane ll = LinkLane::make<R, DOF>(lane);
    const R* Ln = reinterpret_cast<const R*>(L + (size_t)n * 5 * DOF * DOF);
    acc_t acc[NT][MT];
#pragma unroll
    for (int k = 0; k < NT; k++)
#pragma unroll
        for (int tt = 0; tt < MT; tt++) acc[k][tt] = acc_t{0, 0, 0, 0};
    // one hop at a time: its link and vector loads are issued together, hops are not
    // interleaved (bounded register use inside larger kernels)
#pragma unroll 1
    for (int h = 0; h < 5; h++) {
        // B operand: part (g & 1) of column 2s + (g >> 1) -> real index 4s + g of the site
        R bv[NT][KS];
        const int nh = nbr(h);
        const R sh = sgn(h);
#pragma unroll
        for (int k = 0; k < NT; k++) {
            const R* vb = vin[k] + (size_t)nh * DOF * 2 + g;
#pragma unroll
            for (int s2 = 0; s2 < KS; s2++) {
                const R v = vb[4 * s2];
                bv[k][s2] = on[k] ? sh * v : R(0);
            }
        }
        mfma_hop_nt<R, DOF, NT>(acc, Ln + (size_t)h * DOF * DOF * 2, ll, bv);
    }
    const R mass = m0 + R(2);
#pragma unroll
    for (int k = 0; k < NT; k++) {
        if (!on[k]) continue;
        const size_t base = (size_t)jj[k] * ld + (size_t)n * DOF;
#pragma unroll
        for (int tt = 0; tt < MT; tt++) {
            const int r0 = 8 * tt + 2 * g;
            const cplx<R> v0 = in[base + r0], v1 = in[base + r0 + 1];
            cplx<R> o0 = mk<R>(mass * v0.x - acc[k][tt][0], mass * v0.y - acc[k][tt][1]);
            cplx<R> o1 = mk<R>(mass * v1.x - acc[k][tt][2], mass * v1.y - acc[k][tt][3]);
            if constexpr (RES) {
                out[base + r0] = b[base + r0] - o0;
                out[base + r0 + 1] = b[base + r0 + 1] - o1;
            } else {
                epi(out, base + r0, v0, o0);
                epi(out, base + r0 + 1, v1, o1);
            }
        }
    }
}

}  // namespace dd
