// Shared device/host definitions for the DD-alphaAMG HIP library (gfx950).
//
// Layouts (no halos; the reference's halo-padded arrays are packed/unpacked at the
// boundary):
//   site index        n = x * Nt + t          (t fastest, as the reference's interior)
//   level-0 vector    cplx[nrhs][V][2]         (spin fastest)
//   level-l vector    cplx[nrhs][V][dof]       (dof = colour*2 + spin)
//   gauge field       cplx[V][2]               (U_mu(x), mu = 0 is t, mu = 1 is x)
//   coarse links      cplx[V][5][dof][dof]     matrix k in {A, B_0, B_1, C_0, C_1},
//                                              stored column-major: [col][row]
//   test vectors      cplx[V][Nv][dof]         (prolongator of the level)
// Batches of right-hand sides are addressed by a 64-bit mask passed by value: bit r
// set means system r takes part in the launch (grid.y runs over systems).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

namespace dd {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
    R x, y;
};

template <typename R>
__host__ __device__ __forceinline__ cplx<R> mk(R a, R b) {
    cplx<R> c;
    c.x = a;
    c.y = b;
    return c;
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator-(cplx<R> a) { return mk(-a.x, -a.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> operator*(R s, cplx<R> a) { return mk(s * a.x, s * a.y); }
template <typename R>
__host__ __device__ __forceinline__ cplx<R> conj(cplx<R> a) { return mk(a.x, -a.y); }
// a * conj(b)
template <typename R>
__host__ __device__ __forceinline__ cplx<R> mulc(cplx<R> a, cplx<R> b) {
    return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) * b
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
    return mk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// i * a
template <typename R>
__host__ __device__ __forceinline__ cplx<R> times_i(cplx<R> a) { return mk(-a.y, a.x); }
template <typename R>
__host__ __device__ __forceinline__ R norm2(cplx<R> a) { return a.x * a.x + a.y * a.y; }
// complex division a / b (b may be complex); plain formula, as the reference's
// std::complex arithmetic does for the Hessenberg scalars
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cdiv(cplx<R> a, cplx<R> b) {
    R d = b.x * b.x + b.y * b.y;
    return mk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
// a / b with one division (by |b|^2) and two multiplications: the block solves'
// back-substitution, where the quotient's last-bit rounding does not matter
template <typename R>
__host__ __device__ __forceinline__ cplx<R> cdiv_recip(cplx<R> a, cplx<R> b) {
    const R id = R(1) / (b.x * b.x + b.y * b.y);
    return mk((a.x * b.x + a.y * b.y) * id, (a.y * b.x - a.x * b.y) * id);
}
template <typename R>
__host__ __device__ __forceinline__ cplx<R> czero() { return mk<R>(0, 0); }

template <typename To, typename From>
__host__ __device__ __forceinline__ cplx<To> cast(cplx<From> a) {
    return mk<To>((To)a.x, (To)a.y);
}

__device__ __forceinline__ bool rhs_on(uint64_t mask, int r) { return (mask >> r) & 1ull; }

// wave64 butterfly sum over `width` lanes (power of two); every lane gets the same bits
template <typename R>
__device__ __forceinline__ R wave_sum(R v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <typename R>
__device__ __forceinline__ cplx<R> wave_sum(cplx<R> v, int width) {
    for (int o = width >> 1; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o, 64);
        v.y += __shfl_xor(v.y, o, 64);
    }
    return v;
}

// Work-group sum of a complex value over all threads (blockDim.x a multiple of 64,
// at most 1024).  `scratch` holds 2*16 R.  Every thread receives the same total,
// summed in a fixed order.  Must be reached by the whole work-group.
template <typename R>
__device__ __forceinline__ cplx<R> block_sum(cplx<R> v, R* scratch) {
    v = wave_sum(v, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) {
        scratch[2 * wid] = v.x;
        scratch[2 * wid + 1] = v.y;
    }
    __syncthreads();
    cplx<R> t = czero<R>();
    for (int w = 0; w < nw; w++) {
        t.x += scratch[2 * w];
        t.y += scratch[2 * w + 1];
    }
    return t;
}

}  // namespace dd

#define DD_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            throw dd::HipFailure(e_);                                                  \
        }                                                                              \
    } while (0)

namespace dd {
struct HipFailure {
    hipError_t err;
    explicit HipFailure(hipError_t e) : err(e) {}
};
struct BadShape {
    const char* what;
    explicit BadShape(const char* w) : what(w) {}
};
}  // namespace dd
