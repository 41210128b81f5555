// Persistent cooperative coarsest-level GMRES (gfx950), see coarsest.h.
//
// One launch solves the coarsest system for a batch: a grid of G co-resident
// work-groups (8 waves each) owns the lattice in slices of `spw` sites; work-group s
// also owns the scalar state of system s (Givens rotations, right-hand side of the
// least-squares problem, last Hessenberg column) in its LDS.  Grid barriers separate the
// phases of every Arnoldi step.
//
// Arnoldi with delayed re-orthogonalisation (DCGS2).  Before step j the basis
// q_0..q_{j-1} is final and u_j = A q_{j-1} - sum_i h1_{i,j-1} q_i has had one
// projection.  Step j:
//   A  w = A u_j (matrix-core stencil); per work-group partial sums of
//      a_i = <q_i,u_j>, b_i = <q_i,w> (i < j), c = <u_j,u_j>, d = <u_j,w>: one pass over
//      the basis, 8 (FP32) or 4 (FP64) vectors in flight, their 32 or 16 partial sums
//      reduced across the wave with about as many shuffles;
//   B  (work-group s) beta = sqrt(c - |a|^2); column j-1 of the Hessenberg matrix is
//      h1_{.,j-1} + a with beta below (the second projection, one step late); Givens
//      rotation of that column and the break test |g_j| < tol ||eta||; first-pass
//      coefficients of the next column h1_{i,j} = b_i/beta, h1_{j,j} = (d - a^H b)/beta^2;
//   C  q_j = (u_j - Q a)/beta,  u_{j+1} = w/beta - Q_j h1_{.,j}  (one pass).
// The basis may be stored in IEEE binary16 (compressed-basis GMRES, Aliaga, Anzt,
// Gruetzmacher, Quintana-Orti, Tomas 2022): q_j is rounded once when it is written and
// every product is formed in the working precision from the stored values, so the
// orthogonalisation and the Arnoldi relation refer to the stored basis; the passes then
// move half the bytes.
// Using w/beta for A q_j leaves an error A Q a/beta inside span(Q_j); the next step's
// projection removes it and adds it to the Hessenberg column, so the Arnoldi relation
// holds to working precision.  The break of column j-1 is seen at step j, one stencil
// application later than with modified Gram-Schmidt; the solution only uses the
// columns before the break.

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "coarsest.h"
#include "lanes.h"
#include "mfma.h"


namespace dd {

namespace {

constexpr int kCT = 512;     // threads per work-group
constexpr int kNW = kCT / 64;
constexpr int kSpwMax = 8;   // sites per work-group at most
constexpr int kGmax = 256;   // work-groups at most (one per compute unit)
constexpr int kMaxM = 432;   // restart length at most (LDS state below 64 KB)
constexpr int kBB = 32;      // block of the triangular solve
// diagnostic clocks per work-group: 9 phases, then each wave's own reduction and update
// pass times (kNW each)
constexpr int kClk = 9 + 2 * (512 / 64);

enum : int { F_LIVE = 1, F_RUN = 2 };

struct Ctrl {
    int flags, maxit, iters, pad;
};

// Grid barrier of the persistent kernel.  Every work-group arrives on one counter; the
// last to arrive resets it and advances the generation that every work-group waits for.
// (A flat counter measured faster on 256 work-groups than per-group counters: the
// arrivals are device-scope atomics of the memory side, and one level fewer of them
// shortens the chain.)  Thread 0 of each work-group carries the device-scope release
// (before arriving) and acquire (after the generation moved) for the whole work-group,
// between two work-group barriers.  Waits are bounded: a wait that exceeds kBarSpins
// polls (seconds) raises `fail`, which every waiting work-group also watches, and the
// kernel ends; the host reports it.  All fields are zero when a launch starts.
constexpr unsigned kBarSpins = 1u << 25;
constexpr int kFlagStride = 16;   // words between two flags (each flag in its own 64-byte line)
constexpr int kPolyMax = 32;      // degree of the polynomial preconditioner at most

// Each word that work-groups update or poll sits in its own 256-byte region (one memory
// channel's interleave unit), so that the barrier's atomics, the polling of its generation
// and the flag polls do not queue behind each other at one channel.
struct GridBar {
    unsigned cnt;
    unsigned pad0[63];
    unsigned gen;
    unsigned pad1[63];
    unsigned fail;
    unsigned pad2[63];
    // per system: (Arnoldi step << 1) | stop, stored by the system's owner with release
    // once the step's update coefficients are written (see wait_coefs)
    unsigned cflag[64 * kFlagStride];
    // per work-group: the last Arnoldi step whose update pass it has finished (release),
    // read by the work-groups whose stencil needs its rows (see wait_neighbours)
    unsigned udone[kGmax * kFlagStride];
    // per work-group: the last Arnoldi step whose partial sums it has written (release),
    // read by the owners of the running systems (see phaseB)
    unsigned pdone[kGmax * kFlagStride];
    // per work-group: the number of polynomial-preconditioner stages whose rows it has
    // written (release), read by the work-groups whose stencil needs them (poly_stages)
    unsigned zdone[kGmax * kFlagStride];
};

// Executed by every wave that stored data before the work-group barrier that precedes a
// flag release by one thread: waits until this wave's global stores are acknowledged by
// the L2.  The releasing thread's agent-scope release then writes back the L2 (all waves'
// lines) and waits only for its own wave's stores.
__device__ __forceinline__ void release_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// returns false when the kernel must end (a barrier timed out somewhere)
__device__ __forceinline__ bool grid_barrier(GridBar* b, unsigned& gen, int wg, int G, int* sflag) {
    (void)wg;
    release_stores();
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned old = __hip_atomic_fetch_add(&b->cnt, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == (unsigned)G - 1) {
            __hip_atomic_store(&b->cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned spins = 0;
        while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
            if ((++spins & 63u) == 0 &&
                (__hip_atomic_load(&b->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || spins > kBarSpins)) {
                __hip_atomic_store(&b->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        gen++;
        *sflag = ok;
    }
    __syncthreads();
    return *sflag != 0;
}

// complex number of two IEEE binary16 values (basis storage)
struct chalf {
    _Float16 x, y;
};
template <typename R>
__device__ __forceinline__ cplx<R> widen(cplx<R> v) { return v; }
__device__ __forceinline__ cplx<float> widen(chalf v) { return mk<float>((float)v.x, (float)v.y); }
template <typename R>
__device__ __forceinline__ cplx<R> narrow(cplx<R>* d, cplx<R> v) {   // store, return the stored value
    *d = v;
    return v;
}
__device__ __forceinline__ cplx<float> narrow(chalf* d, cplx<float> v) {
    chalf h;
    h.x = (_Float16)v.x;
    h.y = (_Float16)v.y;
    *d = h;
    return widen(h);
}

// Complex multiply-accumulates as four fused multiply-adds each (a stored binary16
// operand is widened inside the FMA): acc += conj(q) v and acc -= c q.
template <typename R>
__device__ __forceinline__ void cmac_conj(cplx<R>& acc, cplx<R> q, cplx<R> v) {
    acc.x = fma(q.x, v.x, fma(q.y, v.y, acc.x));
    acc.y = fma(q.x, v.y, fma(-q.y, v.x, acc.y));
}
template <typename R>
__device__ __forceinline__ void cmsub(cplx<R>& acc, cplx<R> c, cplx<R> q) {
    acc.x = fma(-c.x, q.x, fma(c.y, q.y, acc.x));
    acc.y = fma(-c.x, q.y, fma(-c.y, q.x, acc.y));
}
template <typename R>
__device__ __forceinline__ void cmadd(cplx<R>& acc, cplx<R> c, cplx<R> q) {
    acc.x = fma(c.x, q.x, fma(-c.y, q.y, acc.x));
    acc.y = fma(c.x, q.y, fma(c.y, q.x, acc.y));
}

// Epilogue of one stage of the polynomial preconditioner's product form: z = v - c D v
// (c = 1/theta_i, v the stage's input row), stored as out = z; for the last stage of an
// Arnoldi step (extra = u_j) as out = u_j - z, which is D p(D) u_j; with xacc also
// xacc += cx z (the solution's accumulation at the end of a cycle).
template <typename R>
struct PolyEpi {
    const cplx<R>* extra;
    cplx<R>* xacc;
    cplx<R> c, cx;
    __device__ __forceinline__ void operator()(cplx<R>* __restrict__ out, size_t i, cplx<R> v, cplx<R> Dv) const {
        cplx<R> z = v;
        cmsub(z, c, Dv);
        if (extra) z = extra[i] - z;
        out[i] = z;
        if (xacc) {
            cplx<R> xv = xacc[i];
            cmadd(xv, cx, z);
            xacc[i] = xv;
        }
    }
};

template <typename R, typename S>
struct CArgs {
    const cplx<R>* L;       // links of the coarsest level
    const cplx<R>* eta;     // right-hand sides [nsys][n]
    cplx<R>* x;             // solutions [nsys][n]
    S* Q;                   // basis [m+1][nsys][n] (storage type S)
    cplx<R>* u;             // [nsys][n]
    cplx<R>* w;             // [nsys][n]
    cplx<R>* part;          // [G][nsys][2m+2] per-work-group partial sums (working precision:
                            // they are sums formed in it, the totals are taken in FP64)
    cplx<double>* Hrot;     // [nsys][m+1][m] rotated (triangular) Hessenberg columns
    cplx<R>* etals;         // [nsys][m] least-squares solution (working precision)
    cplx<R>* coef;          // [nsys][2m+2]: a_i (i<j) | h1_{i,j} (i<=j) | 1/beta
    double* nrm;            // [nsys][2]: ||eta||, ||r||
    Ctrl* ctrl;             // [nsys]
    double* stats;          // [2 + 8]: algorithmic bytes, Arnoldi steps, phase clocks of
                            // work-group 0 (A, barrier, B, barrier, C, barrier, cycle end)
    GridBar* bar;           // grid barrier state (zero at launch)
    double* wgclk;          // [G][kClk] per-work-group phase clocks and per-wave pass times
                            // (diagnostics), or null
    uint64_t mask;
    const uint64_t* dmask;  // device word ANDed into mask at the start (nullptr: none)
    int Nx, Nt, nsys, m, restarts, spw;
    int lds_links;          // 1: the slice's links are cached in LDS (after the scalar state)
    R m0;
    double tol;
    // GMRES polynomial right preconditioner p(D) of degree npoly (0: none): the Arnoldi
    // process runs on D p(D) and the solution is x0 + p(D) Q y
    int npoly;
    const cplx<R>* pc;      // [npoly] 1/theta_i of its roots in order of application
    cplx<R>* z;             // [2][cap][n]: the stages of its product form (ping-pong),
    size_t zs;              // the second buffer at z + zs
};

// LDS of the kernel: the scalar state of the owned system and scratch, then (lds_links)
// the links of the slice's sites, 16-byte aligned
__host__ __device__ constexpr size_t coarsest_lds(int m) {
    return ((sizeof(cplx<double>) * ((size_t)(7 * m + 5) + kBB * kBB) + 64) + 15) & ~(size_t)15;
}

// One site of the level stencil (as coarse_site in mfma.h, one 16-system tile) with the
// site's links in LDS (Ls: [5][DOF][DOF] column-major complex, as reals): the five hops'
// vector loads and the own rows are issued together, then the products take their A
// operands from LDS, so the site costs one global-load latency instead of five.
// The K index of the 16x16x4 products is ordered k = KS*g + s (lane group g, K-step s;
// mfma.h uses k = 4s + g; any bijection works when A and B agree), so a lane's B operand
// is KS consecutive reals of the neighbour's vector: KS/4 16-byte loads per hop.  The A
// element (row m = 2r + p, k = 2c + q) is Lr for p == q, Li for p != q, negated for
// (p, q) = (0, 1); with c = (KS/2) g + s/2 and q = s & 1 its LDS offset is a lane part
// plus a constant per (s, t).  FP32 only (the accumulator rows m - 16t = i of mfma.h).
// Without RES the result goes through epi(out, row index, own input value, result).
template <typename R, int DOF, bool RES, typename Epi = StoreEpi<R>>
__device__ __forceinline__ void coarse_site_lds(const R* Ls, const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out,
                                                const cplx<R>* __restrict__ b, int Nx, int Nt, R m0, uint64_t mask,
                                                size_t ld, int nrhs, int n, int tile, int lane,
                                                const Epi& epi = Epi()) {
    using acc_t = typename Mfma<R>::acc_t;
    constexpr int MT = DOF / 8, KS = DOF / 2;
    const int g = lane >> 4;
    const int jj = tile * 16 + (lane & 15);
    const bool on = jj < nrhs && ((mask >> jj) & 1ull);
    const size_t sysoff = (size_t)(on ? jj : 0) * ld;
    const R* vin = reinterpret_cast<const R*>(in + sysoff);
    const int x = n / Nt, t = n - x * Nt;
    const int nb[5] = {n, (t == Nt - 1) ? n - (Nt - 1) : n + 1, (x == Nx - 1) ? t : n + Nt,
                       (t == 0) ? n + (Nt - 1) : n - 1, (x == 0) ? (Nx - 1) * Nt + t : n - Nt};
    const R sg[5] = {R(1), (t == Nt - 1) ? R(-1) : R(1), R(1), (t == 0) ? R(-1) : R(1), R(1)};
    static_assert(sizeof(R) == 4 && KS % 4 == 0, "FP32, K-steps in groups of four reals");
    R bv[5][KS];
#pragma unroll
    for (int h = 0; h < 5; h++) {
        const float4* vb = reinterpret_cast<const float4*>(vin + (size_t)nb[h] * DOF * 2 + KS * g);
#pragma unroll
        for (int s4 = 0; s4 < KS / 4; s4++) {
            const float4 v4 = vb[s4];
            bv[h][4 * s4] = v4.x;
            bv[h][4 * s4 + 1] = v4.y;
            bv[h][4 * s4 + 2] = v4.z;
            bv[h][4 * s4 + 3] = v4.w;
        }
    }
    const size_t base = sysoff + (size_t)n * DOF;
    cplx<R> own[MT][2], rhs[MT][2];
#pragma unroll
    for (int tt = 0; tt < MT; tt++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
            own[tt][e] = in[base + 8 * tt + 2 * g + e];
            if constexpr (RES) rhs[tt][e] = b[base + 8 * tt + 2 * g + e];
            else rhs[tt][e] = czero<R>();
        }
#pragma unroll
    for (int h = 0; h < 5; h++)
#pragma unroll
        for (int s2 = 0; s2 < KS; s2++) bv[h][s2] = on ? sg[h] * bv[h][s2] : R(0);
    // lane part of the A offsets: row r = 8t + i/2, p = i & 1; columns (KS/2) g + s/2
    const int i = lane & 15, p = i & 1;
    const int lbase = 2 * ((KS / 2) * g * DOF + (i >> 1));
    const int oe = lbase + p, oo = lbase + (1 - p);   // even / odd K-steps (q = 0 / 1)
    const bool negodd = p == 0;
    acc_t acc[1][MT];
#pragma unroll
    for (int tt = 0; tt < MT; tt++) acc[0][tt] = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int h = 0; h < 5; h++) {
        const R* Lh = Ls + (size_t)h * DOF * DOF * 2;
#pragma unroll
        for (int s2 = 0; s2 < KS; s2++) {
#pragma unroll
            for (int tt = 0; tt < MT; tt++) {
                const int c = 2 * ((s2 >> 1) * DOF + 8 * tt);
                R av = (s2 & 1) ? Lh[oo + c] : Lh[oe + c];
                if ((s2 & 1) && negodd) av = -av;
                acc[0][tt] = Mfma<R>::mac(av, bv[h][s2], acc[0][tt]);
            }
        }
    }
    if (!on) return;
    const R mass = m0 + R(2);
#pragma unroll
    for (int tt = 0; tt < MT; tt++) {
        const int r0 = 8 * tt + 2 * g;
        const cplx<R> o0 = mk<R>(mass * own[tt][0].x - acc[0][tt][0], mass * own[tt][0].y - acc[0][tt][1]);
        const cplx<R> o1 = mk<R>(mass * own[tt][1].x - acc[0][tt][2], mass * own[tt][1].y - acc[0][tt][3]);
        if constexpr (RES) {
            out[base + r0] = rhs[tt][0] - o0;
            out[base + r0 + 1] = rhs[tt][1] - o1;
        } else {
            epi(out, base + r0, own[tt][0], o0);
            epi(out, base + r0 + 1, own[tt][1], o1);
        }
    }
}

// coarse_site_lds in two phases, for a caller that waits for its neighbours in between:
// PHASE 0 clears acc and accumulates the hops whose neighbour site lies inside the slice
// [s0, s1) (rows the work-group wrote itself), PHASE 1 adds the other hops and runs the
// epilogue.  The caller's slices are runs of at least two sites of one lattice row, so a
// phase has at most three hops: they are gathered into three slots (shifted in, so that
// every register index stays static), the last selected hop first.
template <typename R, int DOF, int PHASE, typename Epi>
__device__ __forceinline__ void coarse_site_lds_phase(const R* Ls, const cplx<R>* __restrict__ in,
                                                      cplx<R>* __restrict__ out, int Nx, int Nt, R m0, uint64_t mask,
                                                      size_t ld, int nrhs, int n, int tile, int lane, int s0, int s1,
                                                      typename Mfma<R>::acc_t (&acc)[DOF / 8], const Epi& epi) {
    constexpr int MT = DOF / 8, KS = DOF / 2;
    static_assert(sizeof(R) == 4 && KS % 4 == 0, "FP32, K-steps in groups of four reals");
    const int g = lane >> 4;
    const int jj = tile * 16 + (lane & 15);
    const bool on = jj < nrhs && ((mask >> jj) & 1ull);
    const size_t sysoff = (size_t)(on ? jj : 0) * ld;
    const R* vin = reinterpret_cast<const R*>(in + sysoff);
    const int x = n / Nt, t = n - x * Nt;
    const int nb[5] = {n, (t == Nt - 1) ? n - (Nt - 1) : n + 1, (x == Nx - 1) ? t : n + Nt,
                       (t == 0) ? n + (Nt - 1) : n - 1, (x == 0) ? (Nx - 1) * Nt + t : n - Nt};
    const R sg[5] = {R(1), (t == Nt - 1) ? R(-1) : R(1), R(1), (t == 0) ? R(-1) : R(1), R(1)};
    // the phase's hops: neighbour site, boundary sign, hop index per slot
    int cn0 = n, cn1 = n, cn2 = n, ch0 = 0, ch1 = 0, ch2 = 0, nh = 0;
    R cg0 = R(0), cg1 = R(0), cg2 = R(0);
#pragma unroll
    for (int h = 0; h < 5; h++) {
        const bool local = nb[h] >= s0 && nb[h] < s1;
        if ((local ? 0 : 1) == PHASE) {
            cn2 = cn1;
            cn1 = cn0;
            cn0 = nb[h];
            cg2 = cg1;
            cg1 = cg0;
            cg0 = sg[h];
            ch2 = ch1;
            ch1 = ch0;
            ch0 = h;
            ++nh;
        }
    }
    if constexpr (PHASE == 0) {
#pragma unroll
        for (int tt = 0; tt < MT; tt++) acc[tt] = typename Mfma<R>::acc_t{0, 0, 0, 0};
    }
    R bv[3][KS];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (c >= nh) continue;
        const int nbc = c == 0 ? cn0 : c == 1 ? cn1 : cn2;
        const float4* vb = reinterpret_cast<const float4*>(vin + (size_t)nbc * DOF * 2 + KS * g);
#pragma unroll
        for (int s4 = 0; s4 < KS / 4; s4++) {
            const float4 v4 = vb[s4];
            bv[c][4 * s4] = v4.x;
            bv[c][4 * s4 + 1] = v4.y;
            bv[c][4 * s4 + 2] = v4.z;
            bv[c][4 * s4 + 3] = v4.w;
        }
    }
    const size_t base = sysoff + (size_t)n * DOF;
    cplx<R> own[MT][2];
    if constexpr (PHASE == 1) {
#pragma unroll
        for (int tt = 0; tt < MT; tt++)
#pragma unroll
            for (int e = 0; e < 2; e++) own[tt][e] = in[base + 8 * tt + 2 * g + e];
    }
    const int i = lane & 15, p = i & 1;
    const int lbase = 2 * ((KS / 2) * g * DOF + (i >> 1));
    const int oe = lbase + p, oo = lbase + (1 - p);
    const bool negodd = p == 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (c >= nh) continue;
        const int hc = c == 0 ? ch0 : c == 1 ? ch1 : ch2;
        const R gc = c == 0 ? cg0 : c == 1 ? cg1 : cg2;
        const R* Lh = Ls + (size_t)hc * DOF * DOF * 2;
#pragma unroll
        for (int s2 = 0; s2 < KS; s2++) {
            const R bb = on ? gc * bv[c][s2] : R(0);
#pragma unroll
            for (int tt = 0; tt < MT; tt++) {
                const int cc = 2 * ((s2 >> 1) * DOF + 8 * tt);
                R av = (s2 & 1) ? Lh[oo + cc] : Lh[oe + cc];
                if ((s2 & 1) && negodd) av = -av;
                acc[tt] = Mfma<R>::mac(av, bb, acc[tt]);
            }
        }
    }
    if constexpr (PHASE == 1) {
        if (!on) return;
        const R mass = m0 + R(2);
#pragma unroll
        for (int tt = 0; tt < MT; tt++) {
            const int r0 = 8 * tt + 2 * g;
            const cplx<R> o0 = mk<R>(mass * own[tt][0].x - acc[tt][0], mass * own[tt][0].y - acc[tt][1]);
            const cplx<R> o1 = mk<R>(mass * own[tt][1].x - acc[tt][2], mass * own[tt][1].y - acc[tt][3]);
            epi(out, base + r0, own[tt][0], o0);
            epi(out, base + r0 + 1, own[tt][1], o1);
        }
    }
}

__device__ __forceinline__ cplx<double> cdivr(cplx<double> a, double d) { return mk<double>(a.x / d, a.y / d); }

// Sums NV values (a power of two, 4 <= NV <= 32) over each group of LPS lanes (32 or 64)
// with NV lane exchanges (each step hands half of the values to the partner lane,
// through the DPP / permlane networks): afterwards lane li = lane mod LPS holds the
// group's total of value li >> (log2 LPS - log2 NV).  The stages are instantiated per
// offset, so every array index is a constant and the values stay in registers.
constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
template <typename R, int NV, int LPS, int ST>
__device__ __forceinline__ void bfly_stage(R (&F)[NV], int lane) {
    constexpr int O = (LPS / 2) >> ST, HALF = (NV / 2) >> ST;
#pragma unroll
    for (int k = 0; k < HALF; k++) {
        R A = F[k], B = F[k + HALF];
        pair_swap<O>(A, B, lane);
        F[k] = A + B;
    }
    if constexpr (ST + 1 < ilog2c(NV)) bfly_stage<R, NV, LPS, ST + 1>(F, lane);
}
template <typename R, int O>
__device__ __forceinline__ R bfly_tail(R t, int lane) {   // t + partners at offsets O, O/2, ..., 1
    if constexpr (O >= 1) return bfly_tail<R, O / 2>(t + lane_xor<O>(t, lane), lane);
    else return t;
}
template <typename R, int NV, int LPS>
__device__ __forceinline__ R butterfly(R (&F)[NV], int lane) {
    static_assert(NV >= 4 && NV <= 32 && (NV & (NV - 1)) == 0, "NV must be a power of two in 4..32");
    static_assert(LPS == 32 || LPS == 64, "lane groups of 32 or 64");
    bfly_stage<R, NV, LPS, 0>(F, lane);
    return bfly_tail<R, ((LPS / 2) >> ilog2c(NV))>(F[0], lane);
}

constexpr int pow2_floor(int v) {
    int p = 1;
    while (2 * p <= v) p *= 2;
    return p;
}
// basis vectors in flight per wave for storage S and RPL rows per lane: the reduction
// pass holds up to 96 registers of basis values (binary16, FP32; 4 vectors for FP64), the
// update passes half as many next to their coefficients
template <typename S, int RPL>
struct Chunk {
    static constexpr int words = (int)sizeof(S) / 4;
    static constexpr int RED = sizeof(S) == 16 ? 4 : pow2_floor(96 / (RPL * words) < 32 ? 96 / (RPL * words) : 32);
    static constexpr int UPD = sizeof(S) == 16 ? 4 : (RED / 2 < 4 ? 4 : RED / 2);
};

// Lanes and rows: a system's rows of the slice (spw sites of DOF rows) are spread over a
// group of LPS lanes, RPL = DOF / 8 consecutive rows per lane, so LPS = 32 holds up to 4
// sites and LPS = 64 up to 8; with LPS = 32 a wave works on two systems at once.
// POLY: the variant with the polynomial preconditioner (a.npoly > 0); the plain one
// compiles without its code.
template <typename R, typename S, int DOF, int LPS, bool POLY>
__global__ __launch_bounds__(kCT) void k_coarsest(CArgs<R, S> a) {
    constexpr int RPL = DOF / 8, SPI = 64 / LPS;
    const int npoly = POLY ? a.npoly : 0;
    // the stage buffers by address arithmetic (a select between two pointers compiles to
    // a table in scratch memory)
    cplx<R>* const zb0 = a.z;
    const size_t zstride = a.zs;
    auto zbuf = [zb0, zstride](int i) { return zb0 + (size_t)(i & 1) * zstride; };
    static_assert(DOF % 8 == 0, "rows per lane must divide the site's rows");
    // kCH basis vectors loaded per pass step, reduced in groups of kGV (NVAL sums each)
    constexpr int kCH = Chunk<S, RPL>::RED, kCU = Chunk<S, RPL>::UPD, kGV = kCH < 8 ? kCH : 8, NVAL = 4 * kGV,
                  SH = ilog2c(LPS) - ilog2c(NVAL);
    static_assert(kCH % kGV == 0 && (kGV == 8 || kGV == 4), "chunk of 4, 8 or a multiple of 8 vectors");
    // systems decided on the device (a device-controlled caller): every work-group reads
    // the same word; a launch with no system left ends at once
    if (a.dmask) a.mask &= *a.dmask;
    if (!a.mask) return;
    unsigned bar_gen = 0;
    extern __shared__ unsigned char smem_raw[];
    const int m = a.m, VALS = 2 * m + 2, CST = 2 * m + 2;
    // LDS: per-system state of the work-group's system, then scratch
    cplx<double>* tot = reinterpret_cast<cplx<double>*>(smem_raw);   // [VALS] step totals
    cplx<double>* h1p = tot + VALS;                                    // [m+1] first-pass column
    cplx<double>* csl = h1p + (m + 1);                                 // [m]
    cplx<double>* snl = csl + m;                                       // [m]
    cplx<double>* gls = snl + m;                                       // [m+1]
    cplx<double>* hcol = gls + (m + 1);                                // [m+1]
    double* sc = reinterpret_cast<double*>(hcol + (m + 1));            // [4]
    int* si = reinterpret_cast<int*>(sc + 4);                          // [4]
    uint64_t* smask = reinterpret_cast<uint64_t*>(si + 4);             // [1]
    cplx<double>* blk = reinterpret_cast<cplx<double>*>(smask + 2);    // [kBB][kBB] triangular solve
    cplx<double>* red = blk;   // [kNW][64] wave sums of the totals (phase B; blk is used after it)
    static_assert(kBB * kBB >= kNW * 64, "LDS scratch of the totals");

    // wv is wave-uniform (read from the first lane, so that per-wave loops and addresses
    // stay scalar)
    const int G = gridDim.x, wg = blockIdx.x, tid = threadIdx.x, lane = tid & 63,
              wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int V = a.Nx * a.Nt;
    const size_t n = (size_t)V * DOF, vs = n * (size_t)a.nsys;
    const int s0 = wg * a.spw, s1 = min(V, s0 + a.spw);
    const int nrow = max(0, s1 - s0) * DOF;   // at most LPS * RPL
    const size_t row0 = (size_t)s0 * DOF;
    const int ntile = (a.nsys + 15) / 16;
    const R* lnk = reinterpret_cast<const R*>(smem_raw + coarsest_lds(m));   // lds_links: [spw][5][DOF][DOF]
    const double sR = sizeof(cplx<R>), sQ = sizeof(S);
    // the lane's system within a wave's pair (h) and its rows lb..lb+RPL-1 (lanes past the
    // slice read its last rows, masked)
    const int h = lane / LPS, li = lane % LPS;
    const bool lrow = li * RPL < nrow;
    const int lb = lrow ? li * RPL : nrow - RPL;
    const int sys = wg;   // the system whose scalars this work-group owns (G >= nsys)
    // lane (u, d) = (lane >> 2, lane & 3) of wave 0: the work-group owning the neighbour of
    // own site s0 + u in direction d (t+1, x+1, t-1, x-1), or -1 (none, or this work-group)
    int nbw = -1;
    {
        const int u = lane >> 2, d = lane & 3;
        if (wv == 0 && s0 + u < s1 && u < kSpwMax) {
            const int site = s0 + u, x = site / a.Nt, t = site - x * a.Nt;
            const int nx = d == 1 ? (x + 1) % a.Nx : d == 3 ? (x + a.Nx - 1) % a.Nx : x;
            const int nt = d == 0 ? (t + 1) % a.Nt : d == 2 ? (t + a.Nt - 1) % a.Nt : t;
            const int w2 = (nx * a.Nt + nt) / a.spw;
            nbw = w2 == wg ? -1 : w2;
        }
    }
    const bool owner = sys < a.nsys && ((a.mask >> sys) & 1ull);
    // partial sums: work-group-major [G][nsys][VALS], each work-group writes its own
    // contiguous block
    auto pidx = [&](int q, int v, int w) -> size_t { return ((size_t)w * a.nsys + q) * VALS + v; };

    const double links_bytes = (double)V * 5 * DOF * DOF * sR;
    // links read by one stencil sweep: none once they sit in LDS (loaded once per launch)
    const double sweep_links = a.lds_links ? 0.0 : links_bytes;
    auto count = [&](double bytes, double steps) {
        if (wg == 0 && tid == 0) {
            a.stats[0] += bytes;
            a.stats[1] += steps;
        }
    };
    uint64_t tclk = wall_clock64();
    // wall-clock time per phase: work-group 0, and every work-group in the diagnostic
    // mode, where a phase ends when all waves of the work-group are done with it
    auto clk = [&](int slot) {
        if (a.wgclk) __syncthreads();
        if (tid == 0) {
            const uint64_t t = wall_clock64();
            if (wg == 0) a.stats[2 + slot] += (double)(t - tclk);
            if (a.wgclk) a.wgclk[wg * kClk + slot] += (double)(t - tclk);
            tclk = t;
        }
    };

    // systems whose control word has `flag`, read by the 64 lanes of wave 0 at once and
    // broadcast through LDS.  Control words are written between grid barriers and read
    // after them (the barrier orders and publishes them), so plain loads and stores
    // suffice and the result is the same in every work-group.  Owners write them in phase
    // B, before the release of their coefficient flags, and every work-group reads them
    // after it has acquired those flags.
    auto sysmask = [&](int flag, int off = 0) -> uint64_t {
        __syncthreads();
        if (wv == 0) {
            const int q = lane;
            bool on = false;
            if (q < a.nsys && ((a.mask >> q) & 1ull)) {
                const int fl = a.ctrl[q].flags;
                on = (fl & flag) != 0 && (fl & off) == 0;
            }
            const uint64_t mk2 = __ballot(on);
            if (lane == 0) *smask = mk2;
        }
        __syncthreads();
        return *smask;
    };
    auto set_flags = [&](int q, int fl) { a.ctrl[q].flags = fl; };
    // sites of the slice, wave by wave; tiles of 16 systems in pairs (one link load for
    // both).  Slices of at most 4 sites give a wave one (site, tile) each.
    auto stencil = [&](const cplx<R>* in, cplx<R>* out, const cplx<R>* b, uint64_t mk2, bool res) {
        if (LPS == 32) {
            const int ns = s1 - s0;
            for (int u = wv; u < ns * ntile; u += kNW) {
                const int site = s0 + u % ns, tl = u / ns;
                if (!((mk2 >> (16 * tl)) & 0xffffull)) continue;
                if constexpr (sizeof(R) == 4 && DOF <= 24) {   // the cases whose links fit the LDS
                    if (a.lds_links) {
                        const R* Ls = lnk + (size_t)(site - s0) * 5 * DOF * DOF * 2;
                        if (res)
                            coarse_site_lds<R, DOF, true>(Ls, in, out, b, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl,
                                                          lane);
                        else
                            coarse_site_lds<R, DOF, false>(Ls, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site,
                                                           tl, lane);
                        continue;
                    }
                }
                if (res)
                    coarse_site<R, DOF, true, 1>(a.L, in, out, b, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl, lane);
                else
                    coarse_site<R, DOF, false, 1>(a.L, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl,
                                                  lane);
            }
            return;
        }
        for (int site = s0 + wv; site < s1; site += kNW)
            for (int tl = 0; tl < ntile; tl += 2) {
                if (!((mk2 >> (16 * tl)) & 0xffffffffull)) continue;
                if (tl + 1 < ntile) {
                    if (res)
                        coarse_site<R, DOF, true, 2>(a.L, in, out, b, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl, lane);
                    else
                        coarse_site<R, DOF, false, 2>(a.L, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site,
                                                      tl, lane);
                } else {
                    if (res)
                        coarse_site<R, DOF, true, 1>(a.L, in, out, b, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl, lane);
                    else
                        coarse_site<R, DOF, false, 1>(a.L, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site,
                                                      tl, lane);
                }
            }
    };
    // the same sweep with the result going through an epilogue (no residual form)
    auto stencil_e = [&](const cplx<R>* in, cplx<R>* out, uint64_t mk2, const auto& epi) {
        if (LPS == 32) {
            const int ns = s1 - s0;
            for (int u = wv; u < ns * ntile; u += kNW) {
                const int site = s0 + u % ns, tl = u / ns;
                if (!((mk2 >> (16 * tl)) & 0xffffull)) continue;
                if constexpr (sizeof(R) == 4 && DOF <= 24) {
                    if (a.lds_links) {
                        const R* Ls = lnk + (size_t)(site - s0) * 5 * DOF * DOF * 2;
                        coarse_site_lds<R, DOF, false>(Ls, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl,
                                                       lane, epi);
                        continue;
                    }
                }
                coarse_site<R, DOF, false, 1>(a.L, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl, lane,
                                              epi);
            }
            return;
        }
        for (int site = s0 + wv; site < s1; site += kNW)
            for (int tl = 0; tl < ntile; tl += 2) {
                if (!((mk2 >> (16 * tl)) & 0xffffffffull)) continue;
                if (tl + 1 < ntile)
                    coarse_site<R, DOF, false, 2>(a.L, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl,
                                                  lane, epi);
                else
                    coarse_site<R, DOF, false, 1>(a.L, in, out, nullptr, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl,
                                                  lane, epi);
            }
    };
    // The polynomial stages' sweeps in two phases around the neighbour wait (one
    // (site, tile) task per wave: FP32 slices of at most 4 sites with their links in LDS):
    // phase 0 the hops inside the slice, phase 1 the others and the epilogue; pacc holds a
    // wave's partial products in between.
    typename Mfma<R>::acc_t pacc[DOF / 8];
    bool split = false;
    if constexpr (POLY && LPS == 32 && sizeof(R) == 4 && DOF <= 24)
        split = a.lds_links && (s1 - s0) * ntile <= kNW && a.spw >= 2 && a.Nt % a.spw == 0 && a.Nx >= 2;
    auto stencil_phase = [&](auto phase_, const cplx<R>* in, cplx<R>* out, uint64_t mk2, const auto& epi) {
        if constexpr (POLY && LPS == 32 && sizeof(R) == 4 && DOF <= 24) {
            constexpr int PH = decltype(phase_)::value;
            const int ns = s1 - s0;
            if (wv >= ns * ntile) return;
            const int site = s0 + wv % ns, tl = wv / ns;
            if (!((mk2 >> (16 * tl)) & 0xffffull)) return;
            const R* Ls = lnk + (size_t)(site - s0) * 5 * DOF * DOF * 2;
            coarse_site_lds_phase<R, DOF, PH>(Ls, in, out, a.Nx, a.Nt, a.m0, mk2, n, a.nsys, site, tl, lane, s0, s1,
                                              pacc, epi);
        }
    };
    auto load_rows = [&](cplx<R> (&r)[RPL], const cplx<R>* vec) {   // vec: system base + row0
#pragma unroll
        for (int k = 0; k < RPL; k++) {
            const cplx<R> v = vec[lb + k];
            r[k] = lrow ? v : czero<R>();
        }
    };
    auto put1 = [&](int q, int v, cplx<R> p, bool act) {   // single work-group partial (lane group)
        p = wave_sum(p, LPS);
        if (li == 0 && act) a.part[pidx(q, v, wg)] = p;
    };
    // the wave's pair of systems qb*SPI + h; loads use a valid system index
    auto sys_of = [&](int qb) { return min(qb * SPI + h, a.nsys - 1); };
    auto pair_mask = [&](uint64_t mk2, int qb) -> unsigned { return (unsigned)(mk2 >> (qb * SPI)) & ((1u << SPI) - 1); };
    // ---- phase A reduction: a_i, c and (hw) b_i, d of system q over the own rows
    auto reduce_step = [&](int qb, int j, bool hw, unsigned pm) {
        const int q = sys_of(qb);
        const bool act = (pm >> h) & 1u;
        cplx<R> ur[RPL], wr[RPL];
        load_rows(ur, a.u + (size_t)q * n + row0);
        if (hw) load_rows(wr, a.w + (size_t)q * n + row0);
        else
#pragma unroll
            for (int k = 0; k < RPL; k++) wr[k] = czero<R>();
        cplx<R> c = czero<R>(), d = czero<R>();
#pragma unroll
        for (int k = 0; k < RPL; k++) {
            c.x += norm2(ur[k]);
            d = d + cmul(ur[k], wr[k]);
        }
        put1(q, 2 * m, c, act);
        if (hw) put1(q, 2 * m + 1, d, act);
        R* pd = reinterpret_cast<R*>(a.part);
        // chunks of NC vectors starting at i0: unconditional loads at clamped indices (no
        // branches between the loads), the terms outside the basis or the own rows masked
        // afterwards.  Full chunks of kCH, the tail in groups of kGV.
        auto chunk = [&](auto nc_, int i0) {
            constexpr int NC = decltype(nc_)::value;
            S qv[NC][RPL];
#pragma unroll
            for (int ii = 0; ii < NC; ii++) {
                const S* qi = a.Q + (size_t)min(i0 + ii, j - 1) * vs + (size_t)q * n + row0 + lb;
#pragma unroll
                for (int k = 0; k < RPL; k++) qv[ii][k] = qi[k];
            }
            // groups of kGV vectors: 4 kGV <= 32 sums per butterfly
#pragma unroll
            for (int g0 = 0; g0 < NC; g0 += kGV) {
                R F[NVAL];
#pragma unroll
                for (int gi = 0; gi < kGV; gi++) {
                    const int ii = g0 + gi;
                    // no mask on the basis values: lanes past the slice hold zero rows of u
                    // and w, and the sums of clamped duplicates (i0 + ii >= j) are not stored
                    cplx<R> pa = czero<R>(), pb = czero<R>();
#pragma unroll
                    for (int k = 0; k < RPL; k++) {
                        const cplx<R> qq = widen(qv[ii][k]);
                        cmac_conj(pa, qq, ur[k]);
                        cmac_conj(pb, qq, wr[k]);
                    }
                    F[4 * gi] = pa.x;
                    F[4 * gi + 1] = pa.y;
                    F[4 * gi + 2] = pb.x;
                    F[4 * gi + 3] = pb.y;
                }
                const R tsum = butterfly<R, NVAL, LPS>(F, lane);
                const int f = li >> SH, ii = g0 + (f >> 2), comp = f & 3;
                if (act && (li & ((1 << SH) - 1)) == 0 && i0 + ii < j && (hw || comp < 2)) {
                    const int v = (comp < 2 ? 0 : m) + i0 + ii;
                    pd[pidx(q, v, wg) * 2 + (comp & 1)] = tsum;
                }
            }
        };
        int i0 = 0;
        for (; i0 + kCH <= j; i0 += kCH) chunk(std::integral_constant<int, kCH>{}, i0);
        // the remainder as one chunk of the next multiple of kGV vectors, so that all its
        // loads are in flight together (at most kGV - 1 clamped duplicates, masked)
        const int ng = (j - i0 + kGV - 1) / kGV;
        if (ng == 1) chunk(std::integral_constant<int, kGV>{}, i0);
        if constexpr (2 * kGV <= kCH)
            if (ng == 2) chunk(std::integral_constant<int, 2 * kGV>{}, i0);
        if constexpr (3 * kGV <= kCH)
            if (ng == 3) chunk(std::integral_constant<int, 3 * kGV>{}, i0);
        if constexpr (4 * kGV <= kCH)
            if (ng == 4) chunk(std::integral_constant<int, 4 * kGV>{}, i0);
        static_assert(kCH <= 4 * kGV, "remainder cases");
    };
    auto reduce_norm = [&](uint64_t mk2, int v, const cplx<R>* vec) {
        for (int qb = wv; qb * SPI < a.nsys; qb += kNW) {
            const unsigned pm = pair_mask(mk2, qb);
            if (!pm) continue;
            const int q = sys_of(qb);
            cplx<R> r[RPL];
            load_rows(r, vec + (size_t)q * n + row0);
            cplx<R> c = czero<R>();
#pragma unroll
            for (int k = 0; k < RPL; k++) c.x += norm2(r[k]);
            put1(q, v, c, (pm >> h) & 1u);
        }
    };
    // total of value v of system q over the work-groups (whole work-group; every thread)
    auto total1 = [&](int q, int v) -> cplx<double> {
        if (wv == 0) {
            cplx<double> t = czero<double>();
            for (int k = lane; k < G; k += 64) t = t + cast<double>(a.part[pidx(q, v, k)]);
            t = wave_sum(t, 64);
            if (lane == 0) tot[v] = t;
        }
        __syncthreads();
        const cplx<double> r = tot[v];
        __syncthreads();
        return r;
    };
    // ---- phase B for the owned system at step j (whole work-group)
    auto phaseB = [&](int j, unsigned gstep) {
        const int q = sys;
        const bool hw = j < m;
        const int nv = hw ? 2 * j + 2 : j + 1;
        {   // totals: a lane per value, wave w adds the work-groups w, w + 8, ... (all
            // their partials loaded before the first addition); the waves' sums are added
            // in LDS in a fixed order
            constexpr int KW = kGmax / kNW, KB = sizeof(R) == 4 ? KW / 2 : KW / 4;
            // each wave first waits for the partial-sum flags of its own work-groups (lane u
            // polls w + 8u) and acquires them, so that a wave whose work-groups are done
            // gathers while the others still wait; a timed-out wait clears si[2].  Every
            // work-group publishes its partials after its stencil and reduction of the
            // step, which follow its update of the previous step: once an owner has seen
            // all of them, no work-group still reads the coefficients or the control flags
            // the owner is about to rewrite, nor the rows of u that their next updates
            // overwrite.
            {
                bool ok = true;
                const int k = wv + kNW * lane;
                if (lane < KW && k < G) {
                    unsigned spins = 0;
                    while (__hip_atomic_load(&a.bar->pdone[k * kFlagStride], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) < gstep) {
                        if ((++spins & 63u) == 0 &&
                            (__hip_atomic_load(&a.bar->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                             spins > kBarSpins)) {
                            __hip_atomic_store(&a.bar->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            ok = false;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (__ballot(!ok) && lane == 0) si[2] = 0;
            }
            for (int base = 0; base < nv; base += 64) {
                const int idx = base + lane;
                int v = 0;
                if (idx < j) v = idx;
                else if (hw && idx < 2 * j) v = m + (idx - j);
                else v = 2 * m + min(idx - (hw ? 2 * j : j), 1);
                cplx<double> t = czero<double>();
#pragma unroll
                for (int u0 = 0; u0 < KW; u0 += KB) {
                    cplx<R> pv[KB];
#pragma unroll
                    for (int u = 0; u < KB; u++) pv[u] = a.part[pidx(q, v, min(wv + kNW * (u0 + u), G - 1))];
#pragma unroll
                    for (int u = 0; u < KB; u++)
                        if (wv + kNW * (u0 + u) < G) t = t + cast<double>(pv[u]);
                }
                if (idx >= nv) t = czero<double>();
                red[wv * 64 + lane] = t;
                __syncthreads();
                if (wv == 0 && idx < nv) {
                    cplx<double> s2 = red[lane];
#pragma unroll
                    for (int w = 1; w < kNW; w++) s2 = s2 + red[w * 64 + lane];
                    tot[v] = s2;
                }
                __syncthreads();
            }
        }
        clk(8);
        if (wv == 0) {
            double sa = 0;
            cplx<double> ab = czero<double>();
            for (int i = lane; i < j; i += 64) {
                sa += norm2(tot[i]);
                if (hw) ab = ab + cmul(tot[i], tot[m + i]);
            }
            sa = wave_sum(sa, 64);
            ab = wave_sum(ab, 64);
            if (lane == 0) {
                const double b2 = tot[2 * m].x - sa;
                sc[0] = b2 > 0 ? sqrt(b2) : 0.0;
                sc[1] = ab.x;
                sc[2] = ab.y;
            }
        }
        __syncthreads();
        const double beta = sc[0];
        // the last step of the cycle, a breakdown and a break found by the previous step's
        // givens() (si[0]; it set maxIt) stop now: no update follows.  The Givens rotation
        // of column j-1 and its break test are left to givens(), after the coefficients
        // are published.
        const bool conv = si[0] != 0;
        const bool stop_now = j == m || beta == 0.0 || conv;
        if (tid == 0) {
            if (j == 0) gls[0] = mk<double>(beta, 0.0);
            if (stop_now) {
                if (!conv) a.ctrl[q].maxit = j;
                set_flags(q, a.ctrl[q].flags & ~F_RUN);
            }
        }
        // column j-1 of the Arnoldi Hessenberg matrix; then (same thread, same i) the
        // first-pass coefficients of column j and the update coefficients
        const double ib = 1.0 / beta;
        cplx<R>* cf = a.coef + (size_t)q * CST;
        for (int i = tid; i <= j; i += kCT) {
            if (j >= 1) hcol[i] = i < j ? h1p[i] + tot[i] : mk<double>(beta, 0.0);
            if (!stop_now) {
                const cplx<double> h1 =
                    i < j ? ib * tot[m + i] : (ib * ib) * (tot[2 * m + 1] - mk<double>(sc[1], sc[2]));
                h1p[i] = h1;
                cf[m + i] = cast<R>(h1);
                if (i < j) cf[i] = cast<R>(tot[i]);
            }
        }
        if (tid == 0 && !stop_now) cf[2 * m + 1] = mk<R>((R)ib, R(0));
        if (tid == 0) si[1] = stop_now;
        if (wv * 64 <= j) release_stores();   // the waves that stored coefficients
        __syncthreads();
        // publish: every thread's coefficient stores precede thread 0's release (not after
        // a timed-out wait: the kernel ends)
        if (tid == 0 && si[2])
            __hip_atomic_store(&a.bar->cflag[q * kFlagStride], (gstep << 1) | (stop_now ? 1u : 0u), __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
    };
    // Waits (wave 0, a lane per neighbour) until every work-group whose rows this slice's
    // stencil reads has finished the update pass of Arnoldi step gstep, then acquires
    // its stores; si[2] = 0 when the kernel must end.  Replaces a grid barrier after the
    // update pass: a slice waits for its (at most 4 spw) neighbours.  No work-group can
    // overwrite rows that a neighbour's stencil still reads: its next update waits for
    // coefficients that need every work-group's next partial sums, which follow that
    // work-group's stencil.
    auto wait_nb = [&](const unsigned* flags, unsigned target) {
        __syncthreads();
        if (wv == 0) {
            bool ok = true;
            if (nbw >= 0) {
                unsigned spins = 0;
                while (__hip_atomic_load(&flags[nbw * kFlagStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                    if ((++spins & 63u) == 0 &&
                        (__hip_atomic_load(&a.bar->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                         spins > kBarSpins)) {
                        __hip_atomic_store(&a.bar->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const uint64_t bad = __ballot(!ok);
            if (lane == 0) si[2] = bad ? 0 : 1;
        }
        __syncthreads();
    };
    auto wait_neighbours = [&](unsigned gstep) { wait_nb(a.bar->udone, gstep); };
    // w = D p(D) u_j for the systems of mk2 by the product form z_0 = u_j,
    // z_{i+1} = z_i - D z_i / theta_i, the last stage stored as w = u_j - z_npoly
    // (= D p(D) u_j).  Between two stages the slice publishes its rows of z_{i+1} and waits
    // for the neighbours whose rows its stencil reads (the same argument as for the update
    // flags keeps the ping-pong buffers safe: a work-group overwrites a stage's rows only
    // after its neighbours have published the next stage, which follows their reads of
    // them).  false when a wait timed out.
    unsigned zst = 0;   // stages published in this launch (the same count in every work-group)
    // With `split`, the caller has run phase 0 of the first sweep (on u_j) before its wait
    // for the neighbours' update flags, and every later sweep's phase 0 runs between the
    // publication of its input and the wait for the neighbours' copies of it.
    auto poly_stages = [&](uint64_t mk2) -> bool {
        const cplx<R>* zin = a.u;
        for (int i = 0; i < npoly; i++) {
            const bool fin = i == npoly - 1;
            cplx<R>* zout = fin ? a.w : zbuf(i + 1);
            const PolyEpi<R> pe{fin ? a.u : nullptr, nullptr, a.pc[i], czero<R>()};
            if (split) stencil_phase(std::integral_constant<int, 1>{}, zin, zout, mk2, pe);
            else stencil_e(zin, zout, mk2, pe);
            if (fin) break;
            ++zst;
            release_stores();
            __syncthreads();
            if (tid == 0)
                __hip_atomic_store(&a.bar->zdone[wg * kFlagStride], zst, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            // this work-group's rows of z_{i+1} are complete and visible to its waves
            if (split) stencil_phase(std::integral_constant<int, 0>{}, zout, nullptr, mk2, pe);
            wait_nb(a.bar->zdone, zst);
            if (!si[2]) return false;
            zin = zout;
        }
        return true;
    };
    // Waits (wave 0, a lane per system) until the owner of every system in `run` has
    // published step gstep's coefficients, then acquires them for the work-group.  Returns
    // the systems that continue (the owners' stop bits cleared), identical in every
    // work-group; si[2] = 0 when the kernel must end (a wait timed out somewhere).  Replaces
    // a grid barrier: a work-group waits for the owners, not for all work-groups.
    auto wait_coefs = [&](uint64_t run, unsigned gstep) -> uint64_t {
        __syncthreads();
        if (wv == 0) {
            const int q = lane;
            const bool need = (run >> q) & 1ull;
            unsigned v = 0, spins = 0;
            bool ok = true;
            if (need) {
                while (((v = __hip_atomic_load(&a.bar->cflag[q * kFlagStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 1) !=
                       gstep) {
                    if ((++spins & 63u) == 0 &&
                        (__hip_atomic_load(&a.bar->fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                         spins > kBarSpins)) {
                        __hip_atomic_store(&a.bar->fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const uint64_t stopm = __ballot(need && ok && (v & 1u));
            const uint64_t bad = __ballot(!ok);
            if (lane == 0) {
                *smask = run & ~stopm;
                si[2] = bad ? 0 : 1;
            }
        }
        __syncthreads();
        return *smask;
    };
    // ---- Givens rotation of column c = j-1 of the owned system (src/fgmres.cpp:99-115),
    // its rotated entries to Hrot and the break test |g_j| < tol ||eta||, by one thread
    // while the work-group's other waves already run the update pass
    auto givens = [&](int j) {
        if (si[0]) return;   // stopped by the previous column's break (maxIt = j - 1)
        const int q = sys, c = j - 1;
        cplx<double> carry = hcol[0];
        for (int i = 0; i < c; i++) {
            const cplx<double> h1 = hcol[i + 1], cn = csl[i], sn = snl[i];
            hcol[i] = cmul(cn, carry) + cmul(sn, h1);
            carry = cn * h1 - sn * carry;
        }
        const cplx<double> hs = hcol[j];
        const double den = sqrt(norm2(carry) + norm2(hs));
        const cplx<double> sn = den > 0 ? cdivr(hs, den) : czero<double>();
        const cplx<double> cn = den > 0 ? cdivr(carry, den) : mk<double>(1.0, 0.0);
        snl[c] = sn;
        csl[c] = cn;
        hcol[c] = cmul(cn, carry) + cmul(sn, hs);
        gls[c + 1] = -(sn * gls[c]);
        gls[c] = cmul(cn, gls[c]);
        cplx<double>* Ho = a.Hrot + (size_t)q * (m + 1) * m;
        for (int i = 0; i < j; i++) Ho[(size_t)i * m + c] = hcol[i];
        if (!si[1] && sqrt(norm2(gls[c + 1])) < a.tol * a.nrm[2 * q]) {
            a.ctrl[q].maxit = j;
            si[0] = 1;   // the next step's phase B stops the system
        }
    };
    // ---- phase C: q_j and u_{j+1} on the own rows (wave per system); the update
    // coefficients are read per chunk, next to the basis vectors they multiply
    auto update = [&](int j, uint64_t mk2) {
        for (int qb = wv; qb * SPI < a.nsys; qb += kNW) {
            const unsigned pm = pair_mask(mk2, qb);
            if (!pm) continue;
            const int q = sys_of(qb);
            const bool act = (pm >> h) & 1u;
            const cplx<R>* cf = a.coef + (size_t)q * CST;
            const R ib = cf[2 * m + 1].x;
            cplx<R> qj[RPL], un[RPL];
            load_rows(qj, a.u + (size_t)q * n + row0);
            load_rows(un, a.w + (size_t)q * n + row0);
#pragma unroll
            for (int k = 0; k < RPL; k++) un[k] = ib * un[k];
            // newest vectors first: the reduction pass just read the basis oldest-first, so
            // its most recent vectors are still in the last-level cache
            // chunks [i0, i0 + NC) from the newest vectors down: full chunks of kCU, the
            // oldest remainder in groups of 4 (indices below 0 masked)
            // coefficients of a chunk: lane li of each lane group loads those of vector i0 + li
            // (NC <= LPS), the group's lanes take them from there (cross-lane reads instead of
            // two loads per vector and lane)
            const int gbase = lane & ~(LPS - 1);
            auto chunk = [&](auto nc_, int i0) {
                constexpr int NC = decltype(nc_)::value;
                static_assert(NC <= LPS, "one coefficient per lane of the group");
                S qv[NC][RPL];
                const int icl = max(0, min(i0 + min(li, NC - 1), j - 1));
                const cplx<R> al = cf[icl], cl = cf[m + icl];
#pragma unroll
                for (int ii = 0; ii < NC; ii++) {
                    const int ic = max(0, min(i0 + ii, j - 1));
                    const S* qi = a.Q + (size_t)ic * vs + (size_t)q * n + row0 + lb;
#pragma unroll
                    for (int k = 0; k < RPL; k++) qv[ii][k] = qi[k];
                }
#pragma unroll
                for (int ii = 0; ii < NC; ii++) {
                    const bool in = i0 + ii >= 0 && i0 + ii < j;
                    const cplx<R> ab = mk<R>(__shfl(al.x, gbase + ii, 64), __shfl(al.y, gbase + ii, 64));
                    const cplx<R> cb = mk<R>(__shfl(cl.x, gbase + ii, 64), __shfl(cl.y, gbase + ii, 64));
                    const cplx<R> aa = in ? ab : czero<R>(), cc = in ? cb : czero<R>();
#pragma unroll
                    for (int k = 0; k < RPL; k++) {
                        const cplx<R> qq = widen(qv[ii][k]);
                        cmsub(qj[k], aa, qq);
                        cmsub(un[k], cc, qq);
                    }
                }
            };
            int i1 = j;
            for (; i1 >= kCU; i1 -= kCU) chunk(std::integral_constant<int, kCU>{}, i1 - kCU);
            // the oldest remainder as one chunk of the next multiple of 4 vectors (indices
            // below 0 masked)
            const int ng = (i1 + 3) / 4;
            if (ng == 1) chunk(std::integral_constant<int, 4>{}, i1 - 4);
            if constexpr (8 <= kCU)
                if (ng == 2) chunk(std::integral_constant<int, 8>{}, i1 - 8);
            if constexpr (12 <= kCU)
                if (ng == 3) chunk(std::integral_constant<int, 12>{}, i1 - 12);
            if constexpr (16 <= kCU)
                if (ng == 4) chunk(std::integral_constant<int, 16>{}, i1 - 16);
            static_assert(kCU <= 16, "remainder cases");
            // u_{j+1} is orthogonalised against the stored q_j
            const cplx<R> cj = cf[m + j];
            S* qo = a.Q + (size_t)j * vs + (size_t)q * n + row0 + lb;
            cplx<R>* uo = a.u + (size_t)q * n + row0 + lb;
            if (lrow && act) {
#pragma unroll
                for (int k = 0; k < RPL; k++) {
                    const cplx<R> qs = narrow(qo + k, ib * qj[k]);
                    uo[k] = un[k] - cj * qs;
                }
            }
        }
    };
    // ---- triangular solve of the owned system (src/fgmres.cpp:118-127), blocked:
    // kBB x kBB diagonal blocks in LDS on one wave, the rest as work-group updates
    auto backsub = [&]() {
        const int q = sys;
        const int nI = a.ctrl[q].maxit;
        const cplx<double>* Ho = a.Hrot + (size_t)q * (m + 1) * m;
        for (int b1 = nI; b1 > 0; b1 -= kBB) {
            const int b0 = max(0, b1 - kBB), nb = b1 - b0;
            for (int e = tid; e < nb * nb; e += kCT) {
                const int r = e / nb, c = e - r * nb;
                blk[r * kBB + c] = r <= c ? Ho[(size_t)(b0 + r) * m + (b0 + c)] : czero<double>();
            }
            __syncthreads();
            if (wv == 0) {
                cplx<double> rh = lane < nb ? gls[b0 + lane] : czero<double>();
                for (int i = nb - 1; i >= 0; i--) {
                    cplx<double> ri;
                    ri.x = __shfl(rh.x, i, 64);
                    ri.y = __shfl(rh.y, i, 64);
                    const cplx<double> e = cdiv(ri, blk[i * kBB + i]);
                    if (lane == i) rh = e;
                    if (lane < i) rh = rh - blk[lane * kBB + i] * e;
                }
                if (lane < nb) gls[b0 + lane] = rh;
            }
            __syncthreads();
            for (int k = tid; k < b0; k += kCT) {
                cplx<double> acc = czero<double>();
                const cplx<double>* row = Ho + (size_t)k * m + b0;
                for (int c = 0; c < nb; c++) acc = acc + row[c] * gls[b0 + c];
                gls[k] = gls[k] - acc;
            }
            __syncthreads();
        }
        cplx<R>* et = a.etals + (size_t)q * m;
        for (int i = tid; i < nI; i += kCT) et[i] = cast<R>(gls[i]);
    };
    // ---- x += sum_{i < maxit} eta_i q_i on the own rows (wave per system)
    auto xupdate = [&](uint64_t mk2) {
        for (int qb = wv; qb * SPI < a.nsys; qb += kNW) {
            const unsigned pm = pair_mask(mk2, qb);
            if (!pm) continue;
            const int q = sys_of(qb);
            const bool act = (pm >> h) & 1u;
            // the pair's longer basis bounds the loop; the other system's terms past its
            // own length are masked
            const int nIo = a.ctrl[q].maxit;
            int nI = nIo;
#pragma unroll
            for (int o = LPS; o < 64; o *= 2) nI = max(nI, __shfl_xor(nI, o, 64));
            const cplx<R>* et = a.etals + (size_t)q * m;
            cplx<R> xr[RPL];
            if (npoly)   // s = Q eta; x += p(D) s follows
#pragma unroll
                for (int k = 0; k < RPL; k++) xr[k] = czero<R>();
            else
                load_rows(xr, a.x + (size_t)q * n + row0);
            for (int i0 = 0; i0 < nI; i0 += kCU) {
                S qv[kCU][RPL];
                cplx<R> e[kCU];
#pragma unroll
                for (int ii = 0; ii < kCU; ii++) {
                    const int ic = max(0, min(i0 + ii, nIo - 1));
                    e[ii] = et[ic];
                    const S* qi = a.Q + (size_t)ic * vs + (size_t)q * n + row0 + lb;
#pragma unroll
                    for (int k = 0; k < RPL; k++) qv[ii][k] = qi[k];
                }
#pragma unroll
                for (int ii = 0; ii < kCU; ii++) {
                    const cplx<R> ee = i0 + ii < nIo ? e[ii] : czero<R>();
#pragma unroll
                    for (int k = 0; k < RPL; k++) cmadd(xr[k], ee, widen(qv[ii][k]));
                }
            }
            cplx<R>* xo = a.x + (size_t)q * n + row0 + lb;
            if (lrow && act) {
                if (npoly) {   // z_0 = s, x += s / theta_0
                    cplx<R>* zo = zb0 + (size_t)q * n + row0 + lb;
                    const cplx<R> c0 = a.pc[0];
#pragma unroll
                    for (int k = 0; k < RPL; k++) {
                        zo[k] = xr[k];
                        cplx<R> xv = xo[k];
                        cmadd(xv, c0, xr[k]);
                        xo[k] = xv;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < RPL; k++) xo[k] = xr[k];
                }
            }
        }
    };
    // ---- x0 = eta, r = eta - A x0, ||eta||, ||r||
    if (wg == 0 && tid == 0)
        for (int k = 0; k < 12; k++) a.stats[k] = 0;
    for (int q = wv; q < a.nsys; q += kNW) {
        if (!((a.mask >> q) & 1ull)) continue;
        for (int row = lane; row < nrow; row += 64) a.x[(size_t)q * n + row0 + row] = a.eta[(size_t)q * n + row0 + row];
    }
    // the slice's links into LDS, read by every stencil of the launch (the barrier below
    // orders these stores before the first read)
    if (a.lds_links) {
        const float4* src = reinterpret_cast<const float4*>(a.L + (size_t)s0 * 5 * DOF * DOF);
        float4* dst = reinterpret_cast<float4*>(smem_raw + coarsest_lds(m));
        const int nq = (int)((size_t)(s1 - s0) * 5 * DOF * DOF * sizeof(cplx<R>) / 16);
        for (int i = tid; i < nq; i += kCT) dst[i] = src[i];
    }
    if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;
    stencil(a.x, a.u, a.eta, a.mask, true);
    __syncthreads();
    reduce_norm(a.mask, 2 * m, a.u);
    reduce_norm(a.mask, 2 * m + 1, a.eta);
    const int nmask = __builtin_popcountll(a.mask);
    count(links_bytes + nmask * n * sR * 6, 0);
    if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;
    if (owner) {
        const double err = sqrt(total1(sys, 2 * m).x), nphi = sqrt(total1(sys, 2 * m + 1).x);
        if (tid == 0) {
            a.nrm[2 * sys] = nphi;
            a.nrm[2 * sys + 1] = err;
            a.ctrl[sys].maxit = m;
            a.ctrl[sys].iters = err == 0.0 ? 0 : a.restarts * m;
            set_flags(sys, err == 0.0 ? 0 : (F_LIVE | F_RUN));
        }
    }
    if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;

    unsigned gstep = 0;   // Arnoldi steps of this launch (the same in every work-group)
    for (int k = 0; k < a.restarts; k++) {
        const uint64_t live = sysmask(F_LIVE);
        if (!live) break;
        const int nl = __builtin_popcountll(live);
        if (tid == 0) si[0] = 0;   // no break pending (owner)
        for (int j = 0; j <= m; j++) {
            const uint64_t run = sysmask(F_RUN);
            if (!run) break;
            const int nr = __builtin_popcountll(run);
            const bool hw = j < m;
            // the rows of u_j the stencil reads were written by the neighbours' update of
            // the previous step (j = 0: before the last grid barrier); the hops inside the
            // slice read only this work-group's rows and may run before the wait
            if constexpr (POLY) {   // (cleared every step, so no value is carried across the passes below)
#pragma unroll
                for (int tt = 0; tt < DOF / 8; tt++) pacc[tt] = typename Mfma<R>::acc_t{0, 0, 0, 0};
            }
            if (hw && npoly && split)
                stencil_phase(std::integral_constant<int, 0>{}, a.u, nullptr, run, PolyEpi<R>{nullptr, nullptr, czero<R>(), czero<R>()});
            if (j > 0) {
                wait_neighbours(gstep);
                if (!si[2]) return;
            }
            if (hw) {
                if (npoly == 0) stencil(a.u, a.w, nullptr, run, false);
                else if (!poly_stages(run)) return;
            }
            __syncthreads();
            clk(7);
            const uint64_t tw0 = a.wgclk ? wall_clock64() : 0;
            for (int qb = wv; qb * SPI < a.nsys; qb += kNW) {
                const unsigned pm = pair_mask(run, qb);
                if (pm) reduce_step(qb, j, hw, pm);
            }
            if (a.wgclk && lane == 0) a.wgclk[wg * kClk + 9 + wv] += (double)(wall_clock64() - tw0);
            clk(0);
            // publish this slice's partial sums of the step; only the owners of running
            // systems wait for every work-group's (a grid barrier's worth of ordering for
            // them, none for the rest)
            ++gstep;
            release_stores();
            if (tid == 0) si[2] = 1;   // no timed-out wait in phase B (yet)
            __syncthreads();
            if (tid == 0) __hip_atomic_store(&a.bar->pdone[wg * kFlagStride], gstep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            clk(1);
            if (owner && ((run >> sys) & 1ull)) {
                phaseB(j, gstep);
                if (!si[2]) return;
            }
            clk(2);
            const uint64_t cont = wait_coefs(run, gstep);
            if (!si[2]) return;
            clk(3);
            if (tid == 0 && j >= 1 && owner && ((run >> sys) & 1ull)) givens(j);
            const uint64_t tw1 = a.wgclk ? wall_clock64() : 0;
            update(j, cont);
            if (a.wgclk && lane == 0) a.wgclk[wg * kClk + 9 + kNW + wv] += (double)(wall_clock64() - tw1);
            clk(4);
            const int nc = __builtin_popcountll(cont);
            // stencil: each sweep reads its input and writes its output (the last poly stage
            // also reads u_j)
            const double sweeps = npoly ? npoly : 1;
            count((hw ? sweeps * (sweep_links + 2.0 * nr * n * sR) + (npoly ? nr * n * sR : 0.0) : 0.0) +
                      nr * n * (sQ * j + sR * (hw ? 2 : 1)) + nc * n * (sQ * (j + 1) + sR * 3),
                  1);
            // publish this slice's update (every thread's stores precede thread 0's release)
            release_stores();
            __syncthreads();
            if (tid == 0) __hip_atomic_store(&a.bar->udone[wg * kFlagStride], gstep, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            clk(5);
        }
        if (owner && ((live >> sys) & 1ull)) backsub();
        if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;
        xupdate(live);
        // x += p(D) s: the stages z_{i+1} = z_i - D z_i / theta_i from z_0 = s, each added
        // to x with 1/theta_{i+1} (z_0's term was added by xupdate)
        for (int i = 0; i + 1 < npoly; i++) {
            if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;
            stencil_e(zbuf(i), zbuf(i + 1), live, PolyEpi<R>{nullptr, a.x, a.pc[i], a.pc[i + 1]});
            count(sweep_links + nl * n * sR * 4, 0);
        }
        if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;
        stencil(a.x, a.u, a.eta, live, true);
        __syncthreads();
        reduce_norm(live, 2 * m, a.u);
        count(sweep_links + nl * n * sR * 6, 0);
        if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;
        if (owner && ((live >> sys) & 1ull)) {
            const double err = sqrt(total1(sys, 2 * m).x);
            if (tid == 0) {
                a.nrm[2 * sys + 1] = err;
                int fl = F_LIVE | F_RUN;
                if (err < a.tol * a.nrm[2 * sys]) {
                    a.ctrl[sys].iters = k * m + a.ctrl[sys].maxit;
                    fl = 0;
                }
                set_flags(sys, fl);
            }
        }
        if (!grid_barrier(a.bar, bar_gen, wg, G, &si[3])) return;
        clk(6);
    }
}

// Clears the grid-barrier state before a launch (a kernel, not a memset, so that graph
// captures hold no memset node), after folding the previous launch's fail flag into the
// sticky word.
__global__ __launch_bounds__(256) void k_bar_reset(GridBar* __restrict__ b, unsigned* __restrict__ sticky) {
    __shared__ unsigned failed;
    if (threadIdx.x == 0) failed = b->fail;
    __syncthreads();
    if (threadIdx.x == 0 && failed) *sticky = 1u;
    unsigned* w = reinterpret_cast<unsigned*>(b);
    for (size_t i = threadIdx.x; i < sizeof(GridBar) / sizeof(unsigned); i += blockDim.x) w[i] = 0u;
}

template <typename R, typename S, int DOF, int LPS, bool POLY>
void launch_coarsest(const CArgs<R, S>& a, int G, size_t sh, hipStream_t s) {
    static bool attr = false;   // dynamic LDS above 64 KB (links cached in LDS)
    if (!attr) {
        DD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_coarsest<R, S, DOF, LPS, POLY>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    void* args[] = {const_cast<CArgs<R, S>*>(&a)};
    DD_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_coarsest<R, S, DOF, LPS, POLY>), dim3(G), dim3(kCT), args,
                                      (unsigned)sh, s));
}

// Work-groups that are certainly co-resident: the kernel is bounded to 512 threads
// (at most 256 VGPRs per lane, two waves per SIMD) and its LDS request (the scalar state,
// checked to be at most 64 KB, plus the slice's links when they are cached, at most
// 160 KB in all) fits one compute unit, so every CU holds one work-group.  0 when the
// device cannot launch cooperatively.
int coresident_workgroups() {
    int dev = 0, ncu = 0, coop = 0;
    DD_HIP(hipGetDevice(&dev));
    DD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    DD_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    return coop ? ncu : 0;
}


}  // namespace

template <typename R>
CoarsestGmres<R>::CoarsestGmres(hipStream_t s, int Nx, int Nt, int dof, int m, int restarts, double tol, int cap,
                                bool half_basis)
    : s_(s), Nx_(Nx), Nt_(Nt), dof_(dof), m_(m), restarts_(restarts), cap_(cap), m_max_(m), tol_(tol),
      half_(half_basis && sizeof(R) == 4) {
    if (m < 1 || m > kMaxM || restarts < 1 || cap < 1 || cap > 64) return;
    const int V = Nx * Nt;
    const size_t sh = coarsest_lds(m);
    if (sh > 64 * 1024) return;
    const int maxg = coresident_workgroups();
    if (dof != 8 && dof != 16 && dof != 24 && dof != 32) return;
    // slices of spw sites (a power of two, at most kSpwMax): the smallest that fits the
    // lattice into the co-resident grid, so that every compute unit streams the basis;
    // one work-group per system for the scalar state (G >= cap)
    spw_ = 1;
    while (spw_ < kSpwMax && (V + spw_ - 1) / spw_ > std::max(maxg, 1)) spw_ *= 2;
    G_ = (V + spw_ - 1) / spw_;
    if (std::getenv("DDAMG_VERBOSE"))
        fprintf(stderr, "coarsest persistent kernel: V %d dof %d cap %d m %d lds %zu co-resident %d -> %d x %d sites\n",
                V, dof, cap, m, sh, maxg, G_, spw_);
    if (G_ > maxg || G_ > kGmax || G_ < cap) return;
    const size_t n = (size_t)V * dof, vb = sizeof(cplx<R>) * n * cap;
    Q_.alloc((half_ ? sizeof(chalf) * n * cap : vb) * (m + 1));
    u_.alloc(vb);
    w_.alloc(vb);
    part_.alloc(sizeof(cplx<double>) * (size_t)cap * (2 * m + 2) * G_);
    Hrot_.alloc(sizeof(cplx<double>) * (size_t)cap * (m + 1) * m);
    etals_.alloc(sizeof(cplx<double>) * (size_t)cap * m);
    coef_.alloc(sizeof(cplx<double>) * (size_t)cap * (2 * m + 2));
    nrm_.alloc(sizeof(double) * (2 * cap + 12));
    ctrl_.alloc(sizeof(Ctrl) * cap);
    bar_.alloc(sizeof(GridBar));
    sticky_.alloc(sizeof(unsigned));
    ctrlh_.alloc(sizeof(Ctrl) * cap + sizeof(double) * 12 + sizeof(GridBar) + sizeof(unsigned));
    DD_HIP(hipMemsetAsync(bar_.as<void>(), 0, sizeof(GridBar), s_));
    DD_HIP(hipMemsetAsync(sticky_.as<void>(), 0, sizeof(unsigned), s_));
    DD_HIP(hipStreamSynchronize(s_));
    ok_ = true;
}

template <typename R>
void CoarsestGmres<R>::check_failure() {
    if (!ok_) return;
    // fold the last launch's flag in, then read the sticky word
    k_bar_reset<<<1, 256, 0, s_>>>(bar_.as<GridBar>(), sticky_.as<unsigned>());
    DD_HIP(hipGetLastError());
    unsigned* h = reinterpret_cast<unsigned*>(reinterpret_cast<unsigned char*>(ctrlh_.as<void>()) + sizeof(Ctrl) * cap_ +
                                              sizeof(double) * 12 + sizeof(GridBar));
    DD_HIP(hipMemcpyAsync(h, sticky_.as<void>(), sizeof(unsigned), hipMemcpyDeviceToHost, s_));
    DD_HIP(hipStreamSynchronize(s_));
    if (*h) {
        DD_HIP(hipMemsetAsync(sticky_.as<void>(), 0, sizeof(unsigned), s_));
        DD_HIP(hipStreamSynchronize(s_));
        fprintf(stderr, "coarsest GMRES: grid barrier timed out\n");
        throw HipFailure(hipErrorLaunchTimeOut);
    }
}

template <typename R>
void CoarsestGmres<R>::set_poly(const std::vector<std::complex<double>>& theta) {
    if (!ok_) return;
    if ((int)theta.size() > kPolyMax) throw BadShape("coarsest GMRES: polynomial degree above 32");
    npoly_ = (int)theta.size();
    if (!npoly_) return;
    if (!pc_.bytes()) {
        pc_.alloc(sizeof(cplx<R>) * kPolyMax);
        z_.alloc(sizeof(cplx<R>) * 2 * (size_t)Nx_ * Nt_ * dof_ * cap_);
    }
    std::vector<cplx<R>> h(kPolyMax, mk<R>(R(0), R(0)));
    for (int i = 0; i < npoly_; i++) {
        const std::complex<double> c = 1.0 / theta[i];
        h[i] = mk<R>((R)c.real(), (R)c.imag());
    }
    DD_HIP(hipMemcpyAsync(pc_.as<void>(), h.data(), sizeof(cplx<R>) * kPolyMax, hipMemcpyHostToDevice, s_));
    DD_HIP(hipStreamSynchronize(s_));
}

template <typename R>
void CoarsestGmres<R>::solve(const cplx<R>* L, double m0, const cplx<R>* eta, cplx<R>* psi, uint64_t mask,
                             int* iters) {
    if (!ok_) throw BadShape("coarsest GMRES: persistent kernel not available for this level");
    const int nsys = mask ? 64 - __builtin_clzll(mask) : 0;
    if (nsys == 0) return;
    if (nsys > cap_ || nsys > G_) throw BadShape("coarsest GMRES: batch larger than capacity");
    if constexpr (sizeof(R) == 4) {
        if (half_) {
            run<chalf>(L, m0, eta, psi, mask, nsys, iters);
            return;
        }
    }
    run<cplx<R>>(L, m0, eta, psi, mask, nsys, iters);
}

template <typename R>
template <typename S>
void CoarsestGmres<R>::run(const cplx<R>* L, double m0, const cplx<R>* eta, cplx<R>* psi, uint64_t mask, int nsys,
                           int* iters) {
    CArgs<R, S> a;
    a.L = L;
    a.eta = eta;
    a.x = psi;
    a.Q = Q_.as<S>();
    a.u = u_.as<cplx<R>>();
    a.w = w_.as<cplx<R>>();
    a.part = part_.as<cplx<R>>();
    a.Hrot = Hrot_.as<cplx<double>>();
    a.etals = etals_.as<cplx<R>>();
    a.coef = coef_.as<cplx<R>>();
    a.nrm = nrm_.as<double>();
    a.ctrl = ctrl_.as<Ctrl>();
    a.stats = nrm_.as<double>() + 2 * cap_;
    a.bar = bar_.as<GridBar>();
    static const bool verbose = std::getenv("DDAMG_VERBOSE") != nullptr;
    if (verbose && !deferred && !wgclk_.bytes()) {
        wgclk_.alloc(sizeof(double) * kClk * G_);
        DD_HIP(hipMemsetAsync(wgclk_.as<void>(), 0, sizeof(double) * kClk * G_, s_));
    }
    a.wgclk = (verbose && !deferred) ? wgclk_.as<double>() : nullptr;
    a.mask = mask;
    a.dmask = dev_mask();
    a.Nx = Nx_;
    a.Nt = Nt_;
    a.nsys = nsys;
    a.m = m_;
    a.restarts = restarts_;
    a.spw = spw_;
    a.m0 = (R)m0;
    a.tol = tol_;
    a.npoly = npoly_;
    a.pc = pc_.as<cplx<R>>();
    a.z = z_.as<cplx<R>>();
    a.zs = (size_t)Nx_ * Nt_ * dof_ * cap_;
    // FP32 slices of up to 4 sites keep their links in LDS when they fit beside the state
    const size_t lbytes = (size_t)spw_ * 5 * dof_ * dof_ * sizeof(cplx<R>);
    a.lds_links = (sizeof(R) == 4 && dof_ <= 24 && spw_ <= 4 && coarsest_lds(m_) + lbytes <= 160 * 1024) ? 1 : 0;
    const size_t sh = coarsest_lds(m_) + (a.lds_links ? lbytes : 0);
    // one cooperative grid in flight per process: two grids that each hold part of the
    // device could wait on each other's barriers (contexts solving from several threads)
    static std::mutex coop_mutex;
    std::lock_guard<std::mutex> lock(coop_mutex);
    k_bar_reset<<<1, 256, 0, s_>>>(a.bar, sticky_.as<unsigned>());
    DD_HIP(hipGetLastError());
    timer_begin(s_, K_COARSEST);
    switch (dof_) {
#define DD_CASE(D)                                                   \
    case D:                                                          \
        if (spw_ <= 4 && npoly_) launch_coarsest<R, S, D, 32, true>(a, G_, sh, s_);   \
        else if (spw_ <= 4) launch_coarsest<R, S, D, 32, false>(a, G_, sh, s_);  \
        else if (npoly_) launch_coarsest<R, S, D, 64, true>(a, G_, sh, s_);      \
        else launch_coarsest<R, S, D, 64, false>(a, G_, sh, s_);                 \
        break;
        DD_CASE(8) DD_CASE(16) DD_CASE(24) DD_CASE(32)
#undef DD_CASE
        default:
            throw BadShape("coarsest GMRES: unsupported dof");
    }
    timer_end(s_, K_COARSEST, 0.0);
    if (deferred) return;   // device work only (graph capture); failures via check_failure()
    Ctrl* ch = ctrlh_.as<Ctrl>();
    double* st = reinterpret_cast<double*>(ch + cap_);
    DD_HIP(hipMemcpyAsync(ch, a.ctrl, sizeof(Ctrl) * nsys, hipMemcpyDeviceToHost, s_));
    DD_HIP(hipMemcpyAsync(st, a.stats, sizeof(double) * 12, hipMemcpyDeviceToHost, s_));
    // only the fail word of the barrier state comes back
    unsigned* bfail = reinterpret_cast<unsigned*>(st + 12);
    DD_HIP(hipMemcpyAsync(bfail, &a.bar->fail, sizeof(unsigned), hipMemcpyDeviceToHost, s_));
    DD_HIP(hipStreamSynchronize(s_));
    if (*bfail) {
        fprintf(stderr, "coarsest GMRES: grid barrier timed out\n");
        throw HipFailure(hipErrorLaunchTimeOut);
    }
    if (verbose && wgclk_.bytes() && ++solves_ % 19 == 0) {   // spread of the phase times over the work-groups
        std::vector<double> h((size_t)kClk * G_);
        DD_HIP(hipMemcpy(h.data(), a.wgclk, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        const char* nm[9] = {"reduce", "pub-part", "scalar", "wait-cf", "update", "pub-upd", "cycle", "nbr+sten",
                             "gather"};
        for (int k = 0; k < 9; k++) {
            double mn = 1e300, mx = 0, sum = 0;
            int imx = 0;
            for (int g = 0; g < G_; g++) {
                const double v = h[(size_t)g * kClk + k];
                mn = std::min(mn, v);
                if (v > mx) {
                    mx = v;
                    imx = g;
                }
                sum += v;
            }
            fprintf(stderr, "coarsest wg clocks %-8s min %.3g mean %.3g max %.3g (wg %d)\n", nm[k], mn, sum / G_, mx,
                    imx);
        }
        // each wave's own pass times: mean over the work-groups per wave index, and the mean
        // spread (slowest minus fastest wave) within a work-group
        for (int pass = 0; pass < 2; pass++) {
            double spread = 0;
            std::vector<double> per(kNW, 0.0);
            for (int g = 0; g < G_; g++) {
                double mn = 1e300, mx = 0;
                for (int w = 0; w < kNW; w++) {
                    const double v = h[(size_t)g * kClk + 9 + pass * kNW + w];
                    per[w] += v / G_;
                    mn = std::min(mn, v);
                    mx = std::max(mx, v);
                }
                spread += (mx - mn) / G_;
            }
            fprintf(stderr, "coarsest wave clocks %s:", pass ? "update" : "reduce");
            for (int w = 0; w < kNW; w++) fprintf(stderr, " %.3g", per[w]);
            fprintf(stderr, " | mean spread in a work-group %.3g\n", spread);
        }
    }
    for (int k = 0; k < 9; k++) phase_clk[k] += st[2 + k];
    timer_add_bytes(K_COARSEST, st[0]);
    steps += (long long)st[1];
    for (int q = 0; q < nsys; q++)
        if ((mask >> q) & 1ull) iters[q] = ch[q].iters;
}

template class CoarsestGmres<float>;
template class CoarsestGmres<double>;

}  // namespace dd
