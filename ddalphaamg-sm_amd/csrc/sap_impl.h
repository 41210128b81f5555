// SAP smoother block solves (gfx950): level-0 wave kernel and the batched coarse-level
// work-group kernel.  Reference behaviour restated (paths relative to the reference
// root): SAP_C::I_D_B_1_It src/sap.cpp:48-77, the block GMRES (FGMRES::fgmres with
// SAP_C::GMRES_D_B hooks) src/fgmres.cpp:5-127 and include/sap.h:14-83, and the
// block operator Level::SAP_level_l::D_local src/level.cpp:607-653.
#pragma once
#include <type_traits>
#include <vector>

#include "kernels.h"
#include "lanes.h"

namespace dd {

#define DD_LAUNCH_CHECK() DD_HIP(hipGetLastError())
#define DD_DOF_LIST(X) X(8) X(16) X(24) X(32)

static inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
static inline double nact(uint64_t mask, int nrhs) {
    return (double)__builtin_popcountll(mask & ((nrhs >= 64) ? ~0ull : ((1ull << nrhs) - 1)));
}

// ======================================================================================
// Block GMRES of the SAP smoother: FGMRES::fgmres with the identity preconditioner,
// initial guess = block right-hand side, restart length M, the break on
// |g_{j+1}| < tol ||phi||, maxIt kept across restarts and the true-residual test after
// each cycle (src/fgmres.cpp:5-97, src/sap.cpp:65).  Every thread of a block holds one
// element of each Krylov vector and the (uniform) Hessenberg scalars.
// ======================================================================================
template <int M>
struct Tri {
    static constexpr int at(int i, int j) { return j * (j + 1) / 2 + i; }   // i <= j
    static constexpr int size = M * (M + 1) / 2;
};

template <typename R, int M, typename E, typename Op, typename Red>
__device__ __forceinline__ E block_gmres(const E& phi, Op&& A, Red&& red, int restarts, R tol) {
    E x = phi;
    E r = phi - A(x);
    const R nphi = sqrt(red(phi.norm2()).x);
    R err = sqrt(red(r.norm2()).x);
    int maxIt = M;
    E V[M + 1];
#pragma unroll
    for (int i = 0; i <= M; i++) V[i] = E::zero();
    cplx<R> H[Tri<M>::size];
    cplx<R> g[M + 1], cn[M];
    R sn[M];   // the rotations' sines are real: h_{j+1,j} / den with h_{j+1,j} a norm
#pragma unroll
    for (int i = 0; i < M; i++) {
        cn[i] = czero<R>();
        sn[i] = R(0);
    }
    for (int k = 0; k < restarts; k++) {
        V[0] = (R(1) / err) * r;
        g[0] = mk<R>(err, 0);
#pragma unroll
        for (int j = 0; j < M; j++) {
            E w = A(V[j]);
#pragma unroll
            for (int i = 0; i <= j; i++) {
                const cplx<R> h = red(w.dotc(V[i]));
                H[Tri<M>::at(i, j)] = h;
                w = w - h * V[i];
            }
            const R hn = sqrt(red(w.norm2()).x);
            if (j + 1 < M && hn > R(0)) V[j + 1] = (R(1) / hn) * w;   // V_M is never read
            // Givens rotations of column j (src/fgmres.cpp:99-115)
#pragma unroll
            for (int i = 0; i < j; i++) {
                const cplx<R> a = H[Tri<M>::at(i, j)], bb = H[Tri<M>::at(i + 1, j)];
                H[Tri<M>::at(i, j)] = cmul(cn[i], a) + sn[i] * bb;
                H[Tri<M>::at(i + 1, j)] = cn[i] * bb - sn[i] * a;
            }
            const cplx<R> hjj = H[Tri<M>::at(j, j)];
            const R iden = R(1) / sqrt(norm2(hjj) + hn * hn);   // one division per rotation
            sn[j] = hn * iden;
            cn[j] = mk<R>(hjj.x * iden, hjj.y * iden);
            H[Tri<M>::at(j, j)] = cmul(cn[j], hjj) + mk<R>(sn[j] * hn, R(0));
            g[j + 1] = -(sn[j] * g[j]);
            g[j] = cmul(cn[j], g[j]);
            if (sqrt(norm2(g[j + 1])) < tol * nphi) {
                maxIt = j + 1;
                break;
            }
        }
        // back substitution (src/fgmres.cpp:118-127) and x += sum eta_j Z_j
        cplx<R> eta[M];
#pragma unroll
        for (int i = M - 1; i >= 0; i--) {
            if (i < maxIt) {
                cplx<R> e = g[i];
#pragma unroll
                for (int jj = i + 1; jj < M; jj++)
                    if (jj < maxIt) e = e - H[Tri<M>::at(i, jj)] * eta[jj];
                eta[i] = cdiv_recip(e, H[Tri<M>::at(i, i)]);
            }
        }
#pragma unroll
        for (int j = 0; j < M; j++)
            if (j < maxIt) x = x + eta[j] * V[j];
        r = phi - A(x);
        err = sqrt(red(r.norm2()).x);
        if (err < tol * nphi) break;
    }
    return x;
}

// Colour-c block number k -> block coordinates (red iff BX + BT even).
__device__ __forceinline__ void colour_block(int k, int colour, int nbx, int nbt, int& BX, int& BT) {
    if ((nbt & 1) == 0) {
        const int h = nbt >> 1;
        BX = k / h;
        BT = 2 * (k - BX * h) + ((BX + colour) & 1);
    } else {
        const int bidx = 2 * k + colour;
        BX = bidx / nbt;
        BT = bidx - BX * nbt;
    }
}

}  // namespace dd
#include "sapm.h"
namespace dd {

// ---- level 0: one site (two spin components) per lane; S = bx*bt lanes per block,
// 64/S blocks per wave; neighbours through cross-lane permutes, sums through
// butterflies inside the S-lane segment.  No LDS and no barriers.
template <typename R>
struct Spin2 {
    cplx<R> a, b;
    __device__ static Spin2 zero() { return {czero<R>(), czero<R>()}; }
    __device__ Spin2 operator+(const Spin2& o) const { return {a + o.a, b + o.b}; }
    __device__ Spin2 operator-(const Spin2& o) const { return {a - o.a, b - o.b}; }
    __device__ cplx<R> dotc(const Spin2& o) const { return mulc(a, o.a) + mulc(b, o.b); }
    __device__ cplx<R> norm2() const { return mk<R>(dd::norm2(a) + dd::norm2(b), R(0)); }
};
template <typename R>
__device__ __forceinline__ Spin2<R> operator*(cplx<R> s, const Spin2<R>& v) { return {s * v.a, s * v.b}; }
template <typename R>
__device__ __forceinline__ Spin2<R> operator*(R s, const Spin2<R>& v) { return {s * v.a, s * v.b}; }

template <typename R>
__device__ __forceinline__ cplx<R> shfl_c(cplx<R> v, int src) {
    return mk<R>(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

template <typename R>
__device__ __forceinline__ cplx<R> row_shift_c(cplx<R> v, int d) {   // d: +1, -1, +4, -4
    switch (d) {
        case 1: return mk<R>(row_shift<1>(v.x), row_shift<1>(v.y));
        case -1: return mk<R>(row_shift<-1>(v.x), row_shift<-1>(v.y));
        case 4: return mk<R>(row_shift<4>(v.x), row_shift<4>(v.y));
        default: return mk<R>(row_shift<-4>(v.x), row_shift<-4>(v.y));
    }
}

// B44: blocks of 4 x 4 sites, one per 16-lane row: the neighbours inside the block are DPP
// row shifts by 1 (t) and 4 (x) and the block sums a DPP butterfly over the row, in the
// order of wave_sum (the same bits); otherwise LDS permutes for any block shape.
// FR: rr is the SAP's right-hand side v and the block right-hand side v - D x is formed
// here at the block's sites (the operator of k_wilson: periodic in x, antiperiodic in t);
// the neighbours it reads belong to blocks of the other colour, which no work-group of
// this launch changes.
template <typename R, int M, bool B44, bool FR>
__global__ __launch_bounds__(64) void k_sap0(const cplx<R>* __restrict__ U, const cplx<R>* __restrict__ rr,
                                             cplx<R>* __restrict__ xx, int Nx, int Nt, int bx, int bt, int colour,
                                             R m0, int restarts, R tol, uint64_t mask, size_t ld,
                                             const int* __restrict__ gate, int nblk_colour) {
    const int rh = blockIdx.y;
    if (!rhs_on(mask, rh)) return;
    if (gate && gate[rh] == 0) return;
    const int S = bx * bt;
    const int lane = threadIdx.x;
    const int seg = lane / S, k = lane - seg * S;
    const int blk = blockIdx.x * (64 / S) + seg;
    if (blk >= nblk_colour) return;
    const int nbx = Nx / bx, nbt = Nt / bt;
    int BX, BT;
    colour_block(blk, colour, nbx, nbt, BX, BT);
    const int lx = k / bt, lt = k - lx * bt;
    const int X = BX * bx + lx, T = BT * bt + lt;
    const int n = X * Nt + T;
    const int base = seg * S;
    const bool hasTp = lt < bt - 1, hasTm = lt > 0, hasXp = lx < bx - 1, hasXm = lx > 0;
    const int sTp = hasTp ? lane + 1 : lane, sTm = hasTm ? lane - 1 : lane;
    const int sXp = hasXp ? lane + bt : lane, sXm = hasXm ? lane - bt : lane;
    // links of the hops that leave the block are zero, so the neighbour values shifted in
    // from outside the block (finite: other sites of the wave, or 0) drop out unmasked
    const cplx<R> U0 = hasTp ? U[2 * n] : czero<R>(), U1 = hasXp ? U[2 * n + 1] : czero<R>();
    const cplx<R> U0m = hasTm ? conj(U[2 * (n - 1)]) : czero<R>();
    const cplx<R> U1m = hasXm ? conj(U[2 * (n - Nt) + 1]) : czero<R>();
    const R mass = m0 + R(2);
    (void)base;
    // D_B restricted to the block, zero outside, no boundary sign
    auto A = [&](const Spin2<R>& v) -> Spin2<R> {
        // (1 -+ sigma_mu) of each hop formed on the own site before the lane move, one
        // complex value per hop: t+ a - b, x+ a + i b, t- a + b, x- a - i b (spin 0's
        // rows; spin 1's rows of the same projectors are -1, -i, 1 and i times them)
        const cplx<R> pt = v.a - v.b, px = v.a + times_i(v.b), mt = v.a + v.b, mx = v.a - times_i(v.b);
        cplx<R> dtp, dxp, dtm, dxm;
        if constexpr (B44) {
            dtp = row_shift_c(pt, 1);
            dxp = row_shift_c(px, 4);
            dtm = row_shift_c(mt, -1);
            dxm = row_shift_c(mx, -4);
        } else {
            dtp = shfl_c(pt, sTp);
            dxp = shfl_c(px, sXp);
            dtm = shfl_c(mt, sTm);
            dxm = shfl_c(mx, sXm);
        }
        const cplx<R> Tp = U0 * dtp, Xp = U1 * dxp, Tm = U0m * dtm, Xm = U1m * dxm;
        const cplx<R> a0 = Tp + Xp + Tm + Xm;
        const cplx<R> a1 = Tm - Tp + times_i(Xm) - times_i(Xp);
        return {mass * v.a - R(0.5) * a0, mass * v.b - R(0.5) * a1};
    };
    auto red = [&](cplx<R> v) -> cplx<R> {
        if constexpr (B44) return mk<R>(seg16_sum(v.x, lane), seg16_sum(v.y, lane));
        else return wave_sum(v, S);
    };
    const cplx<R>* rv = rr + rh * ld;
    Spin2<R> phi = {rv[2 * n], rv[2 * n + 1]};
    if constexpr (FR) {   // phi = v - D x at this site
        const cplx<R>* xs = xx + rh * ld;
        const int ntp = (T == Nt - 1) ? n - (Nt - 1) : n + 1, ntm = (T == 0) ? n + (Nt - 1) : n - 1;
        const int nxp = (X == Nx - 1) ? T : n + Nt, nxm = (X == 0) ? (Nx - 1) * Nt + T : n - Nt;
        const R sp = (T == Nt - 1) ? R(-1) : R(1), sm = (T == 0) ? R(-1) : R(1);
        const cplx<R> W0 = U[2 * n], W1 = U[2 * n + 1], W0m = conj(U[2 * ntm]), W1m = conj(U[2 * nxm + 1]);
        const cplx<R> p0 = xs[2 * n], p1 = xs[2 * n + 1];
        const cplx<R> tp0 = xs[2 * ntp], tp1 = xs[2 * ntp + 1], xp0 = xs[2 * nxp], xp1 = xs[2 * nxp + 1];
        const cplx<R> tm0 = xs[2 * ntm], tm1 = xs[2 * ntm + 1], xm0 = xs[2 * nxm], xm1 = xs[2 * nxm + 1];
        const cplx<R> a0 = W0 * (sp * (tp0 - tp1)) + W1 * (xp0 + times_i(xp1)) + W0m * (sm * (tm0 + tm1)) +
                           W1m * (xm0 - times_i(xm1));
        const cplx<R> a1 = W0 * (sp * (tp1 - tp0)) + W1 * (xp1 - times_i(xp0)) + W0m * (sm * (tm0 + tm1)) +
                           W1m * (times_i(xm0) + xm1);
        phi.a = phi.a - (mass * p0 - R(0.5) * a0);
        phi.b = phi.b - (mass * p1 - R(0.5) * a1);
    }
    const Spin2<R> e = block_gmres<R, M>(phi, A, red, restarts, tol);
    cplx<R>* xv = xx + rh * ld;
    xv[2 * n] = xv[2 * n] + e.a;
    xv[2 * n + 1] = xv[2 * n + 1] + e.b;
}

template <typename R>
void sap_blocks0(hipStream_t s, const cplx<R>* U, const cplx<R>* r, cplx<R>* x, int Nx, int Nt, int bx, int bt,
                 int colour, double m0, int m, int restarts, double tol, int nrhs, uint64_t mask, size_t ld,
                 const int* gate, bool form_residual) {
    timer_begin(s, K_SAP0);
    const int S = bx * bt;
    if (S > 64 || (S & (S - 1)) != 0) throw BadShape("level-0 SAP block must have a power-of-two site count <= 64");
    if (m != kSapM) throw BadShape("SAP block GMRES restart length must be 5");
    const int nblk = (Nx / bx) * (Nt / bt) / 2;
    if (form_residual && (((Nx / bx) & 1) || ((Nt / bt) & 1)))
        throw BadShape("level-0 SAP: forming the residual in the block kernel needs even block counts");
    dim3 grid(cdivu(nblk, 64 / S), nrhs);
    auto go = [&](auto b44, auto fr) {
        k_sap0<R, kSapM, decltype(b44)::value, decltype(fr)::value><<<grid, 64, 0, s>>>(
            U, r, x, Nx, Nt, bx, bt, colour, (R)m0, restarts, (R)tol, mask, ld, gate, nblk);
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (bx == 4 && bt == 4) {
        if (form_residual) go(T_{}, T_{});
        else go(T_{}, F_{});
    } else {
        if (form_residual) go(F_{}, T_{});
        else go(F_{}, F_{});
    }
    // links once, x (and with form_residual its neighbours and v) and the residual per system
    timer_end(s, K_SAP0, 0.5 * Nx * Nt * 2 * sizeof(cplx<R>) * (1.0 + (form_residual ? 4.0 : 3.0) * nact(mask, nrhs)));
    DD_LAUNCH_CHECK();
}

// Batched work-group sum: every thread contributes NQ complex values; every thread
// receives the NQ totals (same bits in every thread).  `red` holds 2*NQ*8 R.
template <typename R, int NQ>
__device__ __forceinline__ void bsum(cplx<R> (&v)[NQ], R* red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; q++) v[q] = wave_sum(v[q], 64);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            red[(wid * NQ + q) * 2] = v[q].x;
            red[(wid * NQ + q) * 2 + 1] = v[q].y;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        cplx<R> t = czero<R>();
        for (int w = 0; w < nw; w++) {
            t.x += red[(w * NQ + q) * 2];
            t.y += red[(w * NQ + q) * 2 + 1];
        }
        v[q] = t;
    }
}

// total q of the last bsum<R, NQ>, read back from the LDS partials (valid until the
// next bsum); avoids indexing the register array with a thread-dependent index
template <typename R, int NQ>
__device__ __forceinline__ cplx<R> bsum_total(const R* red, int q) {
    const int nw = blockDim.x >> 6;
    cplx<R> t = czero<R>();
    for (int w = 0; w < nw; w++) {
        t.x += red[(w * NQ + q) * 2];
        t.y += red[(w * NQ + q) * 2 + 1];
    }
    return t;
}

// Coarse-level SAP block solves for NB systems at once.  One work-group per block of
// the colour and group of NB systems; one thread per (site, row).  Each system runs
// the reference's block GMRES (src/fgmres.cpp:5-97 with the GMRES_D_B hooks,
// include/sap.h:14-83) with its own breaks, restarts and convergence; the block
// operator (Level::SAP_level_l::D_local, src/level.cpp:607-653) streams each link
// element once per iteration for all NB systems.  Registers hold the Krylov basis;
// right-hand sides, iterates and the per-system Hessenberg scalars live in LDS (the
// scalars of system q are updated by thread q); per-system flags are uniform bitmasks.
template <typename R, int DOF, int M, int NB, bool FULL>
__global__ __launch_bounds__(512) void k_sapc(const cplx<R>* __restrict__ L, const cplx<R>* __restrict__ rr,
                                              cplx<R>* __restrict__ xx, int Nx, int Nt, int bx, int bt, int colour,
                                              R m0, int restarts, R tol, uint64_t mask, size_t ld,
                                              const int* __restrict__ gate, int nrhs) {
    extern __shared__ unsigned char smem_raw[];
    constexpr int TS = Tri<M>::size;
    const int S = bx * bt, SD = S * DOF;
    cplx<R>* sv = reinterpret_cast<cplx<R>*>(smem_raw);   // [NB][SD] matvec staging
    cplx<R>* sphi = sv + NB * SD;                         // [NB][SD] block right-hand sides
    cplx<R>* sx = sphi + NB * SD;                         // [NB][SD] block iterates
    cplx<R>* Hs = sx + NB * SD;                           // [NB][TS]
    cplx<R>* gs = Hs + NB * TS;                           // [NB][M+1]
    cplx<R>* cns = gs + NB * (M + 1);                     // [NB][M]
    cplx<R>* sns = cns + NB * M;                          // [NB][M]
    cplx<R>* etas = sns + NB * M;                         // [NB][M]
    R* red = reinterpret_cast<R*>(etas + NB * M);         // [8][NB][2]
    R* nphis = red + 8 * NB * 2;                          // [NB]
    R* errs = nphis + NB;                                 // [NB]
    int* maxIts = reinterpret_cast<int*>(errs + NB);      // [NB]
    int* brk = maxIts + NB;                               // [1] bitmask of systems that broke
    const int tid = threadIdx.x;
    const bool act = tid < SD;
    const int k = act ? tid / DOF : 0, row = act ? tid - (tid / DOF) * DOF : 0;
    const int nbx = Nx / bx, nbt = Nt / bt;
    int BX, BT;
    colour_block(blockIdx.x, colour, nbx, nbt, BX, BT);
    const int lx = k / bt, lt = k - lx * bt;
    const int n = (BX * bx + lx) * Nt + BT * bt + lt;
    const bool hasTp = lt < bt - 1, hasTm = lt > 0, hasXp = lx < bx - 1, hasXm = lx > 0;
    const cplx<R>* Ln = L + (size_t)n * 5 * DOF * DOF + row;
    const R mass = m0 + R(2);
    const int r0 = blockIdx.y * NB;
    unsigned onm = 0;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const int rh = r0 + q;
        if (rh < nrhs && rhs_on(mask, rh) && (!gate || gate[rh] != 0)) onm |= 1u << q;
    }
    if (!onm) return;

    // v <- D_B v for all NB systems in place (zero outside the block, no boundary sign)
    auto A = [&](cplx<R> (&v)[NB]) {
        __syncthreads();
        if (act) {
#pragma unroll
            for (int q = 0; q < NB; q++) sv[q * SD + tid] = v[q];
        }
        __syncthreads();
        cplx<R> acc[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) acc[q] = czero<R>();
        if (act) {
            auto hop = [&](const cplx<R>* Lk, const cplx<R>* src) {
                if constexpr (FULL) {
                    // all DOF link loads of the hop issued before the first use
                    cplx<R> l[DOF];
#pragma unroll
                    for (int col = 0; col < DOF; col++) l[col] = Lk[col * DOF];
#pragma unroll
                    for (int col = 0; col < DOF; col++)
#pragma unroll
                        for (int q = 0; q < NB; q++) acc[q] = acc[q] + l[col] * src[q * SD + col];
                } else {
#pragma unroll 4
                    for (int col = 0; col < DOF; col++) {
                        const cplx<R> l = Lk[col * DOF];
#pragma unroll
                        for (int q = 0; q < NB; q++) acc[q] = acc[q] + l * src[q * SD + col];
                    }
                }
            };
            hop(Ln, sv + k * DOF);
            if (hasTp) hop(Ln + 1 * DOF * DOF, sv + (k + 1) * DOF);
            if (hasXp) hop(Ln + 2 * DOF * DOF, sv + (k + bt) * DOF);
            if (hasTm) hop(Ln + 3 * DOF * DOF, sv + (k - 1) * DOF);
            if (hasXm) hop(Ln + 4 * DOF * DOF, sv + (k - bt) * DOF);
        }
#pragma unroll
        for (int q = 0; q < NB; q++) v[q] = mass * v[q] - acc[q];
    };

    cplx<R> w[NB], V[M + 1][NB];
    const cplx<R>* rv = rr + (size_t)r0 * ld + (size_t)n * DOF + row;
#pragma unroll
    for (int q = 0; q < NB; q++) {
        const cplx<R> ph = (act && ((onm >> q) & 1)) ? rv[q * ld] : czero<R>();
        if (act) {
            sphi[q * SD + tid] = ph;
            sx[q * SD + tid] = ph;   // initial guess = block right-hand side
        }
        w[q] = ph;
#pragma unroll
        for (int i = 0; i <= M; i++) V[i][q] = czero<R>();
    }
    // r = phi - D_B x, kept unscaled in V_0 until the cycle starts
    A(w);
    {
        cplx<R> t[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            const cplx<R> ph = act ? sphi[q * SD + tid] : czero<R>();
            V[0][q] = ph - w[q];
            t[q] = mk<R>(norm2(ph), R(0));
        }
        bsum<R, NB>(t, red);
        if (tid < NB) nphis[tid] = sqrt(bsum_total<R, NB>(red, tid).x);
#pragma unroll
        for (int q = 0; q < NB; q++) t[q] = mk<R>(norm2(V[0][q]), R(0));
        bsum<R, NB>(t, red);
        if (tid < NB) {
            errs[tid] = sqrt(bsum_total<R, NB>(red, tid).x);
            maxIts[tid] = M;
        }
    }
    unsigned livem = onm;
    for (int kk = 0; kk < restarts && livem; kk++) {
        __syncthreads();   // errs of the last residual visible to all
        if (kk == 0) {     // a zero block right-hand side has the solution zero: no iterations
#pragma unroll
            for (int q = 0; q < NB; q++)
                if (nphis[q] == R(0)) livem &= ~(1u << q);
            if (!livem) break;
        }
        unsigned runm = livem;
#pragma unroll
        for (int q = 0; q < NB; q++)
            if ((livem >> q) & 1) V[0][q] = (R(1) / errs[q]) * V[0][q];
        if (tid < NB && ((livem >> tid) & 1)) gs[tid * (M + 1)] = mk<R>(errs[tid], R(0));
#pragma unroll
        for (int j = 0; j < M; j++) {
            if (!runm) break;
#pragma unroll
            for (int q = 0; q < NB; q++) w[q] = V[j][q];
            A(w);
#pragma unroll
            for (int i = 0; i <= j; i++) {
                cplx<R> h[NB];
#pragma unroll
                for (int q = 0; q < NB; q++) h[q] = mulc(w[q], V[i][q]);
                bsum<R, NB>(h, red);
                if (tid < NB && ((runm >> tid) & 1)) Hs[tid * TS + Tri<M>::at(i, j)] = bsum_total<R, NB>(red, tid);
#pragma unroll
                for (int q = 0; q < NB; q++)
                    if ((runm >> q) & 1) w[q] = w[q] - h[q] * V[i][q];
            }
            cplx<R> hn[NB];
#pragma unroll
            for (int q = 0; q < NB; q++) hn[q] = mk<R>(norm2(w[q]), R(0));
            bsum<R, NB>(hn, red);
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const R h = sqrt(hn[q].x);
                if (j + 1 <= M && ((runm >> q) & 1) && h > R(0)) V[j + 1][q] = (R(1) / h) * w[q];
            }
            if (tid == 0) *brk = 0;
            __syncthreads();
            if (tid < NB && ((runm >> tid) & 1)) {
                // Givens rotations of column j (src/fgmres.cpp:99-115) for system q = tid
                const int q = tid;
                cplx<R>* H = Hs + q * TS;
                cplx<R>* cn = cns + q * M;
                cplx<R>* sn = sns + q * M;
                cplx<R>* g = gs + q * (M + 1);
                for (int i = 0; i < j; i++) {
                    const cplx<R> a = H[Tri<M>::at(i, j)], bb = H[Tri<M>::at(i + 1, j)];
                    H[Tri<M>::at(i, j)] = cmul(cn[i], a) + cmul(sn[i], bb);
                    H[Tri<M>::at(i + 1, j)] = cn[i] * bb - sn[i] * a;
                }
                const cplx<R> hjj = H[Tri<M>::at(j, j)];
                const cplx<R> hs = mk<R>(sqrt(bsum_total<R, NB>(red, q).x), R(0));
                const R den = sqrt(norm2(hjj) + norm2(hs));
                sn[j] = mk<R>(hs.x / den, hs.y / den);
                cn[j] = mk<R>(hjj.x / den, hjj.y / den);
                H[Tri<M>::at(j, j)] = cmul(cn[j], hjj) + cmul(sn[j], hs);
                g[j + 1] = -(sn[j] * g[j]);
                g[j] = cmul(cn[j], g[j]);
                if (sqrt(norm2(g[j + 1])) < tol * nphis[q]) {
                    maxIts[q] = j + 1;
                    atomicOr(brk, 1 << q);
                }
            }
            __syncthreads();
            runm &= ~(unsigned)(*brk);
        }
        // back substitution (src/fgmres.cpp:118-127) by thread q, then x += sum eta_j V_j
        if (tid < NB && ((livem >> tid) & 1)) {
            const int q = tid, nI = maxIts[q];
            const cplx<R>* H = Hs + q * TS;
            const cplx<R>* g = gs + q * (M + 1);
            cplx<R>* eta = etas + q * M;
            for (int i = nI - 1; i >= 0; i--) {
                cplx<R> e = g[i];
                for (int jj = i + 1; jj < nI; jj++) e = e - H[Tri<M>::at(i, jj)] * eta[jj];
                eta[i] = cdiv(e, H[Tri<M>::at(i, i)]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NB; q++) {
            cplx<R> xv = act ? sx[q * SD + tid] : czero<R>();
            if ((livem >> q) & 1) {
                const int nI = maxIts[q];
#pragma unroll
                for (int j = 0; j < M; j++)
                    if (j < nI) xv = xv + etas[q * M + j] * V[j][q];
                if (act) sx[q * SD + tid] = xv;
            }
            w[q] = xv;
        }
        A(w);
        cplx<R> t[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) {
            V[0][q] = (act ? sphi[q * SD + tid] : czero<R>()) - w[q];
            t[q] = mk<R>(norm2(V[0][q]), R(0));
        }
        bsum<R, NB>(t, red);
        unsigned conv = 0;
#pragma unroll
        for (int q = 0; q < NB; q++)
            if (((livem >> q) & 1) && sqrt(t[q].x) < tol * nphis[q]) conv |= 1u << q;
        if (tid < NB && ((livem >> tid) & 1)) errs[tid] = sqrt(bsum_total<R, NB>(red, tid).x);
        livem &= ~conv;
    }
    if (act) {
        cplx<R>* xv = xx + (size_t)r0 * ld + (size_t)n * DOF + row;
#pragma unroll
        for (int q = 0; q < NB; q++)
            if ((onm >> q) & 1) xv[q * ld] = xv[q * ld] + sx[q * SD + tid];
    }
}

template <typename R, int DOF, int NB, bool FULL>
static void launch_sapc(hipStream_t s, const cplx<R>* L, const cplx<R>* r, cplx<R>* x, int Nx, int Nt, int bx, int bt,
                        int colour, double m0, int restarts, double tol, int nrhs, uint64_t mask, size_t ld,
                        const int* gate) {
    constexpr int M = kSapM;
    const int S = bx * bt;
    const int threads = ((S * DOF + 63) / 64) * 64;
    const int nblk = (Nx / bx) * (Nt / bt) / 2;
    dim3 grid(nblk, (nrhs + NB - 1) / NB);
    const size_t sh = sizeof(cplx<R>) * ((size_t)3 * NB * S * DOF + NB * (Tri<M>::size + (M + 1) + 3 * M)) +
                      sizeof(R) * (8 * NB * 2 + 2 * NB) + sizeof(int) * (NB + 1);
    k_sapc<R, DOF, M, NB, FULL><<<grid, threads, sh, s>>>(L, r, x, Nx, Nt, bx, bt, colour, (R)m0, restarts,
                                                           (R)tol, mask, ld, gate, nrhs);
}

// Kernel variants of the coarse SAP: systems per work-group (NB) and whether the link
// loads of a hop are issued together.  variant = 2*NB + FULL; sapc_variants lists the
// ones compiled for precision R (the autotuner in solver.hip times them).
// Variant 32 (FP32 only) is the matrix-core kernel k_sapm (16 systems per work-group,
// sapm.h), available for dof <= 24 and blocks of at most 16 sites; variant 33 is the same
// kernel reading a binary16 copy of the links.
template <typename R>
std::vector<int> sapc_variants() {
    if (sizeof(R) == 4) return {2 * 1 + 0, 2 * 1 + 1, 2 * 2 + 1, 2 * 4 + 0, 2 * 4 + 1, 2 * 8 + 0, 2 * 8 + 1, 32, 33};
    return {2 * 1 + 0, 2 * 1 + 1, 2 * 2 + 1, 2 * 4 + 0, 2 * 4 + 1};
}

template <typename R>
bool sapc_supported(int variant, int dof, int S) {
    if (variant == 32 || variant == 33) return sizeof(R) == 4 && dof <= 24 && S <= 16;
    return S * dof <= 512;
}

template <typename R>
int sapc_default_variant(int nrhs, int dof, int S, bool half_links) {
    if (sapc_supported<R>(32, dof, S)) return half_links ? 33 : 32;
    if (nrhs <= 1) return 2 * 1 + 0;
    if (sizeof(R) == 8 || nrhs <= 4) return 2 * 4 + 0;
    return 2 * 8 + 0;
}

template <typename R, int D>
void dispatch_sapc(int variant, hipStream_t s, const _Float16* Lh, const cplx<R>* L, const cplx<R>* r, cplx<R>* x,
                   int Nx, int Nt,
                          int bx, int bt, int colour, double m0, int restarts, double tol, int nrhs, uint64_t mask,
                          size_t ld, const int* gate) {
#define DD_V(NB, F) launch_sapc<R, D, NB, F>(s, L, r, x, Nx, Nt, bx, bt, colour, m0, restarts, tol, nrhs, mask, ld, gate)
    switch (variant) {
        case 2: DD_V(1, false); break;
        case 3: DD_V(1, true); break;
        case 5: DD_V(2, true); break;
        case 8: DD_V(4, false); break;
        case 9: DD_V(4, true); break;
        case 16:
            if constexpr (sizeof(R) == 4) DD_V(8, false);
            else throw BadShape("sap_blocksc: variant not compiled for FP64");
            break;
        case 17:
            if constexpr (sizeof(R) == 4) DD_V(8, true);
            else throw BadShape("sap_blocksc: variant not compiled for FP64");
            break;
        case 32:
        case 33:
            if constexpr (sizeof(R) == 4 && D <= 24) {
                if (variant == 33 && !Lh) throw BadShape("sap_blocksc: variant 33 needs the binary16 links");
                launch_sapm<D>(s, L, variant == 33 ? Lh : nullptr, r, x, Nx, Nt, bx, bt, colour, m0, restarts, tol, nrhs,
                               mask, ld, gate);
            } else throw BadShape("sap_blocksc: matrix-core variant needs FP32 and dof <= 24");
            break;
        default:
            throw BadShape("sap_blocksc: unknown variant");
    }
#undef DD_V
}

template <typename R, int D>
bool dispatch_sap_fused(hipStream_t s, const cplx<R>* L, const _Float16* Lh, const cplx<R>* v, cplx<R>* x, int Nx,
                        int Nt, int bx, int bt, double m0, int restarts, double tol, int nrhs, uint64_t mask, size_t ld,
                        int variant, const SapFuse& fz, int nslice) {
    if constexpr (sizeof(R) == 4 && D <= 24) {
        if (variant != 32 && variant != 33) return false;
        if (variant == 33 && !Lh) return false;
        launch_sapm_fused<D>(s, L, variant == 33 ? Lh : nullptr, v, x, Nx, Nt, bx, bt, m0, restarts, tol, nrhs, mask,
                             ld, fz, nslice);
        return true;
    } else {
        return false;
    }
}

// Each (precision, dof) pair of the coarse SAP dispatch is compiled in its own
// translation unit (sap_<prec>_d<dof>.hip) to keep builds parallel.
#define DD_SAPC_DECL(R, D)                                                                                         \
    extern template void dispatch_sapc<R, D>(int, hipStream_t, const _Float16*, const cplx<R>*, const cplx<R>*,     \
                                             cplx<R>*, int, int,                                                 \
                                             int, int, int, double, int, double, int, uint64_t, size_t, const int*); \
    extern template bool dispatch_sap_fused<R, D>(hipStream_t, const cplx<R>*, const _Float16*, const cplx<R>*,     \
                                                  cplx<R>*, int, int, int, int, double, int, double, int, uint64_t,  \
                                                  size_t, int, const SapFuse&, int);
#ifndef DD_SAPC_UNIT
DD_SAPC_DECL(float, 8)
DD_SAPC_DECL(float, 16)
DD_SAPC_DECL(float, 24)
DD_SAPC_DECL(float, 32)
DD_SAPC_DECL(double, 8)
DD_SAPC_DECL(double, 16)
DD_SAPC_DECL(double, 24)
DD_SAPC_DECL(double, 32)
#endif

template <typename R>
bool sap_fused_iteration(hipStream_t s, const cplx<R>* L, const _Float16* Lh, const cplx<R>* v, cplx<R>* x, int Nx,
                         int Nt, int dof, int bx, int bt, double m0, int m, int restarts, double tol, int nrhs,
                         uint64_t mask, size_t ld, int variant, const SapFuse& fz, int nslice) {
    const int S = bx * bt;
    if (m != kSapM || nrhs > 64 || bx < 1 || bt < 1 || Nx % bx || Nt % bt) return false;
    const int nbx = Nx / bx, nbt = Nt / bt;
    if ((nbx & 1) || (nbt & 1) || nbx < 2 || nbt < 2) return false;   // a true checkerboard with periodic wrap
    if (variant < 0) variant = sapc_default_variant<R>(nrhs, dof, S, Lh != nullptr);
    if ((variant != 32 && variant != 33) || !sapc_supported<R>(variant, dof, S)) return false;
    bool done = false;
    timer_begin(s, K_SAPC);
    switch (dof) {
#define DD_CASE(D)                                                                                              \
    case D:                                                                                                     \
        done = dispatch_sap_fused<R, D>(s, L, Lh, v, x, Nx, Nt, bx, bt, m0, restarts, tol, nrhs, mask, ld, variant, \
                                        fz, nslice);                                                            \
        break;
        DD_DOF_LIST(DD_CASE)
#undef DD_CASE
        default:
            break;
    }
    // both colours' block solves (links once per colour) and, with nslice > 0, the norm
    // pass (links and the iterate once)
    const double lb = variant == 33 ? 2.0 * sizeof(_Float16) : (double)sizeof(cplx<R>);
    const double lnorm = nslice > 0 ? (double)sizeof(cplx<R>) : 0.0;
    timer_end(s, K_SAPC, done ? Nx * Nt * dof * (5.0 * dof * (lb + lnorm) +
                                               3.0 * nact(mask, nrhs) * sizeof(cplx<R>))
                              : 0.0);
    DD_LAUNCH_CHECK();
    return done;
}

template <typename R>
void sap_blocksc(hipStream_t s, const cplx<R>* L, const cplx<R>* r, cplx<R>* x, int Nx, int Nt, int dof, int bx,
                 int bt, int colour, double m0, int m, int restarts, double tol, int nrhs, uint64_t mask, size_t ld,
                 const int* gate, int variant, const _Float16* Lh) {
    timer_begin(s, K_SAPC);
    const int S = bx * bt;
    if (m != kSapM) throw BadShape("SAP block GMRES restart length must be 5");
    if (variant < 0) variant = sapc_default_variant<R>(nrhs, dof, S, Lh != nullptr);
    if (!sapc_supported<R>(variant, dof, S)) throw BadShape("coarse SAP: variant unsupported for this block/dof");
    switch (dof) {
#define DD_CASE(D)                                                                                             \
    case D:                                                                                                    \
        dispatch_sapc<R, D>(variant, s, Lh, L, r, x, Nx, Nt, bx, bt, colour, m0, restarts, tol, nrhs, mask, ld,   \
                            gate);                                                                                \
        break;
        DD_DOF_LIST(DD_CASE)
#undef DD_CASE
        default:
            throw BadShape("sap_blocksc: unsupported dof");
    }
    const double lb = variant == 33 ? 2.0 * sizeof(_Float16) : (double)sizeof(cplx<R>);   // bytes per link element
    timer_end(s, K_SAPC, 0.5 * Nx * Nt * dof * (5.0 * dof * lb + 3.0 * nact(mask, nrhs) * sizeof(cplx<R>)));
    DD_LAUNCH_CHECK();
}

}  // namespace dd
