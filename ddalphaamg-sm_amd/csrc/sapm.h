// Coarse-level SAP block solves on the matrix cores (FP32), 16 systems per work-group.
//
// Same algorithm as k_sapc (block GMRES(5) of the SAP smoother with the reference's
// control flow, src/fgmres.cpp:5-127 with the GMRES_D_B hooks include/sap.h:14-83,
// x0 = block right-hand side src/sap.cpp:65; block operator
// Level::SAP_level_l::D_local src/level.cpp:607-653), organised around the MFMA
// complex block product of mfma.h:
//   * wave w owns the block sites w + W*u (u < SPW); a lane owns, for system j = lane&15
//     and each owned site, the complex rows 8t + 2(lane>>4) + {0, 1} of every row tile t
//     (the accumulator layout of the 16x16x4 product), so the Krylov basis, w and the
//     iterate updates live in registers in exactly the layout the products produce;
//   * D_B v: v is staged in LDS as [site][row][system] (the B operand layout, read
//     without bank conflicts), every link element of the block is read once per
//     product for all 16 systems;
//   * per-system reductions: xor-shuffles over the four lane groups, then one LDS
//     slot per wave (double-buffered), summed in a fixed order so every lane of a
//     system sees the same bits;
//   * the Givens/back-substitution scalars of system j live in LDS and are updated by
//     lane j of wave 0; per-system flags are uniform 16-bit masks (ballots);
//   * the work-groups of one block's system tiles are numbered 8 apart (1-D grid), so
//     the round-robin dispatch puts them on the same XCD at the same time and the second
//     tile reads the block's links from that XCD's L2;
//   * the links are read as FP32 or as a binary16 copy (LT = _Float16, widened on load):
//     the kernel streams every link of its block once per operator application, so the
//     copy halves its dominant memory traffic.  The copy perturbs each link element by at
//     most 2^-11 relative; it serves only the smoother's block solves (the level operator
//     and the Galerkin products keep the FP32 links).
#pragma once
#include "lanes.h"
#include "mfma.h"

namespace dd {

template <int DOF, int SPW>
struct SapmVec {
    static constexpr int MT = DOF / 8, N = SPW * MT * 2;
    cplx<float> e[N];   // e[(u*MT + t)*2 + h]: site u, complex row 8t + 2g + h
};

template <int DOF>
constexpr size_t sapm_lds_bytes(int S) {
    return sizeof(cplx<float>) * ((size_t)3 * S * DOF * 16 + 2 * 16 * 16 + 16 * (Tri<kSapM>::size + (kSapM + 1) + 3 * kSapM)) +
           sizeof(float) * 16 + sizeof(int) * 20;
}

constexpr unsigned kFuseSpins = 1u << 24;

// waits (thread 0 of the work-group) until *w == want; false after kFuseSpins polls or
// when the sticky failure word is set
__device__ __forceinline__ bool fuse_wait_eq(const unsigned* w, unsigned want, unsigned* fail) {
    unsigned spins = 0;
    while (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != want) {
        if ((++spins & 63u) == 0 &&
            (__hip_atomic_load(fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || spins > kFuseSpins)) {
            __hip_atomic_store(fail, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}

// One block solve of a colour for one 16-system tile (task `gid` of the colour).
// FUSED = false: the block right-hand side is read from rr (the residual v - D x).
// FUSED = true (one launch per SAP iteration, k_sapm_fused): rr is v, and the block
// right-hand side v - D x is formed here from the iterate xx and the FP32 links L32 at
// the block's sites; a black block first waits until its four (red) neighbour blocks of
// the same tile have stored their iterates, and every block publishes its own when done.
template <int DOF, int SPW, typename LT, bool FUSED>
__device__ __forceinline__ void sapm_task(const LT* __restrict__ L, const float* __restrict__ L32,
                                          const cplx<float>* __restrict__ rr, cplx<float>* __restrict__ xx, int Nx,
                                          int Nt, int bx, int bt, int colour, float m0, int restarts, float tol,
                                          uint64_t mask, size_t ld, const int* __restrict__ gate, int nrhs, int ntile,
                                          int gid, const SapFuse& fz) {
    using R = float;
    using acc_t = Mfma<R>::acc_t;
    using Vec = SapmVec<DOF, SPW>;
    constexpr int M = kSapM, MT = DOF / 8, KS = DOF / 2, TS = Tri<M>::size, NE = Vec::N;
    extern __shared__ unsigned char smem_raw[];
    const int S = bx * bt, W = blockDim.x >> 6;
    cplx<R>* stg = reinterpret_cast<cplx<R>*>(smem_raw);   // [S][DOF][16] product operand
    cplx<R>* sphi = stg + S * DOF * 16;                   // [S][DOF][16] block right-hand sides
    cplx<R>* sx = sphi + S * DOF * 16;                    // [S][DOF][16] block iterates
    cplx<R>* red = sx + S * DOF * 16;                     // [2][16 waves max][16]
    cplx<R>* Hs = red + 2 * 16 * 16;                      // [16][TS]
    cplx<R>* gs = Hs + 16 * TS;                           // [16][M+1]
    cplx<R>* cns = gs + 16 * (M + 1);                     // [16][M]
    cplx<R>* sns = cns + 16 * M;                          // [16][M]
    cplx<R>* etas = sns + 16 * M;                         // [16][M]
    int* maxIts = reinterpret_cast<int*>(reinterpret_cast<R*>(etas + 16 * M) + 16);   // [16]
    int* flag = maxIts + 16;                                                          // [1]

    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j = lane & 15, g = lane >> 4;
    const bool lead = (wv == 0) && (g == 0);   // lane j of wave 0: scalars of system j
    const int nbx = Nx / bx, nbt = Nt / bt;
    // work-group id -> (block, system tile): 8 consecutive blocks per group of 8 ntile ids
    const int grp = gid / (8 * ntile), rr8 = gid % (8 * ntile);
    const int blk = grp * 8 + (rr8 & 7), tile = rr8 >> 3;
    if (blk >= (nbx * nbt) / 2) return;
    int BX, BT;
    colour_block(blk, colour, nbx, nbt, BX, BT);
    const int r0sys = tile * 16;
    unsigned onm = 0;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int rh = r0sys + q;
        if (rh < nrhs && rhs_on(mask, rh) && (!gate || gate[rh] != 0)) onm |= 1u << q;
    }
    const int gblk = BX * nbt + BT;   // block index over both colours
    if (FUSED) {
        // the iterates of the neighbour blocks this block's residual reads: red blocks read
        // black ones, which no task of this launch changes before this block is done (a
        // black block waits for its red neighbours); black blocks wait for the red ones
        if (colour == 1) {
            if (threadIdx.x == 0) {
                const int nb[4] = {((BX + 1) % nbx) * nbt + BT, ((BX + nbx - 1) % nbx) * nbt + BT,
                                   BX * nbt + (BT + 1) % nbt, BX * nbt + (BT + nbt - 1) % nbt};
                for (int k = 0; k < 4; k++)
                    if (!fuse_wait_eq(fz.done + (size_t)nb[k] * ntile + tile, fz.epoch, fz.fail)) break;
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            __syncthreads();
        }
        if (!onm) {   // nothing to solve: publish at once
            if (threadIdx.x == 0) {
                if (colour == 0)
                    __hip_atomic_store(fz.done + (size_t)gblk * ntile + tile, fz.epoch, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
                else
                    __hip_atomic_fetch_add(fz.ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
            return;
        }
    }
    if (!onm) return;
    const int rh = r0sys + j;
    const bool on = (onm >> j) & 1;

    // owned sites: local index, global site, neighbour local indices (-1 outside the block)
    int kl[SPW], gn[SPW], nbl[SPW][5];
    bool own[SPW];
#pragma unroll
    for (int u = 0; u < SPW; u++) {
        const int k = wv + W * u;
        own[u] = k < S;
        const int kk = own[u] ? k : 0;
        const int lx = kk / bt, lt = kk - lx * bt;
        kl[u] = kk;
        gn[u] = (BX * bx + lx) * Nt + BT * bt + lt;
        nbl[u][0] = kk;
        nbl[u][1] = lt < bt - 1 ? kk + 1 : -1;
        nbl[u][2] = lx < bx - 1 ? kk + bt : -1;
        nbl[u][3] = lt > 0 ? kk - 1 : -1;
        nbl[u][4] = lx > 0 ? kk - bt : -1;
    }
    auto lidx = [&](int u, int t, int h) { return (kl[u] * DOF + 8 * t + 2 * g + h) * 16 + j; };
    const LinkLane ll = LinkLane::make<R, DOF>(lane);
    const R mass = m0 + R(2);
    const R* stgf = reinterpret_cast<const R*>(stg);

    // v <- D_B v (zero outside the block, no boundary sign).  The caller reaches a
    // work-group barrier (a reduction) before the next call re-stages.
    int napp = 0;   // block-operator applications (flop count)
    auto apply = [&](Vec& v) {
        ++napp;
#pragma unroll
        for (int u = 0; u < SPW; u++)
            if (own[u]) {
#pragma unroll
                for (int t = 0; t < MT; t++)
#pragma unroll
                    for (int h = 0; h < 2; h++) stg[lidx(u, t, h)] = v.e[(u * MT + t) * 2 + h];
            }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < SPW; u++) {
            if (!own[u]) continue;
            acc_t acc[MT];
#pragma unroll
            for (int t = 0; t < MT; t++) acc[t] = acc_t{0, 0, 0, 0};
            const LT* Ln = L + (size_t)gn[u] * 5 * DOF * DOF * 2;
#pragma unroll 1
            for (int hop = 0; hop < 5; hop++) {
                const int nk = nbl[u][hop];
                if (nk < 0) continue;
                R bv[KS];
#pragma unroll
                for (int s2 = 0; s2 < KS; s2++) bv[s2] = stgf[((nk * DOF + 2 * s2 + (g >> 1)) * 16 + j) * 2 + (g & 1)];
                mfma_hop<R, DOF, (DOF / 2) % 4 == 0 ? 4 : DOF / 2, LT>(acc, Ln + (size_t)hop * DOF * DOF * 2, ll, bv);
            }
#pragma unroll
            for (int t = 0; t < MT; t++) {
                cplx<R>& a0 = v.e[(u * MT + t) * 2];
                cplx<R>& a1 = v.e[(u * MT + t) * 2 + 1];
                a0 = mk<R>(mass * a0.x - acc[t][0], mass * a0.y - acc[t][1]);
                a1 = mk<R>(mass * a1.x - acc[t][2], mass * a1.y - acc[t][3]);
            }
        }
    };
    // total over the block of this lane's partial, per system j (identical bits in every
    // lane of the system)
    int buf = 0;
    auto rsum = [&](cplx<R> p) -> cplx<R> {
        p.x += lane_xor<16>(p.x, lane);   // permlane swaps: the lane network, not LDS
        p.y += lane_xor<16>(p.y, lane);
        p.x += lane_xor<32>(p.x, lane);
        p.y += lane_xor<32>(p.y, lane);
        cplx<R>* rb = red + buf * 256;
        if (g == 0) rb[wv * 16 + j] = p;
        __syncthreads();
        cplx<R> t = czero<R>();
        for (int w = 0; w < W; w++) t = t + rb[w * 16 + j];
        buf ^= 1;
        return t;
    };
    auto dotc = [&](const Vec& a, const Vec& b) -> cplx<R> {   // sum a * conj(b) over owned rows
        cplx<R> p = czero<R>();
#pragma unroll
        for (int u = 0; u < SPW; u++)
            if (own[u])
#pragma unroll
                for (int e = 0; e < 2 * MT; e++) p = p + mulc(a.e[u * 2 * MT + e], b.e[u * 2 * MT + e]);
        return p;
    };
    auto nrm2 = [&](const Vec& a) -> R {
        R p = 0;
#pragma unroll
        for (int u = 0; u < SPW; u++)
            if (own[u])
#pragma unroll
                for (int e = 0; e < 2 * MT; e++) p += norm2(a.e[u * 2 * MT + e]);
        return rsum(mk<R>(p, R(0))).x;
    };
    auto load_lds = [&](Vec& v, const cplx<R>* src) {
#pragma unroll
        for (int u = 0; u < SPW; u++)
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int h = 0; h < 2; h++) v.e[(u * MT + t) * 2 + h] = own[u] ? src[lidx(u, t, h)] : czero<R>();
    };

    auto resid = [&](Vec& r, const Vec& dw) {   // r = phi - dw
#pragma unroll
        for (int u = 0; u < SPW; u++)
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const int e = (u * MT + t) * 2 + h;
                    r.e[e] = (own[u] ? sphi[lidx(u, t, h)] : czero<R>()) - dw.e[e];
                }
    };

    // right-hand sides (initial guess = right-hand side)
    Vec w, V[M];
    if (FUSED) {   // v - D x at the owned sites (the global stencil: neighbours anywhere, t sign)
        const cplx<R>* rv = rr + (size_t)(on ? rh : 0) * ld;
        const R* xs = reinterpret_cast<const R*>(xx + (size_t)(on ? rh : 0) * ld);
#pragma unroll
        for (int u = 0; u < SPW; u++) {
            const int n = gn[u], x0 = n / Nt, t0 = n - x0 * Nt;
            acc_t acc[MT];
#pragma unroll
            for (int t = 0; t < MT; t++) acc[t] = acc_t{0, 0, 0, 0};
            const float* Ln = L32 + (size_t)n * 5 * DOF * DOF * 2;
#pragma unroll 1
            for (int hop = 0; hop < 5; hop++) {
                const int nh = hop == 0 ? n
                             : hop == 1 ? ((t0 == Nt - 1) ? n - (Nt - 1) : n + 1)
                             : hop == 2 ? ((x0 == Nx - 1) ? t0 : n + Nt)
                             : hop == 3 ? ((t0 == 0) ? n + (Nt - 1) : n - 1)
                                        : ((x0 == 0) ? (Nx - 1) * Nt + t0 : n - Nt);
                const R sh = (hop == 1 && t0 == Nt - 1) || (hop == 3 && t0 == 0) ? R(-1) : R(1);
                R bv[KS];
#pragma unroll
                for (int s2 = 0; s2 < KS; s2++) bv[s2] = sh * xs[(size_t)nh * DOF * 2 + 4 * s2 + g];
                mfma_hop<R, DOF, (DOF / 2) % 4 == 0 ? 4 : DOF / 2, float>(acc, Ln + (size_t)hop * DOF * DOF * 2, ll,
                                                                          bv);
            }
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const size_t o = (size_t)n * DOF + 8 * t + 2 * g + h;
                    const cplx<R> xv = xx[(size_t)(on ? rh : 0) * ld + o], vv = rv[o];
                    const cplx<R> ph = (own[u] && on) ? mk<R>(vv.x - (mass * xv.x - acc[t][2 * h]),
                                                              vv.y - (mass * xv.y - acc[t][2 * h + 1]))
                                                      : czero<R>();
                    if (own[u]) {
                        sphi[lidx(u, t, h)] = ph;
                        sx[lidx(u, t, h)] = ph;
                    }
                    w.e[(u * MT + t) * 2 + h] = ph;
                }
        }
    } else {
        const cplx<R>* rv = rr + (size_t)(on ? rh : 0) * ld;
#pragma unroll
        for (int u = 0; u < SPW; u++)
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const cplx<R> ph =
                        (own[u] && on) ? rv[(size_t)gn[u] * DOF + 8 * t + 2 * g + h] : czero<R>();
                    if (own[u]) {
                        sphi[lidx(u, t, h)] = ph;
                        sx[lidx(u, t, h)] = ph;
                    }
                    w.e[(u * MT + t) * 2 + h] = ph;
                }
    }
    const R nphi = sqrt(nrm2(w));
    apply(w);
    resid(V[0], w);
    R err = sqrt(nrm2(V[0]));
    if (lead) maxIts[j] = M;

    unsigned livem = onm;
    // a zero block right-hand side has the solution zero: no iterations (no 1/0 scaling)
    livem &= ~((unsigned)__ballot(g == 0 && nphi == R(0)) & 0xffffu);
    for (int kk = 0; kk < restarts && livem; kk++) {
        unsigned runm = livem;
        const bool live = (livem >> j) & 1;
        if (live) {
            const R ie = R(1) / err;
#pragma unroll
            for (int e = 0; e < NE; e++) V[0].e[e] = ie * V[0].e[e];
        }
        if (lead && live) gs[j * (M + 1)] = mk<R>(err, R(0));
#pragma unroll
        for (int jj = 0; jj < M; jj++) {
            if (!runm) break;
            const bool run = (runm >> j) & 1;
            w = V[jj];
            apply(w);
#pragma unroll
            for (int ii = 0; ii <= jj; ii++) {
                const cplx<R> h = rsum(dotc(w, V[ii]));
                if (lead && run) Hs[j * TS + Tri<M>::at(ii, jj)] = h;
                if (run)
#pragma unroll
                    for (int e = 0; e < NE; e++) w.e[e] = w.e[e] - h * V[ii].e[e];
            }
            const R hv = sqrt(nrm2(w));
            if (jj + 1 < M && run && hv > R(0)) {
                const R ih = R(1) / hv;
#pragma unroll
                for (int e = 0; e < NE; e++) V[jj + 1].e[e] = ih * w.e[e];
            }
            bool broke = false;
            if (lead && run) {
                // Givens rotations of column jj (src/fgmres.cpp:99-115) for system j
                cplx<R>* H = Hs + j * TS;
                cplx<R>* cn = cns + j * M;
                cplx<R>* sn = sns + j * M;
                cplx<R>* gg = gs + j * (M + 1);
                for (int i2 = 0; i2 < jj; i2++) {
                    const cplx<R> a = H[Tri<M>::at(i2, jj)], bb = H[Tri<M>::at(i2 + 1, jj)];
                    H[Tri<M>::at(i2, jj)] = cmul(cn[i2], a) + cmul(sn[i2], bb);
                    H[Tri<M>::at(i2 + 1, jj)] = cn[i2] * bb - sn[i2] * a;
                }
                const cplx<R> hjj = H[Tri<M>::at(jj, jj)];
                const cplx<R> hs = mk<R>(hv, R(0));
                const R den = sqrt(norm2(hjj) + norm2(hs));
                sn[jj] = mk<R>(hs.x / den, hs.y / den);
                cn[jj] = mk<R>(hjj.x / den, hjj.y / den);
                H[Tri<M>::at(jj, jj)] = cmul(cn[jj], hjj) + cmul(sn[jj], hs);
                gg[jj + 1] = -(sn[jj] * gg[jj]);
                gg[jj] = cmul(cn[jj], gg[jj]);
                if (sqrt(norm2(gg[jj + 1])) < tol * nphi) {
                    maxIts[j] = jj + 1;
                    broke = true;
                }
            }
            const unsigned brk = (unsigned)__ballot(broke) & 0xffffu;   // wave 0 only
            if (wv == 0 && lane == 0) *flag = (int)brk;
            __syncthreads();
            runm &= ~(unsigned)(*flag);
        }
        // back substitution (src/fgmres.cpp:118-127) by lane j of wave 0
        if (lead && live) {
            const int nI = maxIts[j];
            const cplx<R>* H = Hs + j * TS;
            const cplx<R>* gg = gs + j * (M + 1);
            cplx<R>* eta = etas + j * M;
            for (int i2 = nI - 1; i2 >= 0; i2--) {
                cplx<R> e = gg[i2];
                for (int j2 = i2 + 1; j2 < nI; j2++) e = e - H[Tri<M>::at(i2, j2)] * eta[j2];
                eta[i2] = cdiv(e, H[Tri<M>::at(i2, i2)]);
            }
        }
        __syncthreads();
        load_lds(w, sx);
        if (live) {
            const int nI = maxIts[j];
#pragma unroll
            for (int i2 = 0; i2 < M; i2++)
                if (i2 < nI) {
                    const cplx<R> et = etas[j * M + i2];
#pragma unroll
                    for (int e = 0; e < NE; e++) w.e[e] = w.e[e] + et * V[i2].e[e];
                }
#pragma unroll
            for (int u = 0; u < SPW; u++)
                if (own[u])
#pragma unroll
                    for (int t = 0; t < MT; t++)
#pragma unroll
                        for (int h = 0; h < 2; h++) sx[lidx(u, t, h)] = w.e[(u * MT + t) * 2 + h];
        }
        apply(w);
        resid(V[0], w);
        const R e2 = nrm2(V[0]);
        const bool conv = live && sqrt(e2) < tol * nphi;
        if (live) err = sqrt(e2);
        const unsigned cm = (unsigned)__ballot(conv && g == 0) & 0xffffu;
        livem &= ~cm;
    }
    if (on) {
        cplx<R>* xv = xx + (size_t)rh * ld;
#pragma unroll
        for (int u = 0; u < SPW; u++)
            if (own[u])
#pragma unroll
                for (int t = 0; t < MT; t++)
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const size_t o = (size_t)gn[u] * DOF + 8 * t + 2 * g + h;
                        xv[o] = xv[o] + sx[lidx(u, t, h)];
                    }
    }
    if (fz.flops && threadIdx.x == 0) {
        // matrix-core flops of the task: 128 DOF^2 per (site, hop) product of 16 systems;
        // the block operator has the in-block hops, the fused right-hand side all five
        const unsigned long long pairs = (unsigned long long)napp * (S + 2 * bx * (bt - 1) + 2 * bt * (bx - 1)) +
                                         (FUSED ? 5ull * S : 0ull);
        __hip_atomic_fetch_add(fz.flops, pairs * 128ull * DOF * DOF, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (FUSED) {   // publish this block's iterate: every wave's stores, then one release
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (threadIdx.x == 0) {
            if (colour == 0)
                __hip_atomic_store(fz.done + (size_t)gblk * ntile + tile, fz.epoch, __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            else
                __hip_atomic_fetch_add(fz.ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

template <int DOF, int SPW, typename LT>
__global__ __launch_bounds__(512) void k_sapm(const LT* __restrict__ L, const cplx<float>* __restrict__ rr,
                                              cplx<float>* __restrict__ xx, int Nx, int Nt, int bx, int bt,
                                              int colour, float m0, int restarts, float tol, uint64_t mask, size_t ld,
                                              const int* __restrict__ gate, int nrhs, int ntile,
                                              unsigned long long* flops) {
    SapFuse fz{};
    fz.flops = flops;
    sapm_task<DOF, SPW, LT, false>(L, nullptr, rr, xx, Nx, Nt, bx, bt, colour, m0, restarts, tol, mask, ld, gate,
                                   nrhs, ntile, blockIdx.x, fz);
}

// One SAP iteration (src/sap.cpp:103-156) in one launch: work-groups [0, nt) solve the red
// blocks, [nt, 2nt) the black blocks, [2nt, 2nt + ns) form ||v - D x||^2 per system over
// slices of the sites, and the last slice to finish updates the gate (iterations stop for
// the systems with ||v - D x|| < tol ||v||).  Work-groups start in the order of their
// ids, so a work-group only ever waits for lower ids that are already running: black
// blocks for their red neighbours, norm slices for every black block.  Needs even
// numbers of blocks in both directions (no block borders one of its own colour).
template <int DOF, int SPW, typename LT>
__global__ __launch_bounds__(512) void k_sapm_fused(const LT* __restrict__ L, const float* __restrict__ L32,
                                                    const cplx<float>* __restrict__ v, cplx<float>* __restrict__ xx,
                                                    int Nx, int Nt, int bx, int bt, float m0, int restarts, float tol,
                                                    uint64_t mask, size_t ld, int nrhs, int ntile, int ntask,
                                                    int nslice, SapFuse fz) {
    using R = float;
    using acc_t = Mfma<R>::acc_t;
    constexpr int MT = DOF / 8, KS = DOF / 2;
    const int gid = blockIdx.x;
    if (gid < 2 * ntask) {
        const int colour = gid < ntask ? 0 : 1;
        sapm_task<DOF, SPW, LT, true>(L, L32, v, xx, Nx, Nt, bx, bt, colour, m0, restarts, tol, mask, ld, fz.gate,
                                      nrhs, ntile, gid - colour * ntask, fz);
        return;
    }
    // norm slice: ||v - D x||^2 per system over sites [n0, n1), once every black block is done
    const int sl = gid - 2 * ntask, V = Nx * Nt;
    const int n0 = (int)((long long)V * sl / nslice), n1 = (int)((long long)V * (sl + 1) / nslice);
    __shared__ float red_s[16][64];   // one row per wave (at most 16 waves)
    __shared__ int ok_s;
    if (threadIdx.x == 0) {
        const unsigned nblack = (unsigned)((Nx / bx) * (Nt / bt) / 2 * ntile);   // black tasks that exist
        ok_s = fuse_wait_eq(fz.ctr, nblack, fz.fail) ? 1 : 0;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, W = blockDim.x >> 6;
    const LinkLane ll = LinkLane::make<R, DOF>(lane);
    const R mass = m0 + R(2);
    float sumsq[4] = {0, 0, 0, 0};   // per tile k < ntile (at most 4 tiles: 64 systems)
    for (int n = n0 + wv; n < n1 && ok_s; n += W) {
        const int x0 = n / Nt, t0 = n - x0 * Nt;
        const float* Ln = L32 + (size_t)n * 5 * DOF * DOF * 2;
        for (int k = 0; k < ntile; k++) {
            const int j = k * 16 + (lane & 15);
            const bool on = j < nrhs && rhs_on(mask, j);
            const R* xs = reinterpret_cast<const R*>(xx + (size_t)(on ? j : 0) * ld);
            acc_t acc[MT];
#pragma unroll
            for (int t = 0; t < MT; t++) acc[t] = acc_t{0, 0, 0, 0};
#pragma unroll 1
            for (int hop = 0; hop < 5; hop++) {
                const int nh = hop == 0 ? n
                             : hop == 1 ? ((t0 == Nt - 1) ? n - (Nt - 1) : n + 1)
                             : hop == 2 ? ((x0 == Nx - 1) ? t0 : n + Nt)
                             : hop == 3 ? ((t0 == 0) ? n + (Nt - 1) : n - 1)
                                        : ((x0 == 0) ? (Nx - 1) * Nt + t0 : n - Nt);
                const R sh = (hop == 1 && t0 == Nt - 1) || (hop == 3 && t0 == 0) ? R(-1) : R(1);
                R bv[KS];
#pragma unroll
                for (int s2 = 0; s2 < KS; s2++) bv[s2] = sh * xs[(size_t)nh * DOF * 2 + 4 * s2 + g];
                mfma_hop<R, DOF, (DOF / 2) % 4 == 0 ? 4 : DOF / 2, float>(acc, Ln + (size_t)hop * DOF * DOF * 2,
                                                                          ll, bv);
            }
            float p = 0;
#pragma unroll
            for (int t = 0; t < MT; t++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const size_t o = (size_t)(on ? j : 0) * ld + (size_t)n * DOF + 8 * t + 2 * g + h;
                    const cplx<R> xv = xx[o], vv = v[o];
                    const R re = vv.x - (mass * xv.x - acc[t][2 * h]), im = vv.y - (mass * xv.y - acc[t][2 * h + 1]);
                    p += re * re + im * im;
                }
            sumsq[k] += on ? p : 0.0f;
        }
    }
    // per system: the four lane groups, then the eight waves in order, then the slices
    for (int k = 0; k < 4; k++) {
        float p = sumsq[k];
        p += lane_xor<16>(p, lane);
        p += lane_xor<32>(p, lane);
        if (g == 0) red_s[wv][k * 16 + (lane & 15)] = p;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        float t = 0;
        for (int w = 0; w < W; w++) t += red_s[w][threadIdx.x];
        fz.part[(size_t)sl * 64 + threadIdx.x] = t;
    }
    if (fz.flops && threadIdx.x == 0 && ok_s)   // the slice's stencil: five products per site and tile
        __hip_atomic_fetch_add(fz.flops, (unsigned long long)(n1 - n0) * ntile * 5ull * 128ull * DOF * DOF,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    __shared__ unsigned last_s;
    if (threadIdx.x == 0)
        last_s = __hip_atomic_fetch_add(fz.ctr + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                 (unsigned)nslice - 1;
    __syncthreads();
    if (last_s && threadIdx.x < 64) {   // the last slice: totals in slice order, gate update
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const int j = threadIdx.x;
        if (j < nrhs && rhs_on(mask, j)) {
            double t = 0;
            for (int q = 0; q < nslice; q++) t += (double)fz.part[(size_t)q * 64 + j];
            if (sqrt(t) < fz.tol * sqrt(fz.vnorm[2 * j].x)) fz.gate[j] = 0;
        }
    }
}

template <int DOF, int SPW, typename LT>
static void launch_sapm_k(hipStream_t s, dim3 grid, int threads, size_t sh, const LT* L, const cplx<float>* r,
                          cplx<float>* x, int Nx, int Nt, int bx, int bt, int colour, double m0, int restarts,
                          double tol, int nrhs, uint64_t mask, size_t ld, const int* gate, int ntile) {
    static bool attr = false;
    if (!attr) {
        DD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sapm<DOF, SPW, LT>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr = true;
    }
    k_sapm<DOF, SPW, LT><<<grid, threads, sh, s>>>(L, r, x, Nx, Nt, bx, bt, colour, (float)m0, restarts, (float)tol,
                                                    mask, ld, gate, nrhs, ntile, timer_flops_dev(K_SAPC));
}

// One fused SAP iteration (k_sapm_fused) for the systems in mask; L the FP32 links, Lh
// (nullable) the binary16 copy the block solves then read.
template <int DOF>
static void launch_sapm_fused(hipStream_t s, const cplx<float>* L, const _Float16* Lh, const cplx<float>* v,
                              cplx<float>* x, int Nx, int Nt, int bx, int bt, double m0, int restarts, double tol,
                              int nrhs, uint64_t mask, size_t ld, const SapFuse& fz, int nslice) {
    const int S = bx * bt;
    const size_t sh = sapm_lds_bytes<DOF>(S);
    const int nblk = (Nx / bx) * (Nt / bt) / 2, ntile = (nrhs + 15) / 16;
    const int ntask = ((nblk + 7) / 8) * 8 * ntile;
    const dim3 grid(2 * ntask + nslice);
    const float* Lf = reinterpret_cast<const float*>(L);
    auto go = [&](auto kern, int threads, const auto* Ls) {
        static bool attr = false;
        if (!attr) {   // the dynamic LDS may take what the kernel's static LDS leaves of 160 KB
            hipFuncAttributes fa;
            DD_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kern)));
            DD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)(160 * 1024 - fa.sharedSizeBytes)));
            attr = true;
        }
        if (sh + 4200 > 160 * 1024) throw BadShape("fused coarse SAP: block does not fit the LDS");
        SapFuse fzc = fz;
        fzc.flops = timer_flops_dev(K_SAPC);
        kern<<<grid, threads, sh, s>>>(Ls, Lf, v, x, Nx, Nt, bx, bt, (float)m0, restarts, (float)tol, mask, ld, nrhs,
                                       ntile, ntask, nslice, fzc);
    };
    if (S > 8) {
        if (Lh) go(&k_sapm_fused<DOF, 2, _Float16>, 64 * ((S + 1) / 2), Lh);
        else go(&k_sapm_fused<DOF, 2, float>, 64 * ((S + 1) / 2), Lf);
    } else {
        if (Lh) go(&k_sapm_fused<DOF, 1, _Float16>, 64 * S, Lh);
        else go(&k_sapm_fused<DOF, 1, float>, 64 * S, Lf);
    }
}

// L: FP32 links [V][5][DOF][DOF] complex, or (Lh non-null) their binary16 copy, which is
// then the one read
template <int DOF>
static void launch_sapm(hipStream_t s, const cplx<float>* L, const _Float16* Lh, const cplx<float>* r,
                        cplx<float>* x, int Nx, int Nt, int bx, int bt, int colour, double m0, int restarts,
                        double tol, int nrhs, uint64_t mask, size_t ld, const int* gate) {
    const int S = bx * bt;
    if (S > 16) throw BadShape("matrix-core coarse SAP: at most 16 sites per block");
    const size_t sh = sapm_lds_bytes<DOF>(S);
    if (sh > 160 * 1024) throw BadShape("matrix-core coarse SAP: block does not fit the LDS");
    const int nblk = (Nx / bx) * (Nt / bt) / 2, ntile = (nrhs + 15) / 16;
    const dim3 grid(((nblk + 7) / 8) * 8 * ntile);
    const float* Lf = reinterpret_cast<const float*>(L);
#define DD_SAPM(SPW, THREADS)                                                                                      \
    (Lh ? launch_sapm_k<DOF, SPW, _Float16>(s, grid, THREADS, sh, Lh, r, x, Nx, Nt, bx, bt, colour, m0, restarts, \
                                            tol, nrhs, mask, ld, gate, ntile)                                      \
        : launch_sapm_k<DOF, SPW, float>(s, grid, THREADS, sh, Lf, r, x, Nx, Nt, bx, bt, colour, m0, restarts, tol, \
                                         nrhs, mask, ld, gate, ntile))
    if (S > 8) DD_SAPM(2, 64 * ((S + 1) / 2));   // two sites per wave
    else DD_SAPM(1, 64 * S);
#undef DD_SAPM
}

}  // namespace dd
