// gfx950 kernels of the DD-alphaAMG solve path.  See common.h for the layouts.
//
// Reference behaviour restated here (paths relative to the reference root):
//   k_wilson        D_phi / D_dagger_phi          src/dirac_operator.cpp:24-85
//   k_dirac_links   Level::makeDirac              src/level.cpp:286-331
//   k_coarse        Level::D_operator             src/level.cpp:394-426
//   k_sap0/k_sapc   SAP_C::I_D_B_1_It + the block GMRES (FGMRES::fgmres with
//                   SAP_C::GMRES_D_B hooks) and SAP_level_l::D_local
//                   src/sap.cpp:48-77, src/fgmres.cpp:5-127, include/sap.h:14-83,
//                   src/level.cpp:607-653
//   k_restrict      Level::Pdagg_v                src/level.cpp:145-189
//   k_prolong       Level::P_vc                   src/level.cpp:92-141
//   k_ortho         Level::orthonormalize         src/level.cpp:193-284
//   k_galerkin      Level::makeCoarseLinks        src/level.cpp:431-542
#include "kernels.h"
#include "mfma.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace dd {

#define DD_LAUNCH_CHECK() DD_HIP(hipGetLastError())

// ---- live kernel timing ------------------------------------------------------------
namespace {
struct TimerSlot {
    bool on = false;
    std::vector<hipEvent_t> pool;   // start/end pairs
    size_t used = 0;
    double bytes = 0;
};
TimerSlot g_timers[K_NUM];
unsigned long long* g_flops = nullptr;   // [K_NUM] device counters, allocated on first enable
}  // namespace

void timer_enable(int id, bool on) {
    if (id < 0 || id >= K_NUM) return;
    g_timers[id].on = on;
    if (on && !g_flops) {
        DD_HIP(hipMalloc(reinterpret_cast<void**>(&g_flops), sizeof(unsigned long long) * K_NUM));
        DD_HIP(hipMemset(g_flops, 0, sizeof(unsigned long long) * K_NUM));
    }
}
void timer_reset() {
    for (auto& t : g_timers) {
        t.used = 0;
        t.bytes = 0;
    }
    if (g_flops) {
        DD_HIP(hipDeviceSynchronize());
        DD_HIP(hipMemset(g_flops, 0, sizeof(unsigned long long) * K_NUM));
    }
}
unsigned long long* timer_flops_dev(int id) {
    return (id >= 0 && id < K_NUM && g_timers[id].on && g_flops) ? g_flops + id : nullptr;
}
double timer_read_flops(int id) {
    if (!g_flops || id < 0 || id >= K_NUM) return 0.0;
    unsigned long long h = 0;
    DD_HIP(hipDeviceSynchronize());
    DD_HIP(hipMemcpy(&h, g_flops + id, sizeof(h), hipMemcpyDeviceToHost));
    return (double)h;
}
void timer_begin(hipStream_t s, int id) {
    TimerSlot& t = g_timers[id];
    if (!t.on) return;
    if (t.used + 2 > t.pool.size()) {
        for (int k = 0; k < 64; k++) {
            hipEvent_t e;
            DD_HIP(hipEventCreate(&e));
            t.pool.push_back(e);
        }
    }
    DD_HIP(hipEventRecord(t.pool[t.used], s));
}
void timer_end(hipStream_t s, int id, double bytes) {
    TimerSlot& t = g_timers[id];
    if (!t.on) return;
    DD_HIP(hipEventRecord(t.pool[t.used + 1], s));
    t.used += 2;
    t.bytes += bytes;
}
void timer_release() {
    for (auto& t : g_timers) {
        for (hipEvent_t e : t.pool) (void)hipEventDestroy(e);
        t.pool.clear();
        t.used = 0;
        t.bytes = 0;
    }
    if (g_flops) (void)hipFree(g_flops);
    g_flops = nullptr;
}
bool timer_any() {
    for (const auto& t : g_timers)
        if (t.on) return true;
    return false;
}
void timer_add_bytes(int id, double bytes) {
    TimerSlot& t = g_timers[id];
    if (t.on) t.bytes += bytes;
}
double timer_read(int id, long long* count, double* bytes) {
    TimerSlot& t = g_timers[id];
    double tot = 0;
    for (size_t k = 0; k + 1 < t.used; k += 2) {
        DD_HIP(hipEventSynchronize(t.pool[k + 1]));
        float ms = 0;
        DD_HIP(hipEventElapsedTime(&ms, t.pool[k], t.pool[k + 1]));
        tot += ms;
    }
    if (count) *count = (long long)(t.used / 2);
    if (bytes) *bytes = t.bytes;
    return tot;
}

// ---- device-side system mask of the calling thread's launches ---------------------------
namespace {
thread_local const uint64_t* tl_dmask = nullptr;
}  // namespace
const uint64_t* dev_mask() { return tl_dmask; }
DevMaskScope::DevMaskScope(const uint64_t* m) : prev(tl_dmask) { tl_dmask = m; }
DevMaskScope::~DevMaskScope() { tl_dmask = prev; }

static inline unsigned cdivu(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
static inline double nact(uint64_t mask, int nrhs) { return (double)__builtin_popcountll(mask & ((nrhs >= 64) ? ~0ull : ((1ull << nrhs) - 1))); }

// ======================================================================================
// Level-0 Wilson-Dirac stencil.  One thread per site and group of up to RB systems: the
// four link matrices of the site are loaded once for the group.  Work-groups are numbered
// XCD by XCD: the hardware deals consecutive work-group ids round-robin to the 8 XCDs, so
// work-group id p works on logical block (p % 8) * (G / 8) + p / 8 (when 8 divides the
// number of work-groups G).  Each XCD then sweeps a contiguous band of sites, and the
// x-neighbour rows a block reads were fetched by the blocks just before it into the
// same XCD's L2.
// ======================================================================================
__device__ __forceinline__ int xcd_block(int p, int G) { return (G & 7) ? p : (p & 7) * (G >> 3) + (p >> 3); }

template <typename R, bool DAG, bool RES, int RB>
__global__ __launch_bounds__(256) void k_wilson(const cplx<R>* __restrict__ U, const cplx<R>* __restrict__ in,
                                                cplx<R>* __restrict__ out, const cplx<R>* __restrict__ b, int Nx,
                                                int Nt, R m0, uint64_t mask, size_t ld, int nsb, int nrhs, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int L = xcd_block(blockIdx.x, gridDim.x);
    const int sb = L % nsb, r0 = (L / nsb) * RB;
    const int n = sb * blockDim.x + threadIdx.x;
    if (n >= Nx * Nt) return;
    const int x = n / Nt, t = n - x * Nt;
    const int ntp = (t == Nt - 1) ? n - (Nt - 1) : n + 1;
    const int ntm = (t == 0) ? n + (Nt - 1) : n - 1;
    const int nxp = (x == Nx - 1) ? t : n + Nt;
    const int nxm = (x == 0) ? (Nx - 1) * Nt + t : n - Nt;
    const R sp = (t == Nt - 1) ? R(-1) : R(1);
    const R sm = (t == 0) ? R(-1) : R(1);
    const cplx<R> U0 = U[2 * n], U1 = U[2 * n + 1];
    const cplx<R> U0m = conj(U[2 * ntm]), U1m = conj(U[2 * nxm + 1]);
    const R mass = m0 + R(2);
#pragma unroll
    for (int q = 0; q < RB; q++) {
        const int r = r0 + q;
        if (r >= nrhs || !rhs_on(mask, r)) continue;
        const cplx<R>* v = in + r * ld;
        const cplx<R> p0 = v[2 * n], p1 = v[2 * n + 1];
        const cplx<R> tp0 = v[2 * ntp], tp1 = v[2 * ntp + 1];
        const cplx<R> xp0 = v[2 * nxp], xp1 = v[2 * nxp + 1];
        const cplx<R> tm0 = v[2 * ntm], tm1 = v[2 * ntm + 1];
        const cplx<R> xm0 = v[2 * nxm], xm1 = v[2 * nxm + 1];
        cplx<R> a0, a1;
        if (!DAG) {
            a0 = U0 * (sp * (tp0 - tp1)) + U1 * (xp0 + times_i(xp1)) + U0m * (sm * (tm0 + tm1)) +
                 U1m * (xm0 - times_i(xm1));
            a1 = U0 * (sp * (tp1 - tp0)) + U1 * (xp1 - times_i(xp0)) + U0m * (sm * (tm0 + tm1)) +
                 U1m * (times_i(xm0) + xm1);
        } else {
            a0 = U0m * (sm * (tm0 - tm1)) + U1m * (xm0 + times_i(xm1)) + U0 * (sp * (tp0 + tp1)) +
                 U1 * (xp0 - times_i(xp1));
            a1 = U0m * (sm * (tm1 - tm0)) + U1m * (xm1 - times_i(xm0)) + U0 * (sp * (tp0 + tp1)) +
                 U1 * (times_i(xp0) + xp1);
        }
        cplx<R> o0 = mass * p0 - R(0.5) * a0;
        cplx<R> o1 = mass * p1 - R(0.5) * a1;
        cplx<R>* o = out + r * ld;
        if (RES) {
            const cplx<R>* bb = b + r * ld;
            o0 = bb[2 * n] - o0;
            o1 = bb[2 * n + 1] - o1;
        }
        o[2 * n] = o0;
        o[2 * n + 1] = o1;
    }
}

template <typename R, bool DAG, bool RES>
static void launch_wilson(hipStream_t s, const cplx<R>* U, const cplx<R>* in, cplx<R>* out, const cplx<R>* b, int Nx,
                          int Nt, double m0, int nrhs, uint64_t mask, size_t ld) {
    constexpr int RB = 4;
    const int nsb = (int)cdivu((size_t)Nx * Nt, 256), ng = (nrhs + RB - 1) / RB;
    k_wilson<R, DAG, RES, RB><<<dim3((unsigned)nsb * ng), 256, 0, s>>>(U, in, out, b, Nx, Nt, (R)m0, mask, ld, nsb,
                                                                      nrhs, dev_mask());
}

template <typename R>
void wilson(hipStream_t s, const cplx<R>* U, const cplx<R>* in, cplx<R>* out, const cplx<R>* b, int Nx, int Nt,
            double m0, bool dagger, int nrhs, uint64_t mask, size_t ld) {
    timer_begin(s, K_WILSON);
    if (dagger) {
        if (b) launch_wilson<R, true, true>(s, U, in, out, b, Nx, Nt, m0, nrhs, mask, ld);
        else launch_wilson<R, true, false>(s, U, in, out, b, Nx, Nt, m0, nrhs, mask, ld);
    } else {
        if (b) launch_wilson<R, false, true>(s, U, in, out, b, Nx, Nt, m0, nrhs, mask, ld);
        else launch_wilson<R, false, false>(s, U, in, out, b, Nx, Nt, m0, nrhs, mask, ld);
    }
    timer_end(s, K_WILSON, (double)Nx * Nt * 2 * sizeof(cplx<R>) * (1.0 + nact(mask, nrhs) * (b ? 3 : 2)));
    DD_LAUNCH_CHECK();
}

// Level-0 links as generic stencil matrices (A = 0):
//   B_mu = 1/2 (1 - sigma_mu) U_mu(x),   C_mu = 1/2 (1 + sigma_mu) U_mu(x - mu)^*
template <typename R>
__global__ void k_dirac_links(const cplx<R>* __restrict__ U, cplx<R>* __restrict__ L, int Nx, int Nt) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Nx * Nt) return;
    const int x = n / Nt, t = n - x * Nt;
    const int ntm = (t == 0) ? n + (Nt - 1) : n - 1;
    const int nxm = (x == 0) ? (Nx - 1) * Nt + t : n - Nt;
    const cplx<R> u[2] = {U[2 * n], U[2 * n + 1]};
    const cplx<R> um[2] = {conj(U[2 * ntm]), conj(U[2 * nxm + 1])};
    const cplx<R> one = mk<R>(1, 0), I = mk<R>(0, 1);
    // (1 - sigma_mu)[row][col] and (1 + sigma_mu)[row][col]
    const cplx<R> Mm[2][2][2] = {{{one, -one}, {-one, one}}, {{one, I}, {-I, one}}};
    const cplx<R> Pm[2][2][2] = {{{one, one}, {one, one}}, {{one, -I}, {I, one}}};
    cplx<R>* o = L + (size_t)n * 5 * 4;
    for (int e = 0; e < 4; e++) o[e] = czero<R>();
    for (int mu = 0; mu < 2; mu++)
        for (int col = 0; col < 2; col++)
            for (int row = 0; row < 2; row++) {
                o[(1 + mu) * 4 + col * 2 + row] = R(0.5) * (Mm[mu][row][col] * u[mu]);
                o[(3 + mu) * 4 + col * 2 + row] = R(0.5) * (Pm[mu][row][col] * um[mu]);
            }
}

template <typename R>
void dirac_links(hipStream_t s, const cplx<R>* U, cplx<R>* L, int Nx, int Nt) {
    k_dirac_links<R><<<cdivu((size_t)Nx * Nt, 256), 256, 0, s>>>(U, L, Nx, Nt);
    DD_LAUNCH_CHECK();
}

// ======================================================================================
// Generic level stencil  out = (m0+2) v - A v - sum_mu (B_mu s+ v(x+mu) + C_mu s- v(x-mu))
// on the matrix cores: one wave per site and tile of 16 systems; each link element is
// read once per launch for all 16 systems of the tile (real-block 16x16x4 products,
// layout in mfma.h).  The neighbour spinors are read straight from the vectors (four
// lanes per system share each 16-byte segment).
// ======================================================================================
template <typename R, int DOF, bool RES, int NT>
__global__ __launch_bounds__(256) void k_coarse(const cplx<R>* __restrict__ L, const cplx<R>* __restrict__ in,
                                                cplx<R>* __restrict__ out, const cplx<R>* __restrict__ b, int Nx,
                                                int Nt, R m0, uint64_t mask, size_t ld, int nrhs, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    // XCD by XCD, as k_wilson: the neighbour sites' vectors come from the same XCD's L2
    const int n = xcd_block(blockIdx.x, gridDim.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (n >= Nx * Nt) return;   // uniform per wave
    coarse_site<R, DOF, RES, NT>(L, in, out, b, Nx, Nt, m0, mask, ld, nrhs, n, blockIdx.y * NT, threadIdx.x & 63);
}

#define DD_DOF_LIST(X) X(8) X(16) X(24) X(32)

template <typename R>
void coarse_apply(hipStream_t s, const cplx<R>* L, const cplx<R>* in, cplx<R>* out, const cplx<R>* b, int Nx,
                  int Nt, int dof, double m0, int nrhs, uint64_t mask, size_t ld) {
    timer_begin(s, K_COARSE);
    constexpr int kWaves = 4;   // sites per work-group
    const int top = mask ? 64 - __builtin_clzll(mask) : 1;   // systems up to the highest one set
    const int tiles = (int)cdivu((size_t)std::min(top, nrhs), 16);
    // two 16-system tiles per wave share every link load once there are more than 16 systems
    const bool two = tiles >= 2;
    dim3 grid(cdivu((size_t)Nx * Nt, kWaves), two ? cdivu(tiles, 2) : tiles);
    switch (dof) {
#define DD_CASE(D)                                                                                            \
    case D:                                                                                                   \
        if (two) {                                                                                            \
            if (b) k_coarse<R, D, true, 2><<<grid, 64 * kWaves, 0, s>>>(L, in, out, b, Nx, Nt, (R)m0, mask, ld, nrhs, dev_mask()); \
            else k_coarse<R, D, false, 2><<<grid, 64 * kWaves, 0, s>>>(L, in, out, b, Nx, Nt, (R)m0, mask, ld, nrhs, dev_mask()); \
        } else {                                                                                              \
            if (b) k_coarse<R, D, true, 1><<<grid, 64 * kWaves, 0, s>>>(L, in, out, b, Nx, Nt, (R)m0, mask, ld, nrhs, dev_mask()); \
            else k_coarse<R, D, false, 1><<<grid, 64 * kWaves, 0, s>>>(L, in, out, b, Nx, Nt, (R)m0, mask, ld, nrhs, dev_mask()); \
        }                                                                                                     \
        break;
        DD_DOF_LIST(DD_CASE)
#undef DD_CASE
        default:
            throw BadShape("coarse_apply: unsupported dof");
    }
    timer_end(s, K_COARSE, (double)Nx * Nt * dof * sizeof(cplx<R>) * (5.0 * dof + nact(mask, nrhs) * (b ? 3 : 2)));
    DD_LAUNCH_CHECK();
}

// ======================================================================================
// Transfers
// ======================================================================================
// The transfers serve RB systems per thread: every test-vector element is loaded once for
// the group (the test vectors are the dominant operand), and the work-groups are numbered
// XCD by XCD as in k_wilson.
constexpr int kTransferRB = 8;

template <typename R, int RB>
__global__ __launch_bounds__(256) void k_restrict(const cplx<R>* __restrict__ W, const cplx<R>* __restrict__ v,
                                                  cplx<R>* __restrict__ out, int Nx, int Nt, int dof, int ax, int at,
                                                  int nv, uint64_t mask, size_t ldf, size_t ldc, int nrhs, int nsb,
                                                  const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int L = xcd_block(blockIdx.x, gridDim.x);
    const int sb = L % nsb, r0 = (L / nsb) * RB;
    const int Ntc = Nt / at, Vc = (Nx / ax) * Ntc;
    const int idx = sb * blockDim.x + threadIdx.x;
    if (idx >= Vc * 2 * nv) return;
    bool any = false;
#pragma unroll
    for (int k = 0; k < RB; k++) any |= r0 + k < nrhs && rhs_on(mask, r0 + k);
    if (!any) return;
    const int Xc = idx / (2 * nv), q = idx - Xc * 2 * nv, p = q >> 1, sp = q & 1;
    const int X = Xc / Ntc, Tc = Xc - X * Ntc;
    const int col = dof / 2;
    cplx<R> sum[RB];
#pragma unroll
    for (int k = 0; k < RB; k++) sum[k] = czero<R>();
    for (int lx = 0; lx < ax; lx++)
        for (int lt = 0; lt < at; lt++) {
            const size_t n = (size_t)(X * ax + lx) * Nt + Tc * at + lt;
            for (int c = 0; c < col; c++) {
                const int d = c * 2 + sp;
                const cplx<R> w = W[(n * nv + p) * dof + d];
#pragma unroll
                for (int k = 0; k < RB; k++) {
                    const int r = min(r0 + k, nrhs - 1);
                    sum[k] = sum[k] + cmul(w, v[(size_t)r * ldf + n * dof + d]);
                }
            }
        }
#pragma unroll
    for (int k = 0; k < RB; k++) {
        const int r = r0 + k;
        if (r < nrhs && rhs_on(mask, r)) out[(size_t)r * ldc + (size_t)Xc * 2 * nv + q] = sum[k];
    }
}

template <typename R>
void restrict_(hipStream_t s, const cplx<R>* W, const cplx<R>* v, cplx<R>* out, int Nx, int Nt, int dof, int ax,
               int at, int nv, int nrhs, uint64_t mask, size_t ldf, size_t ldc) {
    timer_begin(s, K_RESTRICT);
    const size_t tot = (size_t)(Nx / ax) * (Nt / at) * 2 * nv;
    const int nsb = (int)cdivu(tot, 256), ng = (nrhs + kTransferRB - 1) / kTransferRB;
    k_restrict<R, kTransferRB><<<dim3((unsigned)nsb * ng), 256, 0, s>>>(W, v, out, Nx, Nt, dof, ax, at, nv, mask, ldf,
                                                                        ldc, nrhs, nsb, dev_mask());
    timer_end(s, K_RESTRICT, (double)Nx * Nt * dof * sizeof(cplx<R>) * (nv + nact(mask, nrhs)) + nact(mask, nrhs) * (double)(Nx / ax) * (Nt / at) * 2 * nv * sizeof(cplx<R>));
    DD_LAUNCH_CHECK();
}

template <typename R, int RB>
__global__ __launch_bounds__(256) void k_prolong(const cplx<R>* __restrict__ W, const cplx<R>* __restrict__ vc,
                                                 const cplx<R>* y, cplx<R>* out, int Nx, int Nt, int dof, int ax,
                                                 int at, int nv, uint64_t mask, size_t ldf, size_t ldc, int nrhs,
                                                 int nsb, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int L = xcd_block(blockIdx.x, gridDim.x);
    const int sb = L % nsb, r0 = (L / nsb) * RB;
    const size_t idx = (size_t)sb * blockDim.x + threadIdx.x;
    if (idx >= (size_t)Nx * Nt * dof) return;
    bool any = false;
#pragma unroll
    for (int k = 0; k < RB; k++) any |= r0 + k < nrhs && rhs_on(mask, r0 + k);
    if (!any) return;
    const int n = (int)(idx / dof), d = (int)(idx - (size_t)n * dof), sp = d & 1;
    const int x = n / Nt, t = n - x * Nt;
    const int Xc = (x / ax) * (Nt / at) + t / at;
    const cplx<R>* Wn = W + (size_t)n * nv * dof + d;
    cplx<R> sum[RB];
#pragma unroll
    for (int k = 0; k < RB; k++) sum[k] = czero<R>();
    for (int p = 0; p < nv; p++) {
        const cplx<R> w = Wn[p * dof];
#pragma unroll
        for (int k = 0; k < RB; k++) {
            const int r = min(r0 + k, nrhs - 1);
            sum[k] = sum[k] + w * vc[(size_t)r * ldc + (size_t)Xc * 2 * nv + sp + 2 * p];
        }
    }
#pragma unroll
    for (int k = 0; k < RB; k++) {
        const int r = r0 + k;
        if (r < nrhs && rhs_on(mask, r)) {
            cplx<R> o = sum[k];
            if (y) o = y[(size_t)r * ldf + idx] + o;
            out[(size_t)r * ldf + idx] = o;
        }
    }
}

template <typename R>
void prolong(hipStream_t s, const cplx<R>* W, const cplx<R>* vc, const cplx<R>* y, cplx<R>* out, int Nx, int Nt,
             int dof, int ax, int at, int nv, int nrhs, uint64_t mask, size_t ldf, size_t ldc) {
    timer_begin(s, K_PROLONG);
    const int nsb = (int)cdivu((size_t)Nx * Nt * dof, 256), ng = (nrhs + kTransferRB - 1) / kTransferRB;
    k_prolong<R, kTransferRB><<<dim3((unsigned)nsb * ng), 256, 0, s>>>(W, vc, y, out, Nx, Nt, dof, ax, at, nv, mask,
                                                                       ldf, ldc, nrhs, nsb, dev_mask());
    timer_end(s, K_PROLONG, (double)Nx * Nt * dof * sizeof(cplx<R>) * (nv + (y ? 2 : 1) * nact(mask, nrhs)) + nact(mask, nrhs) * (double)(Nx / ax) * (Nt / at) * 2 * nv * sizeof(cplx<R>));
    DD_LAUNCH_CHECK();
}

// Modified Gram-Schmidt per (aggregate, spin); one work-group, one thread per
// (site, colour) entry; the aggregate's vectors live in LDS.
template <typename R>
__global__ __launch_bounds__(1024) void k_ortho(cplx<R>* __restrict__ W, int Nx, int Nt, int dof, int ax, int at,
                                                int nv) {
    extern __shared__ unsigned char smem_raw[];
    const int col = dof / 2, E = ax * at * col;
    cplx<R>* w = reinterpret_cast<cplx<R>*>(smem_raw);   // [nv][E]
    R* scratch = reinterpret_cast<R*>(w + nv * E);
    const int Ntc = Nt / at;
    const int agg = blockIdx.x >> 1, sp = blockIdx.x & 1;
    const int X = agg / Ntc, Tc = agg - X * Ntc;
    const int e = threadIdx.x;
    const bool act = e < E;
    size_t gidx = 0;
    if (act) {
        const int site = e / col, c = e - site * col;
        const int lx = site / at, lt = site - lx * at;
        const size_t n = (size_t)(X * ax + lx) * Nt + Tc * at + lt;
        gidx = n * nv * dof + c * 2 + sp;
        for (int p = 0; p < nv; p++) w[p * E + e] = W[gidx + (size_t)p * dof];
    }
    __syncthreads();
    for (int nt = 0; nt < nv; nt++) {
        for (int ntt = 0; ntt < nt; ntt++) {
            const cplx<R> part = act ? mulc(w[nt * E + e], w[ntt * E + e]) : czero<R>();
            const cplx<R> proj = block_sum(part, scratch);
            if (act) w[nt * E + e] = w[nt * E + e] - proj * w[ntt * E + e];
        }
        const cplx<R> part = act ? mk<R>(norm2(w[nt * E + e]), R(0)) : czero<R>();
        const R nrm = sqrt(block_sum(part, scratch).x);
        if (act && nrm > R(1e-8)) {
            const cplx<R> v = w[nt * E + e];
            w[nt * E + e] = mk<R>(v.x / nrm, v.y / nrm);
        }
    }
    if (act)
        for (int p = 0; p < nv; p++) W[gidx + (size_t)p * dof] = w[p * E + e];
}

template <typename R>
void orthonormalize(hipStream_t s, cplx<R>* W, int Nx, int Nt, int dof, int ax, int at, int nv) {
    const int E = ax * at * (dof / 2);
    const int threads = ((E + 63) / 64) * 64;
    if (threads > 1024) throw BadShape("orthonormalize: aggregate too large");
    const size_t sh = sizeof(cplx<R>) * nv * E + sizeof(R) * 32;
    const int nagg = (Nx / ax) * (Nt / at);
    k_ortho<R><<<2 * nagg, threads, sh, s>>>(W, Nx, Nt, dof, ax, at, nv);
    DD_LAUNCH_CHECK();
}

// Galerkin coarse links.  One work-group per coarse site; thread (i, j) of the
// dc x dc output (dc = 2 nv) accumulates all five matrices.  For each fine site x of
// the aggregate and each link k of the fine stencil, T = L_k(x) P_y is formed in LDS
// (y the neighbour the link reaches) and then P_x^dagger T is added to A_c, or to
// B_c,mu / C_c,mu when the hop leaves the aggregate through its +mu / -mu face.
template <typename R>
__global__ __launch_bounds__(1024) void k_galerkin(const cplx<R>* __restrict__ Lf, const cplx<R>* __restrict__ W,
                                                   cplx<R>* __restrict__ Lc, int Nx, int Nt, int df, int ax, int at,
                                                   int nv, int has_A) {
    extern __shared__ unsigned char smem_raw[];
    const int dc = 2 * nv, col = df / 2;
    cplx<R>* wx = reinterpret_cast<cplx<R>*>(smem_raw);   // [nv][df]
    cplx<R>* wy = wx + nv * df;                           // [nv][df]
    cplx<R>* T = wy + nv * df;                            // [df][dc]
    const int Ntc = Nt / at;
    const int Xc = blockIdx.x, X = Xc / Ntc, Tc = Xc - X * Ntc;
    const int tid = threadIdx.x;
    const bool own = tid < dc * dc;
    const int i = own ? tid % dc : 0, j = own ? tid / dc : 0;   // output row i=(p,alf), col j=(s,bet)
    const int p = i >> 1, alf = i & 1;
    cplx<R> acc[5];
#pragma unroll
    for (int q = 0; q < 5; q++) acc[q] = czero<R>();
    for (int lx = 0; lx < ax; lx++)
        for (int lt = 0; lt < at; lt++) {
            const int x = X * ax + lx, t = Tc * at + lt;
            const int n = x * Nt + t;
            __syncthreads();
            for (int e = tid; e < nv * df; e += blockDim.x) wx[e] = W[(size_t)n * nv * df + e];
            for (int kk = has_A ? 0 : 1; kk < 5; kk++) {
                int y, target;
                switch (kk) {
                    case 0: y = n; target = 0; break;
                    case 1: y = x * Nt + (t == Nt - 1 ? 0 : t + 1); target = (lt == at - 1) ? 1 : 0; break;
                    case 2: y = (x == Nx - 1 ? 0 : x + 1) * Nt + t; target = (lx == ax - 1) ? 2 : 0; break;
                    case 3: y = x * Nt + (t == 0 ? Nt - 1 : t - 1); target = (lt == 0) ? 3 : 0; break;
                    default: y = (x == 0 ? Nx - 1 : x - 1) * Nt + t; target = (lx == 0) ? 4 : 0; break;
                }
                __syncthreads();
                for (int e = tid; e < nv * df; e += blockDim.x) wy[e] = W[(size_t)y * nv * df + e];
                __syncthreads();
                // T[row=(c,a)][jj=(s,b)] = sum_b' L_k(x)[row][(b',b)] w_s(y, b', b)
                const cplx<R>* Lk = Lf + ((size_t)n * 5 + kk) * df * df;
                for (int e = tid; e < df * dc; e += blockDim.x) {
                    const int row = e / dc, jj = e - row * dc, ss = jj >> 1, bet = jj & 1;
                    cplx<R> sum = czero<R>();
                    for (int bb = 0; bb < col; bb++) {
                        const int cidx = bb * 2 + bet;
                        sum = sum + Lk[cidx * df + row] * wy[ss * df + cidx];
                    }
                    T[e] = sum;
                }
                __syncthreads();
                if (own) {
                    cplx<R> sum = czero<R>();
                    for (int c = 0; c < col; c++) {
                        const int rr = c * 2 + alf;
                        sum = sum + cmul(wx[p * df + rr], T[rr * dc + j]);
                    }
                    // static indexing keeps the accumulators in registers
                    if (target == 0) acc[0] = acc[0] + sum;
                    else if (target == 1) acc[1] = acc[1] + sum;
                    else if (target == 2) acc[2] = acc[2] + sum;
                    else if (target == 3) acc[3] = acc[3] + sum;
                    else acc[4] = acc[4] + sum;
                }
            }
        }
    if (own) {
        cplx<R>* o = Lc + (size_t)Xc * 5 * dc * dc;
#pragma unroll
        for (int q = 0; q < 5; q++) o[(q * dc + j) * dc + i] = acc[q];
    }
}

template <typename R>
void galerkin(hipStream_t s, const cplx<R>* Lf, const cplx<R>* W, cplx<R>* Lc, int Nx, int Nt, int df, int ax, int at,
              int nv, bool has_A) {
    const int dc = 2 * nv;
    if (dc * dc > 1024) throw BadShape("galerkin: too many test vectors");
    const int threads = std::max(64, ((dc * dc + 63) / 64) * 64);
    const size_t sh = sizeof(cplx<R>) * (2 * nv * df + df * dc);
    const int Vc = (Nx / ax) * (Nt / at);
    k_galerkin<R><<<Vc, threads, sh, s>>>(Lf, W, Lc, Nx, Nt, df, ax, at, nv, has_A ? 1 : 0);
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ void k_transpose_tv(const cplx<R>* __restrict__ in, cplx<R>* __restrict__ out, int V, int nv, int dof,
                               int to_site) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t tot = (size_t)V * nv * dof;
    if (idx >= tot) return;
    // idx enumerates the site-major layout [V][nv][dof]
    const int d = (int)(idx % dof);
    const size_t q = idx / dof;
    const int p = (int)(q % nv);
    const size_t n = q / nv;
    const size_t vm = ((size_t)p * V + n) * dof + d;   // vector-major [nv][V][dof]
    if (to_site) out[idx] = in[vm];
    else out[vm] = in[idx];
}

template <typename R>
void transpose_tv(hipStream_t s, const cplx<R>* in, cplx<R>* out, int V, int nv, int dof, bool to_site_major) {
    k_transpose_tv<R><<<cdivu((size_t)V * nv * dof, 256), 256, 0, s>>>(in, out, V, nv, dof, to_site_major ? 1 : 0);
    DD_LAUNCH_CHECK();
}

// ======================================================================================
// BLAS-1 with deterministic two-pass reductions (partials summed in a fixed order).
// ======================================================================================
template <typename R>
__global__ __launch_bounds__(256) void k_dot_partial(const cplx<R>* __restrict__ a, const cplx<R>* __restrict__ b,
                                                     size_t n, size_t ld, uint64_t mask,
                                                     cplx<double>* __restrict__ partial, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    __shared__ R scratch[32];
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    const cplx<R>* ar = a + r * ld;
    const cplx<R>* br = b + r * ld;
    cplx<R> s = czero<R>();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s = s + mulc(ar[i], br[i]);
    s = block_sum(s, scratch);
    if (threadIdx.x == 0) partial[(size_t)r * gridDim.x + blockIdx.x] = mk<double>(s.x, s.y);
}

static inline int red_blocks(size_t n) {
    return (int)std::min<size_t>(kMaxRedBlocks, std::max<size_t>(1, (n + 2047) / 2048));
}

template <typename R>
void dot_partial(hipStream_t s, const cplx<R>* a, const cplx<R>* b, size_t n, size_t ld, int nrhs, uint64_t mask,
                 RedBuf& rb) {
    timer_begin(s, K_DOT);
    rb.nblk = red_blocks(n);
    dim3 grid(rb.nblk, nrhs);
    k_dot_partial<R><<<grid, 256, 0, s>>>(a, b, n, ld, mask, (cplx<double>*)rb.partial, dev_mask());
    timer_end(s, K_DOT, nact(mask, nrhs) * (double)n * 2 * sizeof(cplx<R>));
    DD_LAUNCH_CHECK();
}

// sum of the nblk partials of system r, in a fixed order (lane-strided then butterfly),
// by one wave; every lane returns the same bits
__device__ __forceinline__ cplx<double> sum_partials(const cplx<double>* p, int nblk) {
    const int lane = threadIdx.x & 63;
    cplx<double> s = czero<double>();
    for (int i = lane; i < nblk; i += 64) s = s + p[i];
    return wave_sum(s, 64);
}

__global__ void k_dot_final(const cplx<double>* __restrict__ partial, int nblk, cplx<double>* out, int ostride,
                            uint64_t mask, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.x;
    if (!rhs_on(mask, r)) return;
    const cplx<double> s = sum_partials(partial + (size_t)r * nblk, nblk);
    if (threadIdx.x == 0) out[(size_t)r * ostride] = s;
}

template <typename R>
void dot_final(hipStream_t s, const RedBuf& rb, cplx<double>* out, int ostride, int nrhs, uint64_t mask) {
    k_dot_final<<<nrhs, 64, 0, s>>>((const cplx<double>*)rb.partial, rb.nblk, out, ostride, mask, dev_mask());
    DD_LAUNCH_CHECK();
}

static inline unsigned ew_blocks(size_t n) { return (unsigned)std::min<size_t>(1024, std::max<size_t>(1, (n + 255) / 256)); }

// One modified Gram-Schmidt projection and the next one's partial sums in a single pass:
// h = sum of the partials in `pin` -> H[r*hstride]; w -= h v; then the partials of
// <w, nxt> (nxt = w: |w|^2) -> pout.  The threads walk the elements exactly as
// k_dot_partial does and sum them with the same block_sum, so the partials, and every h
// and norm formed from them, have the bits of an update pass w -= h v followed by
// k_dot_partial.  pin and pout must not overlap (every block reads all of pin before any
// block's write).
template <typename R>
__global__ __launch_bounds__(256) void k_mgs_update_dot(cplx<R>* __restrict__ w, const cplx<R>* __restrict__ v,
                                                        const cplx<R>* __restrict__ nxt, size_t n, size_t ld,
                                                        uint64_t mask, const cplx<double>* __restrict__ pin,
                                                        cplx<double>* __restrict__ pout, cplx<double>* H, int hstride,
                                                        const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    __shared__ cplx<double> hs;
    __shared__ R scratch[32];
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    const int nblk = gridDim.x;
    if (threadIdx.x < 64) {
        const cplx<double> h = sum_partials(pin + (size_t)r * nblk, nblk);
        if (threadIdx.x == 0) {
            hs = h;
            if (blockIdx.x == 0) H[(size_t)r * hstride] = h;
        }
    }
    __syncthreads();
    const cplx<R> h = mk<R>((R)hs.x, (R)hs.y);
    cplx<R>* wr = w + r * ld;
    const cplx<R>* vr = v + r * ld;
    const cplx<R>* nr = nxt ? nxt + r * ld : nullptr;
    cplx<R> s = czero<R>();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)nblk * blockDim.x) {
        const cplx<R> wv = wr[i] - h * vr[i];
        wr[i] = wv;
        s = s + mulc(wv, nr ? nr[i] : wv);
    }
    s = block_sum(s, scratch);
    if (threadIdx.x == 0) pout[(size_t)r * nblk + blockIdx.x] = mk<double>(s.x, s.y);
}

template <typename R>
void mgs_update_dot(hipStream_t s, cplx<R>* w, const cplx<R>* v, const cplx<R>* nxt, size_t n, size_t ld, int nrhs,
                    uint64_t mask, const RedBuf& in, RedBuf& out, cplx<double>* H, int hstride) {
    if (in.partial == out.partial) throw BadShape("mgs_update_dot: partial buffers must differ");
    if (in.nblk != red_blocks(n)) throw BadShape("mgs_update_dot: input partials from another length");
    timer_begin(s, K_DOT);
    out.nblk = in.nblk;
    dim3 grid(in.nblk, nrhs);
    k_mgs_update_dot<R><<<grid, 256, 0, s>>>(w, v, nxt, n, ld, mask, (const cplx<double>*)in.partial,
                                             (cplx<double>*)out.partial, H, hstride, dev_mask());
    timer_end(s, K_DOT, nact(mask, nrhs) * (double)n * (nxt ? 4 : 3) * sizeof(cplx<R>));
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ __launch_bounds__(256) void k_norm_scale(cplx<R>* __restrict__ out, const cplx<R>* __restrict__ w, size_t n,
                                                    size_t ld, uint64_t mask, const cplx<double>* __restrict__ partial,
                                                    int nblk, cplx<double>* H, int hstride, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    __shared__ double hs;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    if (threadIdx.x < 64) {
        const cplx<double> h = sum_partials(partial + (size_t)r * nblk, nblk);
        if (threadIdx.x == 0) {
            // partial sums are accumulated in R; the norm is formed in R as well
            const R hn = sqrt((R)h.x);
            hs = (double)hn;
            if (blockIdx.x == 0) H[(size_t)r * hstride] = mk<double>((double)hn, 0.0);
        }
    }
    __syncthreads();
    const R hn = (R)hs;
    if (!(hn > R(0))) return;
    const R inv = R(1) / hn;
    cplx<R>* o = out + r * ld;
    const cplx<R>* wr = w + r * ld;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = inv * wr[i];
}

template <typename R>
void norm_scale(hipStream_t s, cplx<R>* out, const cplx<R>* w, size_t n, size_t ld, int nrhs, uint64_t mask,
                const RedBuf& rb, cplx<double>* H, int hstride) {
    dim3 grid(ew_blocks(n), nrhs);
    k_norm_scale<R><<<grid, 256, 0, s>>>(out, w, n, ld, mask, (const cplx<double>*)rb.partial, rb.nblk, H, hstride, dev_mask());
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ __launch_bounds__(256) void k_scale_inv(cplx<R>* __restrict__ out, const cplx<R>* __restrict__ w, size_t n,
                                                   size_t ld, uint64_t mask, const cplx<double>* __restrict__ beta,
                                                   int bstride, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    const R bt = (R)beta[(size_t)r * bstride].x;
    const R inv = R(1) / bt;
    cplx<R>* o = out + r * ld;
    const cplx<R>* wr = w + r * ld;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        o[i] = inv * wr[i];
}

template <typename R>
void scale_inv(hipStream_t s, cplx<R>* out, const cplx<R>* w, size_t n, size_t ld, int nrhs, uint64_t mask,
               const cplx<double>* beta, int bstride) {
    dim3 grid(ew_blocks(n), nrhs);
    k_scale_inv<R><<<grid, 256, 0, s>>>(out, w, n, ld, mask, beta, bstride, dev_mask());
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ __launch_bounds__(256) void k_axpy(cplx<R>* out, const cplx<R>* x, const cplx<R>* y, R alpha, size_t n,
                                              size_t ld, uint64_t mask, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[r * ld + i] = x[r * ld + i] + alpha * y[r * ld + i];
}

template <typename R>
void axpy(hipStream_t s, cplx<R>* out, const cplx<R>* x, const cplx<R>* y, double alpha, size_t n, size_t ld, int nrhs,
          uint64_t mask) {
    dim3 grid(ew_blocks(n), nrhs);
    k_axpy<R><<<grid, 256, 0, s>>>(out, x, y, (R)alpha, n, ld, mask, dev_mask());
    DD_LAUNCH_CHECK();
}

// out = c0[r] x + c1[r] y + c2[r] z per system r (coefficients c[3r..3r+2] in FP64 on
// the device; a null operand contributes nothing)
template <typename R>
__global__ __launch_bounds__(256) void k_lincomb(cplx<R>* out, const cplx<R>* x, const cplx<R>* y, const cplx<R>* z,
                                                 const cplx<double>* __restrict__ c, size_t n, size_t ld,
                                                 uint64_t mask, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    const cplx<double> a = c[3 * r], b = c[3 * r + 1], d = c[3 * r + 2];
    const cplx<R> ca = mk<R>((R)a.x, (R)a.y), cb = mk<R>((R)b.x, (R)b.y), cd = mk<R>((R)d.x, (R)d.y);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t o = r * ld + i;
        cplx<R> v = czero<R>();
        if (x) v = v + ca * x[o];
        if (y) v = v + cb * y[o];
        if (z) v = v + cd * z[o];
        out[o] = v;
    }
}

template <typename R>
void lincomb(hipStream_t s, cplx<R>* out, const cplx<R>* x, const cplx<R>* y, const cplx<R>* z,
             const cplx<double>* c, size_t n, size_t ld, int nrhs, uint64_t mask) {
    dim3 grid(ew_blocks(n), nrhs);
    k_lincomb<R><<<grid, 256, 0, s>>>(out, x, y, z, c, n, ld, mask, dev_mask());
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ __launch_bounds__(256) void k_update_x(cplx<R>* __restrict__ x, const cplx<R>* __restrict__ Z,
                                                  size_t zstride, const cplx<double>* __restrict__ eta,
                                                  const int* __restrict__ nit, int m, size_t n, size_t ld,
                                                  uint64_t mask, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    const int k = nit[r];
    cplx<R>* xr = x + r * ld;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        cplx<R> v = xr[i];
        for (int j = 0; j < k; j++) {
            const cplx<double> e = eta[(size_t)r * m + j];
            v = v + mk<R>((R)e.x, (R)e.y) * Z[j * zstride + r * ld + i];
        }
        xr[i] = v;
    }
}

template <typename R>
void update_x(hipStream_t s, cplx<R>* x, const cplx<R>* Z, size_t zstride, const cplx<double>* eta, const int* nit,
              int m, size_t n, size_t ld, int nrhs, uint64_t mask) {
    dim3 grid(ew_blocks(n), nrhs);
    k_update_x<R><<<grid, 256, 0, s>>>(x, Z, zstride, eta, nit, m, n, ld, mask, dev_mask());
    DD_LAUNCH_CHECK();
}


// ---- classical Gram-Schmidt passes (used by CGS2 orthogonalisation) --------------------
template <typename R>
__global__ __launch_bounds__(256) void k_multidot(const cplx<R>* __restrict__ w, const cplx<R>* __restrict__ V,
                                                  size_t vstride, int nvec, size_t n, size_t ld, uint64_t mask,
                                                  cplx<double>* __restrict__ partial, int nblk, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    // grid: (nblk * nvec, nrhs); block b = (vector i, chunk c)
    __shared__ R scratch[32];
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    const int i = blockIdx.x / nblk, c = blockIdx.x - i * nblk;
    const cplx<R>* wr = w + r * ld;
    const cplx<R>* vr = V + i * vstride + r * ld;
    cplx<R> s = czero<R>();
    for (size_t e = (size_t)c * blockDim.x + threadIdx.x; e < n; e += (size_t)nblk * blockDim.x)
        s = s + mulc(wr[e], vr[e]);
    s = block_sum(s, scratch);
    if (threadIdx.x == 0) partial[((size_t)r * nvec + i) * nblk + c] = mk<double>(s.x, s.y);
}

template <typename R>
void multidot(hipStream_t s, const cplx<R>* w, const cplx<R>* V, size_t vstride, int nvec, size_t n, size_t ld,
              int nrhs, uint64_t mask, RedBuf& rb) {
    timer_begin(s, K_MULTIDOT);
    rb.nblk = red_blocks(n);
    dim3 grid(rb.nblk * nvec, nrhs);
    k_multidot<R><<<grid, 256, 0, s>>>(w, V, vstride, nvec, n, ld, mask, (cplx<double>*)rb.partial, rb.nblk, dev_mask());
    timer_end(s, K_MULTIDOT, nact(mask, nrhs) * (double)n * (nvec + 1) * sizeof(cplx<R>));
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ __launch_bounds__(256) void k_multi_axpy(cplx<R>* __restrict__ w, const cplx<R>* __restrict__ V,
                                                    size_t vstride, int nvec, size_t n, size_t ld, uint64_t mask,
                                                    const cplx<double>* __restrict__ partial, int nblk,
                                                    cplx<double>* H, int hstride, int accumulate, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    extern __shared__ unsigned char smem_raw[];
    cplx<double>* hs = reinterpret_cast<cplx<double>*>(smem_raw);
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    const int wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int i = wid; i < nvec; i += nw) {
        const cplx<double> h = sum_partials(partial + ((size_t)r * nvec + i) * nblk, nblk);
        if ((threadIdx.x & 63) == 0) hs[i] = h;
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < nvec; i += blockDim.x) {
            cplx<double>* hp = H + (size_t)r * hstride + i;
            *hp = accumulate ? (*hp + hs[i]) : hs[i];
        }
    cplx<R>* wr = w + r * ld;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        cplx<R> v = wr[e];
        for (int i = 0; i < nvec; i++) v = v - mk<R>((R)hs[i].x, (R)hs[i].y) * V[i * vstride + r * ld + e];
        wr[e] = v;
    }
}

template <typename R>
void multi_axpy(hipStream_t s, cplx<R>* w, const cplx<R>* V, size_t vstride, int nvec, size_t n, size_t ld, int nrhs,
                uint64_t mask, const RedBuf& rb, cplx<double>* H, int hstride, bool accumulate) {
    timer_begin(s, K_MULTIAXPY);
    dim3 grid(ew_blocks(n), nrhs);
    const size_t sh = sizeof(cplx<double>) * nvec;
    k_multi_axpy<R><<<grid, 256, sh, s>>>(w, V, vstride, nvec, n, ld, mask, (const cplx<double>*)rb.partial, rb.nblk, H,
                                          hstride, accumulate ? 1 : 0, dev_mask());
    timer_end(s, K_MULTIAXPY, nact(mask, nrhs) * (double)n * (nvec + 2) * sizeof(cplx<R>));
    DD_LAUNCH_CHECK();
}

template <typename To, typename From>
__global__ __launch_bounds__(256) void k_convert(cplx<To>* __restrict__ out, const cplx<From>* __restrict__ in, size_t n,
                                                 size_t ldo, size_t ldi, uint64_t mask, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[r * ldo + i] = cast<To>(in[r * ldi + i]);
}

template <typename R>
void copy(hipStream_t s, cplx<R>* out, const cplx<R>* in, size_t n, size_t ldo, size_t ldi, int nrhs, uint64_t mask) {
    dim3 grid(ew_blocks(n), nrhs);
    k_convert<R, R><<<grid, 256, 0, s>>>(out, in, n, ldo, ldi, mask, dev_mask());
    DD_LAUNCH_CHECK();
}

template <typename To, typename From>
void convert(hipStream_t s, cplx<To>* out, const cplx<From>* in, size_t n, size_t ldo, size_t ldi, int nrhs,
             uint64_t mask) {
    dim3 grid(ew_blocks(n), nrhs);
    k_convert<To, From><<<grid, 256, 0, s>>>(out, in, n, ldo, ldi, mask, dev_mask());
    DD_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_to_half(const float* __restrict__ in, _Float16* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (_Float16)in[i];
}

void to_half(hipStream_t s, const float* in, _Float16* out, size_t n) {
    k_to_half<<<ew_blocks(n), 256, 0, s>>>(in, out, n);
    DD_LAUNCH_CHECK();
}

__global__ __launch_bounds__(256) void k_from_half(const _Float16* __restrict__ in, double* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (double)in[i];
}

void from_half(hipStream_t s, const _Float16* in, double* out, size_t n) {
    k_from_half<<<ew_blocks(n), 256, 0, s>>>(in, out, n);
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ __launch_bounds__(256) void k_zero(cplx<R>* out, size_t n, size_t ld, uint64_t mask, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[r * ld + i] = czero<R>();
}

template <typename R>
void zero(hipStream_t s, cplx<R>* out, size_t n, size_t ld, int nrhs, uint64_t mask) {
    dim3 grid(ew_blocks(n), nrhs);
    k_zero<R><<<grid, 256, 0, s>>>(out, n, ld, mask, dev_mask());
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ __launch_bounds__(256) void k_sigma3(cplx<R>* out, const cplx<R>* in, size_t n, size_t ld, uint64_t mask, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = blockIdx.y;
    if (!rhs_on(mask, r)) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const cplx<R> v = in[r * ld + i];
        out[r * ld + i] = (i & 1) ? -v : v;
    }
}

template <typename R>
void sigma3(hipStream_t s, cplx<R>* out, const cplx<R>* in, size_t n, size_t ld, int nrhs, uint64_t mask) {
    dim3 grid(ew_blocks(n), nrhs);
    k_sigma3<R><<<grid, 256, 0, s>>>(out, in, n, ld, mask, dev_mask());
    DD_LAUNCH_CHECK();
}

__global__ void k_gate(int* gate, const cplx<double>* err2, const cplx<double>* v2, int stride, double tol,
                       uint64_t mask, int nrhs, const uint64_t* __restrict__ dm) {
    if (dm) mask &= *dm;
    const int r = threadIdx.x;
    if (r >= nrhs || !rhs_on(mask, r)) return;
    if (sqrt(err2[r * stride].x) < tol * sqrt(v2[r * stride].x)) gate[r] = 0;
}

__global__ void k_gate_init(int* gate, const uint64_t* __restrict__ dm) {
    const int r = threadIdx.x;
    gate[r] = dm ? (int)((*dm >> r) & 1ull) : 1;
}

void gate_init(hipStream_t s, int* gate) {
    k_gate_init<<<1, 64, 0, s>>>(gate, dev_mask());
    DD_LAUNCH_CHECK();
}

void gate_update(hipStream_t s, int* gate, const cplx<double>* err2, const cplx<double>* v2, int stride, double tol,
                 int nrhs, uint64_t mask) {
    k_gate<<<1, 64, 0, s>>>(gate, err2, v2, stride, tol, mask, nrhs, dev_mask());
    DD_LAUNCH_CHECK();
}

// ======================================================================================
// Quenched U(1) Metropolis (seeded, counter-based random numbers)
// ======================================================================================
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// One sub-sweep: links of direction mu with parity par (of x for mu = 0, of t for mu = 1)
// share no plaquette and are updated together.  theta_P(x) = th0(x) + th1(x+0) - th0(x+1) - th1(x).
__global__ void k_metro(double* th, int Nx, int Nt, double beta, uint64_t key, int mu, int par, double step) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Nx * Nt / 2) return;
    int x, t;
    if (mu == 0) {
        const int hx = Nx / 2;
        t = idx / hx;
        x = 2 * (idx - t * hx) + par;
    } else {
        const int ht = Nt / 2;
        x = idx / ht;
        t = 2 * (idx - x * ht) + par;
    }
    auto T0 = [&](int xx, int tt) -> double {
        xx = (xx + Nx) % Nx;
        tt = (tt + Nt) % Nt;
        return th[(xx * Nt + tt) * 2];
    };
    auto T1 = [&](int xx, int tt) -> double {
        xx = (xx + Nx) % Nx;
        tt = (tt + Nt) % Nt;
        return th[(xx * Nt + tt) * 2 + 1];
    };
    auto P = [&](int xx, int tt) -> double { return T0(xx, tt) + T1(xx, tt + 1) - T0(xx + 1, tt) - T1(xx, tt); };
    const uint64_t h1 = mix64(key ^ ((uint64_t)(x * Nt + t) * 2 + 0));
    const uint64_t h2 = mix64(h1 ^ 0x5851f42d4c957f2dull);
    const double d = step * (2.0 * u01(h1) - 1.0);
    double pa, pb, na, nb;
    if (mu == 0) {   // in P(x,t) with +, in P(x-1,t) with -
        pa = P(x, t);
        pb = P(x - 1, t);
        na = pa + d;
        nb = pb - d;
    } else {         // in P(x,t) with -, in P(x,t-1) with +
        pa = P(x, t);
        pb = P(x, t - 1);
        na = pa - d;
        nb = pb + d;
    }
    const double dS = -beta * (cos(na) + cos(nb) - cos(pa) - cos(pb));
    if (dS <= 0.0 || u01(h2) < exp(-dS)) th[(x * Nt + t) * 2 + mu] += d;
}

__global__ void k_wrap(double* th, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pi = 3.141592653589793;
    double v = fmod(th[i] + pi, 2.0 * pi);
    if (v < 0) v += 2.0 * pi;
    th[i] = v - pi;
}

void metropolis_u1(hipStream_t s, double* theta, int Nx, int Nt, double beta, int sweeps, uint64_t seed, double step) {
    if ((Nx & 1) || (Nt & 1)) throw BadShape("metropolis: lattice extents must be even");
    DD_HIP(hipMemsetAsync(theta, 0, sizeof(double) * Nx * Nt * 2, s));
    const unsigned blocks = cdivu((size_t)Nx * Nt / 2, 256);
    for (int sw = 0; sw < sweeps; sw++)
        for (int mu = 0; mu < 2; mu++)
            for (int par = 0; par < 2; par++) {
                const uint64_t key = (seed * 0x100000001b3ull) ^ ((uint64_t)sw << 20) ^ ((uint64_t)(mu * 2 + par) << 52);
                k_metro<<<blocks, 256, 0, s>>>(theta, Nx, Nt, beta, key, mu, par, step);
            }
    k_wrap<<<cdivu((size_t)Nx * Nt * 2, 256), 256, 0, s>>>(theta, (size_t)Nx * Nt * 2);
    DD_LAUNCH_CHECK();
}

template <typename R>
__global__ void k_theta_to_u(const double* th, cplx<R>* U, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    U[i] = mk<R>((R)cos(th[i]), (R)sin(th[i]));
}

template <typename R>
void theta_to_u(hipStream_t s, const double* theta, cplx<R>* U, size_t n) {
    k_theta_to_u<R><<<cdivu(n, 256), 256, 0, s>>>(theta, U, n);
    DD_LAUNCH_CHECK();
}

// ======================================================================================
// dD/dtheta contraction.  With U = e^{i theta}: the link U_mu(n) enters row n through
// -1/2 (1 - sigma_mu) s+ U x(n+mu) and row n+mu through -1/2 (1 + sigma_mu) s- U^* x(n).
// grad[n][mu] += scale * Re[ y(n)^H (1-s_mu) s i U x(n+mu) + y(n+mu)^H (1+s_mu) s (-i) U^* x(n) ]
// (scale = 1/2 gives dL/dtheta for L = Re<eta, D^-1 b>, y = D^-dagger eta, x = D^-1 b).
// ======================================================================================
__global__ void k_dtheta(const cplx<double>* __restrict__ U, const cplx<double>* __restrict__ y,
                         const cplx<double>* __restrict__ xv, double* grad, int Nx, int Nt, int nrhs, size_t ld,
                         double scale) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Nx * Nt) return;
    const int x = n / Nt, t = n - x * Nt;
    const cplx<double> I = mk(0.0, 1.0);
    for (int mu = 0; mu < 2; mu++) {
        const int m = (mu == 0) ? x * Nt + (t == Nt - 1 ? 0 : t + 1) : (x == Nx - 1 ? 0 : x + 1) * Nt + t;
        const double sg = (mu == 0 && t == Nt - 1) ? -1.0 : 1.0;
        const cplx<double> u = U[2 * n + mu];
        double acc = 0.0;
        for (int r = 0; r < nrhs; r++) {
            const cplx<double>* yr = y + r * ld;
            const cplx<double>* xr = xv + r * ld;
            const cplx<double> iu = sg * (I * u);
            const cplx<double> f0 = iu * xr[2 * m], f1 = iu * xr[2 * m + 1];
            const cplx<double> mi = sg * (-(I * conj(u)));
            const cplx<double> b0 = mi * xr[2 * n], b1 = mi * xr[2 * n + 1];
            cplx<double> fm0, fm1, bp0, bp1;   // (1 - sigma) f, (1 + sigma) b
            if (mu == 0) {
                fm0 = f0 - f1;
                fm1 = f1 - f0;
                bp0 = b0 + b1;
                bp1 = b0 + b1;
            } else {
                fm0 = f0 + times_i(f1);
                fm1 = f1 - times_i(f0);
                bp0 = b0 - times_i(b1);
                bp1 = times_i(b0) + b1;
            }
            const cplx<double> tv = cmul(yr[2 * n], fm0) + cmul(yr[2 * n + 1], fm1) + cmul(yr[2 * m], bp0) +
                                    cmul(yr[2 * m + 1], bp1);
            acc += tv.x;
        }
        grad[2 * n + mu] += scale * acc;
    }
}

void dtheta_contract(hipStream_t s, const cplx<double>* U, const cplx<double>* y, const cplx<double>* x, double* grad,
                     int Nx, int Nt, int nrhs, size_t ld, double scale) {
    k_dtheta<<<cdivu((size_t)Nx * Nt, 256), 256, 0, s>>>(U, y, x, grad, Nx, Nt, nrhs, ld, scale);
    DD_LAUNCH_CHECK();
}

// ======================================================================================
// Explicit instantiations
// ======================================================================================
#define DD_INST(R)                                                                                                  \
    template void wilson<R>(hipStream_t, const cplx<R>*, const cplx<R>*, cplx<R>*, const cplx<R>*, int, int, double, \
                            bool, int, uint64_t, size_t);                                                           \
    template void dirac_links<R>(hipStream_t, const cplx<R>*, cplx<R>*, int, int);                                  \
    template void coarse_apply<R>(hipStream_t, const cplx<R>*, const cplx<R>*, cplx<R>*, const cplx<R>*, int, int,  \
                                  int, double, int, uint64_t, size_t);                                              \
    template void restrict_<R>(hipStream_t, const cplx<R>*, const cplx<R>*, cplx<R>*, int, int, int, int, int, int, \
                               int, uint64_t, size_t, size_t);                                                      \
    template void prolong<R>(hipStream_t, const cplx<R>*, const cplx<R>*, const cplx<R>*, cplx<R>*, int, int, int,  \
                             int, int, int, int, uint64_t, size_t, size_t);                                         \
    template void orthonormalize<R>(hipStream_t, cplx<R>*, int, int, int, int, int, int);                          \
    template void galerkin<R>(hipStream_t, const cplx<R>*, const cplx<R>*, cplx<R>*, int, int, int, int, int, int,  \
                              bool);                                                                                \
    template void transpose_tv<R>(hipStream_t, const cplx<R>*, cplx<R>*, int, int, int, bool);                      \
    template void dot_partial<R>(hipStream_t, const cplx<R>*, const cplx<R>*, size_t, size_t, int, uint64_t,        \
                                 RedBuf&);                                                                          \
    template void dot_final<R>(hipStream_t, const RedBuf&, cplx<double>*, int, int, uint64_t);                      \
    template void mgs_update_dot<R>(hipStream_t, cplx<R>*, const cplx<R>*, const cplx<R>*, size_t, size_t, int,     \
                                    uint64_t, const RedBuf&, RedBuf&, cplx<double>*, int);                          \
    template void norm_scale<R>(hipStream_t, cplx<R>*, const cplx<R>*, size_t, size_t, int, uint64_t,               \
                                const RedBuf&, cplx<double>*, int);                                                 \
    template void scale_inv<R>(hipStream_t, cplx<R>*, const cplx<R>*, size_t, size_t, int, uint64_t,                \
                               const cplx<double>*, int);                                                           \
    template void axpy<R>(hipStream_t, cplx<R>*, const cplx<R>*, const cplx<R>*, double, size_t, size_t, int,       \
                          uint64_t);                                                                                \
    template void lincomb<R>(hipStream_t, cplx<R>*, const cplx<R>*, const cplx<R>*, const cplx<R>*,                 \
                             const cplx<double>*, size_t, size_t, int, uint64_t);                                   \
    template void update_x<R>(hipStream_t, cplx<R>*, const cplx<R>*, size_t, const cplx<double>*, const int*, int,  \
                              size_t, size_t, int, uint64_t);                                                       \
    template void copy<R>(hipStream_t, cplx<R>*, const cplx<R>*, size_t, size_t, size_t, int, uint64_t);            \
    template void zero<R>(hipStream_t, cplx<R>*, size_t, size_t, int, uint64_t);                                    \
    template void sigma3<R>(hipStream_t, cplx<R>*, const cplx<R>*, size_t, size_t, int, uint64_t);                  \
    template void multidot<R>(hipStream_t, const cplx<R>*, const cplx<R>*, size_t, int, size_t, size_t, int,        \
                              uint64_t, RedBuf&);                                                                   \
    template void multi_axpy<R>(hipStream_t, cplx<R>*, const cplx<R>*, size_t, int, size_t, size_t, int, uint64_t,  \
                                const RedBuf&, cplx<double>*, int, bool);                                           \
    template void theta_to_u<R>(hipStream_t, const double*, cplx<R>*, size_t);

DD_INST(float)
DD_INST(double)
template void convert<float, double>(hipStream_t, cplx<float>*, const cplx<double>*, size_t, size_t, size_t, int,
                                     uint64_t);
template void convert<double, float>(hipStream_t, cplx<double>*, const cplx<float>*, size_t, size_t, size_t, int,
                                     uint64_t);
template void convert<double, double>(hipStream_t, cplx<double>*, const cplx<double>*, size_t, size_t, size_t, int,
                                      uint64_t);
template void convert<float, float>(hipStream_t, cplx<float>*, const cplx<float>*, size_t, size_t, size_t, int,
                                    uint64_t);

}  // namespace dd
