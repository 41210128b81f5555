// Host orchestration of the DD-alphaAMG solve (see solver.h).
//
// Reference correspondence (paths relative to the reference root):
//   Krylov::solve            FGMRES::fgmres / rotation / solve_upper_triangular,
//                            src/fgmres.cpp:5-127; hooks include/fgmres.h:86-102
//   Multigrid::sap           SAP_C::SAP, src/sap.cpp:81-165
//   Multigrid::kcycle/vcycle AlgebraicMG::k_cycle / v_cycle, src/amg.cpp:64-174
//   Multigrid::coarsest      GMRES_level_l, include/level.h:70-140
//   Multigrid::setup         AlgebraicMG::setUpPhase, src/amg.cpp:4-62
//   Solver::solve            FGMRES_AMG_{k,v}_cycle, include/amg.h:147-270
#include "poly.h"
#include "solver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace dd {

// ---------------------------------------------------------------------------------------
void DevMem::alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    DD_HIP(hipMalloc(&p_, bytes));
    bytes_ = bytes;
}
void DevMem::release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
}
void PinMem::alloc(size_t bytes) {
    release();
    if (bytes == 0) return;
    DD_HIP(hipHostMalloc(&p_, bytes, hipHostMallocDefault));
}
void PinMem::release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
}

size_t kview_bytes(int m);   // bytes of the device control state of Krylov::solve_dc

__global__ void k_fill_int(int* p, int v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// deterministic pseudo-random complex entries in [-1, 1) (autotuner inputs)
template <typename R>
__global__ void k_fill_hash(cplx<R>* p, size_t n, uint64_t seed) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    p[i] = mk<R>((R)((double)(z >> 40) / 8388608.0 - 1.0), (R)((double)(z & 0xffffff) / 8388608.0 - 1.0));
}

// ---------------------------------------------------------------------------------------
// Krylov engine
// ---------------------------------------------------------------------------------------
template <typename R>
Krylov<R>::Krylov(hipStream_t s, size_t n, int m, int restarts, double tol, int cap, bool flexible, int ortho)
    : s_(s), n_(n), m_(m), restarts_(restarts), cap_(cap), ortho_(ortho), m_max_(m), tol_(tol), flex_(flexible) {
    if (m < 1 || restarts < 1 || cap < 1 || cap > 64) throw BadShape("Krylov: bad m/restarts/batch");
    const size_t vb = sizeof(cplx<R>) * n * cap;
    V_.alloc(vb * (m + 1));
    if (flex_) Z_.alloc(vb * m);
    r_.alloc(vb);
    w_.alloc(vb);
    H_.alloc(sizeof(cplx<double>) * (size_t)m * cap * (m + 1));
    nrm_.alloc(sizeof(cplx<double>) * cap * 4);
    eta_.alloc(sizeof(cplx<double>) * cap * m);
    nit_.alloc(sizeof(int) * cap);
    const size_t nblk = std::min<size_t>(256, std::max<size_t>(1, (n + 2047) / 2048));
    part_.alloc(sizeof(cplx<double>) * cap * (m + 1) * nblk);
    rb_.partial = part_.as<void>();
    Hh_.alloc(sizeof(cplx<double>) * cap * (m + 1));
    nrmh_.alloc(sizeof(cplx<double>) * cap * 4);
    etah_.alloc(sizeof(cplx<double>) * cap * m);
    nith_.alloc(sizeof(int) * cap);
}

template <typename R>
void Krylov<R>::alloc_ctl() {
    if (!ctl_.bytes()) ctl_.alloc(kview_bytes(m_max_));
}

template <typename R>
RedBuf Krylov<R>::ortho(int j, cplx<R>* w, uint64_t run, int nr) {
    const size_t vs = n_ * cap_;
    cplx<double>* Hj = H_.as<cplx<double>>() + (size_t)j * cap_ * (m_ + 1);
    const cplx<R>* V = V_.as<cplx<R>>();
    if (ortho_ == 0) {
        // modified Gram-Schmidt (src/fgmres.cpp:27-38): projection i updates w and forms
        // the partial sums of projection i + 1 (of |w|^2 after the last) in the same pass;
        // the partials alternate between two slots of part_ ([cap][nblk] each)
        const size_t nblk = std::min<size_t>(kMaxRedBlocks, std::max<size_t>(1, (n_ + 2047) / 2048));
        RedBuf pb[2] = {rb_, rb_};
        pb[1].partial = reinterpret_cast<cplx<double>*>(rb_.partial) + (size_t)cap_ * nblk;   // m_ >= 1: 2 slots fit
        dot_partial<R>(s_, w, V, n_, n_, nr, run, pb[0]);
        for (int i = 0; i <= j; i++)
            mgs_update_dot<R>(s_, w, V + i * vs, i < j ? V + (i + 1) * vs : nullptr, n_, n_, nr, run, pb[i & 1],
                              pb[(i + 1) & 1], Hj + i, m_ + 1);
        return pb[(j + 1) & 1];
    }
    multidot<R>(s_, w, V, vs, j + 1, n_, n_, nr, run, rb_);
    multi_axpy<R>(s_, w, V, vs, j + 1, n_, n_, nr, run, rb_, Hj, m_ + 1, false);
    multidot<R>(s_, w, V, vs, j + 1, n_, n_, nr, run, rb_);
    multi_axpy<R>(s_, w, V, vs, j + 1, n_, n_, nr, run, rb_, Hj, m_ + 1, true);
    dot_partial<R>(s_, w, w, n_, n_, nr, run, rb_);
    return rb_;
}

template <typename R>
uint64_t Krylov<R>::solve(const cplx<R>* phi, const cplx<R>* x0, cplx<R>* x, uint64_t mask, const Func& A,
                          const Prec* M, int* iters, double* relres) {
    using C = std::complex<double>;
    const int nr = mask_count(mask);
    if (nr > cap_) throw BadShape("Krylov: batch larger than capacity");
    const size_t n = n_, vs = n_ * cap_;
    cplx<R>* V = V_.as<cplx<R>>();
    cplx<R>* Z = flex_ ? Z_.as<cplx<R>>() : V;
    cplx<R>* r = r_.as<cplx<R>>();
    cplx<R>* w = w_.as<cplx<R>>();
    cplx<double>* nrm = nrm_.as<cplx<double>>();
    cplx<double>* nrmh = nrmh_.as<cplx<double>>();
    // setZeros (include/fgmres.h:116-146): the basis and the Z vectors start at zero
    for (int i = 0; i <= m_; i++) zero<R>(s_, V + i * vs, n, n, nr, mask);
    if (flex_)
        for (int i = 0; i < m_; i++) zero<R>(s_, Z + i * vs, n, n, nr, mask);
    if (x != x0) copy<R>(s_, x, x0, n, n, n, nr, mask);
    A(x, w, mask);
    axpy<R>(s_, r, phi, w, -1.0, n, n, nr, mask);
    dot_partial<R>(s_, phi, phi, n, n, nr, mask, rb_);
    dot_final<R>(s_, rb_, nrm + 0, 4, nr, mask);
    dot_partial<R>(s_, r, r, n, n, nr, mask, rb_);
    dot_final<R>(s_, rb_, nrm + 1, 4, nr, mask);
    DD_HIP(hipMemcpyAsync(nrmh, nrm, sizeof(cplx<double>) * 4 * nr, hipMemcpyDeviceToHost, s_));
    DD_HIP(hipStreamSynchronize(s_));
    std::vector<double> nphi(nr), err(nr);
    for (int q = 0; q < nr; q++) {
        nphi[q] = std::sqrt(nrmh[4 * q].x);
        err[q] = std::sqrt(nrmh[4 * q + 1].x);
    }
    std::vector<int> maxIt(nr, m_);
    std::vector<std::vector<C>> H(nr, std::vector<C>((size_t)(m_ + 1) * m_));   // H[i*m + j]
    std::vector<std::vector<C>> g(nr, std::vector<C>(m_ + 1)), sn(nr, std::vector<C>(m_)),
        cn(nr, std::vector<C>(m_));
    for (int q = 0; q < nr; q++)
        if ((mask >> q) & 1) iters[q] = restarts_ * m_;
    uint64_t done = ~mask;
    // a zero initial residual is a solution already
    for (int q = 0; q < nr; q++)
        if (((mask >> q) & 1) && err[q] == 0.0) {
            iters[q] = 0;
            done |= 1ull << q;
        }
    cplx<double>* Hh = Hh_.as<cplx<double>>();
    cplx<double>* etah = etah_.as<cplx<double>>();
    int* nith = nith_.as<int>();
    for (int k = 0; k < restarts_; k++) {
        const uint64_t live = mask & ~done;
        if (!live) break;
        // V_0 = r / ||r|| from the partial sums of the last residual norm
        norm_scale<R>(s_, V, r, n, n, nr, live, rb_, nrm + 2, 4);
        for (int q = 0; q < nr; q++)
            if ((live >> q) & 1) g[q][0] = C(err[q], 0.0);
        uint64_t run = live;
        for (int j = 0; j < m_ && run; j++) {
            steps++;
            cplx<R>* Vj = V + j * vs;
            cplx<R>* Zj = Z + j * vs;
            if (flex_) (*M)(Vj, Zj, run);
            A(Zj, w, run);
            const RedBuf wn = ortho(j, w, run, nr);
            cplx<double>* Hj = H_.as<cplx<double>>() + (size_t)j * cap_ * (m_ + 1);
            norm_scale<R>(s_, V + (j + 1) * vs, w, n, n, nr, run, wn, Hj + (j + 1), m_ + 1);
            DD_HIP(hipMemcpyAsync(Hh, Hj, sizeof(cplx<double>) * (m_ + 1) * nr, hipMemcpyDeviceToHost, s_));
            DD_HIP(hipStreamSynchronize(s_));
            for (int q = 0; q < nr; q++) {
                if (!((run >> q) & 1)) continue;
                std::vector<C>& h = H[q];
                for (int i = 0; i <= j + 1; i++) h[(size_t)i * m_ + j] = C(Hh[q * (m_ + 1) + i].x, Hh[q * (m_ + 1) + i].y);
                // rotation (src/fgmres.cpp:99-115)
                for (int i = 0; i < j; i++) {
                    const C temp = std::conj(cn[q][i]) * h[i * m_ + j] + std::conj(sn[q][i]) * h[(i + 1) * m_ + j];
                    h[(i + 1) * m_ + j] = -sn[q][i] * h[i * m_ + j] + cn[q][i] * h[(i + 1) * m_ + j];
                    h[i * m_ + j] = temp;
                }
                const C hjj = h[j * m_ + j], hs = h[(j + 1) * m_ + j];
                const C den = std::sqrt(std::real(std::conj(hjj) * hjj + std::conj(hs) * hs));
                sn[q][j] = hs / den;
                cn[q][j] = hjj / den;
                h[j * m_ + j] = std::conj(cn[q][j]) * hjj + std::conj(sn[q][j]) * hs;
                h[(j + 1) * m_ + j] = 0.0;
                g[q][j + 1] = -sn[q][j] * g[q][j];
                g[q][j] = std::conj(cn[q][j]) * g[q][j];
                if (std::abs(g[q][j + 1]) < tol_ * nphi[q]) {
                    maxIt[q] = j + 1;
                    run &= ~(1ull << q);
                }
            }
        }
        // solve_upper_triangular (src/fgmres.cpp:118-127) and x += sum eta_j Z_j
        for (int q = 0; q < nr; q++) {
            nith[q] = 0;
            if (!((live >> q) & 1)) continue;
            const int nI = maxIt[q];
            std::vector<C> eta(nI);
            for (int i = nI - 1; i >= 0; i--) {
                eta[i] = g[q][i];
                for (int jj = i + 1; jj < nI; jj++) eta[i] -= H[q][(size_t)i * m_ + jj] * eta[jj];
                eta[i] /= H[q][(size_t)i * m_ + i];
            }
            for (int jj = 0; jj < nI; jj++) etah[q * m_ + jj] = mk<double>(eta[jj].real(), eta[jj].imag());
            nith[q] = nI;
        }
        DD_HIP(hipMemcpyAsync(eta_.as<void>(), etah, sizeof(cplx<double>) * m_ * nr, hipMemcpyHostToDevice, s_));
        DD_HIP(hipMemcpyAsync(nit_.as<void>(), nith, sizeof(int) * nr, hipMemcpyHostToDevice, s_));
        update_x<R>(s_, x, Z, vs, eta_.as<cplx<double>>(), nit_.as<int>(), m_, n, n, nr, live);
        A(x, w, live);
        axpy<R>(s_, r, phi, w, -1.0, n, n, nr, live);
        dot_partial<R>(s_, r, r, n, n, nr, live, rb_);
        dot_final<R>(s_, rb_, nrm + 1, 4, nr, live);
        DD_HIP(hipMemcpyAsync(nrmh, nrm, sizeof(cplx<double>) * 4 * nr, hipMemcpyDeviceToHost, s_));
        DD_HIP(hipStreamSynchronize(s_));
        for (int q = 0; q < nr; q++) {
            if (!((live >> q) & 1)) continue;
            err[q] = std::sqrt(nrmh[4 * q + 1].x);
            if (err[q] < tol_ * nphi[q]) {
                iters[q] = k * m_ + maxIt[q];
                done |= 1ull << q;
            }
        }
    }
    if (relres)
        for (int q = 0; q < nr; q++)
            if ((mask >> q) & 1) relres[q] = nphi[q] > 0 ? err[q] / nphi[q] : 0.0;
    return mask & done;
}

// ---- device-controlled solve (Krylov::solve_dc) ---------------------------------------
// Control state: the masks of the systems still to solve in this cycle (live), still in
// the Arnoldi loop (run) and finished (done), then per system ||phi||, the last true
// residual norm, maxIt, the iteration count, the least-squares right-hand side g, the
// rotations (cn, sn) and the rotated Hessenberg matrix Hr[i*m + j] (src/fgmres.cpp:5-127).
struct KCtl {
    uint64_t live, run, done, pad;
};
struct KView {
    KCtl* h;
    double *nphi, *err;
    int *maxit, *iters;
    cplx<double> *g, *cn, *sn, *Hr;
    int m;
    __host__ __device__ static size_t bytes(int cap, int m) {
        return sizeof(KCtl) + (size_t)cap * (2 * sizeof(double) + 2 * sizeof(int)) +
               sizeof(cplx<double>) * (size_t)cap * ((m + 1) + 2 * m + (m + 1) * m) + 64;
    }
    __host__ __device__ static KView at(void* base, int cap, int m) {
        KView v;
        unsigned char* p = static_cast<unsigned char*>(base);
        v.h = reinterpret_cast<KCtl*>(p);
        p += sizeof(KCtl);
        v.nphi = reinterpret_cast<double*>(p);
        v.err = v.nphi + cap;
        p += 2 * sizeof(double) * cap;
        v.maxit = reinterpret_cast<int*>(p);
        v.iters = v.maxit + cap;
        p += 2 * sizeof(int) * cap;
        p = reinterpret_cast<unsigned char*>((reinterpret_cast<uintptr_t>(p) + 15) & ~(uintptr_t)15);
        v.g = reinterpret_cast<cplx<double>*>(p);
        v.cn = v.g + (size_t)cap * (m + 1);
        v.sn = v.cn + (size_t)cap * m;
        v.Hr = v.sn + (size_t)cap * m;
        v.m = m;
        return v;
    }
};

// one lane per system: ||phi||, ||phi - A x0||, maxIt = m, iters = restarts*m; a system
// outside the batch (mask, and the caller's device mask) or with a zero residual is done
__global__ void k_kc_init(KView v, const cplx<double>* nrm, uint64_t mask, const uint64_t* parent, int restarts) {
    const int q = threadIdx.x;
    const uint64_t batch = mask & (parent ? *parent : ~0ull);
    const bool in = (batch >> q) & 1ull;
    bool dn = !in;
    if (in) {
        v.nphi[q] = sqrt(nrm[4 * q].x);
        v.err[q] = sqrt(nrm[4 * q + 1].x);
        v.maxit[q] = v.m;
        v.iters[q] = v.err[q] == 0.0 ? 0 : restarts * v.m;
        dn = v.err[q] == 0.0;
    }
    const uint64_t done = __ballot(dn);
    if (q == 0) {
        v.h->done = done;
        v.h->live = 0;
        v.h->run = 0;
    }
}

// cycle start: live = run = the systems not done; g_0 = ||r||
__global__ void k_kc_start(KView v) {
    const int q = threadIdx.x;
    const uint64_t live = ~v.h->done;
    if ((live >> q) & 1ull) v.g[(size_t)q * (v.m + 1)] = mk<double>(v.err[q], 0.0);
    __syncthreads();
    if (q == 0) {
        v.h->live = live;
        v.h->run = live;
    }
}

// Givens rotation of column j (src/fgmres.cpp:99-115) and the break test, per running system
__global__ void k_kc_givens(KView v, const cplx<double>* Hj, int hstride, int j, double tol) {
    const int q = threadIdx.x, m = v.m;
    const bool on = (v.h->run >> q) & 1ull;
    bool stop = false;
    if (on) {
        cplx<double>* cn = v.cn + (size_t)q * m;
        cplx<double>* sn = v.sn + (size_t)q * m;
        cplx<double>* g = v.g + (size_t)q * (m + 1);
        cplx<double>* H = v.Hr + (size_t)q * (m + 1) * m;
        const cplx<double>* col = Hj + (size_t)q * hstride;
        for (int i = 0; i <= j + 1; i++) H[(size_t)i * m + j] = col[i];
        for (int i = 0; i < j; i++) {
            const cplx<double> a = H[(size_t)i * m + j], b = H[(size_t)(i + 1) * m + j];
            const cplx<double> temp = cmul(cn[i], a) + cmul(sn[i], b);
            H[(size_t)(i + 1) * m + j] = cn[i] * b - sn[i] * a;
            H[(size_t)i * m + j] = temp;
        }
        const cplx<double> hjj = H[(size_t)j * m + j], hs = H[(size_t)(j + 1) * m + j];
        const double den = sqrt(norm2(hjj) + norm2(hs));
        sn[j] = mk<double>(hs.x / den, hs.y / den);
        cn[j] = mk<double>(hjj.x / den, hjj.y / den);
        H[(size_t)j * m + j] = cmul(cn[j], hjj) + cmul(sn[j], hs);
        H[(size_t)(j + 1) * m + j] = czero<double>();
        g[j + 1] = czero<double>() - sn[j] * g[j];
        g[j] = cmul(cn[j], g[j]);
        if (sqrt(norm2(g[j + 1])) < tol * v.nphi[q]) {
            v.maxit[q] = j + 1;
            stop = true;
        }
    }
    const uint64_t st = __ballot(stop);
    if (q == 0) v.h->run &= ~st;
}

// back substitution (src/fgmres.cpp:118-127) of every live system: eta [q*m + i], nit [q]
__global__ void k_kc_backsub(KView v, cplx<double>* eta, int* nit, int cap) {
    const int q = threadIdx.x, m = v.m;
    if (q >= cap) return;
    if (!((v.h->live >> q) & 1ull)) {
        nit[q] = 0;
        return;
    }
    const int nI = v.maxit[q];
    const cplx<double>* H = v.Hr + (size_t)q * (m + 1) * m;
    const cplx<double>* g = v.g + (size_t)q * (m + 1);
    cplx<double>* e = eta + (size_t)q * m;
    for (int i = nI - 1; i >= 0; i--) {
        cplx<double> t = g[i];
        for (int jj = i + 1; jj < nI; jj++) t = t - H[(size_t)i * m + jj] * e[jj];
        e[i] = cdiv(t, H[(size_t)i * m + i]);
    }
    nit[q] = nI;
}

// cycle end: the true residual norm of every live system; converged ones are done with
// k*m + maxIt iterations
__global__ void k_kc_end(KView v, const cplx<double>* nrm, int k, double tol) {
    const int q = threadIdx.x;
    bool dn = false;
    if ((v.h->live >> q) & 1ull) {
        v.err[q] = sqrt(nrm[4 * q + 1].x);
        if (v.err[q] < tol * v.nphi[q]) {
            v.iters[q] = k * v.m + v.maxit[q];
            dn = true;
        }
    }
    const uint64_t d = __ballot(dn);
    if (q == 0) v.h->done |= d;
}

size_t kview_bytes(int m) { return KView::bytes(64, m); }

template <typename R>
void Krylov<R>::solve_dc(const cplx<R>* phi, cplx<R>* x, uint64_t mask, const Func& A, const Prec* M) {
    const int nr = mask_count(mask);
    if (nr > cap_) throw BadShape("Krylov: batch larger than capacity");
    const size_t n = n_, vs = n_ * cap_;
    if (!ctl_.bytes()) throw BadShape("Krylov: device control state not allocated (alloc_ctl)");
    const KView kv = KView::at(ctl_.as<void>(), 64, m_max_);
    KView v = kv;
    v.m = m_;
    cplx<R>* V = V_.as<cplx<R>>();
    cplx<R>* Z = flex_ ? Z_.as<cplx<R>>() : V;
    cplx<R>* r = r_.as<cplx<R>>();
    cplx<R>* w = w_.as<cplx<R>>();
    cplx<double>* nrm = nrm_.as<cplx<double>>();
    const uint64_t* parent = dev_mask();   // the systems the caller runs (nullptr: all of mask)
    // setZeros (include/fgmres.h:116-146), x0 = phi, r = phi - A x0, ||phi||, ||r||
    for (int i = 0; i <= m_; i++) zero<R>(s_, V + i * vs, n, n, nr, mask);
    if (flex_)
        for (int i = 0; i < m_; i++) zero<R>(s_, Z + i * vs, n, n, nr, mask);
    copy<R>(s_, x, phi, n, n, n, nr, mask);
    A(x, w, mask);
    axpy<R>(s_, r, phi, w, -1.0, n, n, nr, mask);
    dot_partial<R>(s_, phi, phi, n, n, nr, mask, rb_);
    dot_final<R>(s_, rb_, nrm + 0, 4, nr, mask);
    dot_partial<R>(s_, r, r, n, n, nr, mask, rb_);
    dot_final<R>(s_, rb_, nrm + 1, 4, nr, mask);
    k_kc_init<<<1, 64, 0, s_>>>(v, nrm, mask, parent, restarts_);
    DD_HIP(hipGetLastError());
    for (int k = 0; k < restarts_; k++) {
        k_kc_start<<<1, 64, 0, s_>>>(v);
        DD_HIP(hipGetLastError());
        {   // V_0 = r / ||r|| from the partial sums of the last residual norm
            DevMaskScope sc(&v.h->live);
            norm_scale<R>(s_, V, r, n, n, nr, mask, rb_, nrm + 2, 4);
        }
        for (int j = 0; j < m_; j++) {
            cplx<double>* Hj = H_.as<cplx<double>>() + (size_t)j * cap_ * (m_ + 1);
            {
                DevMaskScope sc(&v.h->run);
                cplx<R>* Vj = V + j * vs;
                cplx<R>* Zj = Z + j * vs;
                if (flex_) (*M)(Vj, Zj, mask);
                A(Zj, w, mask);
                const RedBuf wn = ortho(j, w, mask, nr);
                norm_scale<R>(s_, V + (j + 1) * vs, w, n, n, nr, mask, wn, Hj + (j + 1), m_ + 1);
            }
            k_kc_givens<<<1, 64, 0, s_>>>(v, Hj, m_ + 1, j, tol_);
            DD_HIP(hipGetLastError());
        }
        k_kc_backsub<<<1, 64, 0, s_>>>(v, eta_.as<cplx<double>>(), nit_.as<int>(), cap_);
        DD_HIP(hipGetLastError());
        {
            DevMaskScope sc(&v.h->live);
            update_x<R>(s_, x, Z, vs, eta_.as<cplx<double>>(), nit_.as<int>(), m_, n, n, nr, mask);
            A(x, w, mask);
            axpy<R>(s_, r, phi, w, -1.0, n, n, nr, mask);
            dot_partial<R>(s_, r, r, n, n, nr, mask, rb_);
            dot_final<R>(s_, rb_, nrm + 1, 4, nr, mask);
        }
        k_kc_end<<<1, 64, 0, s_>>>(v, nrm, k, tol_);
        DD_HIP(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------------------
// Geometry and test-vector streams
// ---------------------------------------------------------------------------------------
std::vector<LevelGeom> make_geometry(int Nx, int Nt, const dd_params& p) {
    if (p.levels < 2 || p.levels > DD_MAX_LEVELS) throw BadShape("levels must be in [2, 8]");
    std::vector<LevelGeom> g(p.levels);
    g[0].Nx = Nx;
    g[0].Nt = Nt;
    g[0].dof = 2;
    for (int l = 0; l < p.levels; l++) {
        if (l < p.levels - 1) {
            g[l].ax = p.agg_x[l];
            g[l].at = p.agg_t[l];
            g[l].nv = p.nv[l];
            g[l].bx = p.sap_x[l];
            g[l].bt = p.sap_t[l];
            if (g[l].ax < 1 || g[l].at < 1 || g[l].Nx % g[l].ax || g[l].Nt % g[l].at)
                throw BadShape("aggregate size must divide the level lattice");
            if (g[l].bx < 1 || g[l].bt < 1 || g[l].Nx % g[l].bx || g[l].Nt % g[l].bt)
                throw BadShape("SAP block size must divide the level lattice");
            const int nbx = g[l].Nx / g[l].bx, nbt = g[l].Nt / g[l].bt;
            if ((nbx * nbt) % 2 || ((nbt & 1) && (nbx & 1))) throw BadShape("SAP needs an even number of blocks");
            if (g[l].nv < 1 || 2 * g[l].nv > 32) throw BadShape("test vectors per level must be in [1, 16]");
            g[l + 1].Nx = g[l].Nx / g[l].ax;
            g[l + 1].Nt = g[l].Nt / g[l].at;
            g[l + 1].dof = 2 * g[l].nv;
        }
    }
    return g;
}

RankStreams::RankStreams(int Rx, int Rt) : Rx_(Rx), Rt_(Rt) {
    if (Rx < 1 || Rt < 1) throw BadShape("virtual rank grid must be positive");
    for (int r = 0; r < Rx * Rt; r++) g_.emplace_back((std::mt19937::result_type)r);
}

void RankStreams::fill(std::complex<double>* out, int first, int count, int Nx, int Nt, int dof) {
    if (Nx % Rx_ || Nt % Rt_) throw BadShape("virtual rank grid must divide the level lattice");
    const int wx = Nx / Rx_, wt = Nt / Rt_;
    std::uniform_real_distribution<double> d(-1.0, 1.0);
    const size_t V = (size_t)Nx * Nt;
    for (int cx = 0; cx < Rx_; cx++)
        for (int ct = 0; ct < Rt_; ct++) {
            std::mt19937& g = g_[cx * Rt_ + ct];
            for (int cc = first; cc < first + count; cc++)
                for (int lx = 0; lx < wx; lx++)
                    for (int lt = 0; lt < wt; lt++) {
                        const size_t nsite = (size_t)(cx * wx + lx) * Nt + ct * wt + lt;
                        for (int k = 0; k < dof; k++) {
                            const double im = d(g);
                            const double re = d(g);
                            out[((size_t)cc * V + nsite) * dof + k] = std::complex<double>(re, im);
                        }
                    }
        }
}

// ---------------------------------------------------------------------------------------
// Multigrid hierarchy
// ---------------------------------------------------------------------------------------
template <typename R>
Multigrid<R>::Multigrid(hipStream_t s, int Nx, int Nt, const dd_params& p, const cplx<R>* U)
    : s_(s), p_(p), g_(make_geometry(Nx, Nt, p)), U_(U) {
    const int L = (int)g_.size();
    const int cap = std::max(1, p.max_rhs);
    if (cap > 64) throw BadShape("max_rhs must be <= 64");
    cap_.resize(L);
    links_.resize(L);
    hlinks_.resize(L);
    tvraw_.resize(L);
    W_.resize(L);
    res_.resize(L);
    eta_.resize(L);
    psi_.resize(L);
    sapr_.resize(L);
    gate_.resize(L);
    nrm_.resize(L);
    rb_.resize(L);
    part_.resize(L);
    kry_.resize(L);
    for (int l = 0; l < L; l++) {
        const LevelGeom& G = g_[l];
        cap_[l] = (l < L - 1) ? std::max(cap, G.nv) : cap;
        const size_t vb = sizeof(cplx<R>) * G.n();
        res_[l].alloc(vb * cap_[l]);
        sapr_[l].alloc(vb * cap_[l]);
        gate_[l].alloc(sizeof(int) * 64);
        nrm_[l].alloc(sizeof(cplx<double>) * 2 * 64);
        part_[l].alloc(sizeof(cplx<double>) * 64 * 256);
        rb_[l].partial = part_[l].as<void>();
        if (l >= 1) {
            eta_[l].alloc(vb * cap);
            psi_[l].alloc(vb * cap);
            links_[l].alloc(sizeof(cplx<R>) * (size_t)G.V() * 5 * G.dof * G.dof);
            // binary16 link copy of an FP32 level smoothed by the matrix-core coarse SAP
            if (sizeof(R) == 4 && p.coarse_sap_links == 1 && l < L - 1 && sapc_supported<R>(32, G.dof, G.bx * G.bt))
                hlinks_[l].alloc(sizeof(_Float16) * 2 * (size_t)G.V() * 5 * G.dof * G.dof);
            kry_[l].reset(new Krylov<R>(s, G.n(), p.kcycle_m, p.kcycle_restarts, p.kcycle_tol, cap, true, p.ortho_inner));
            if (p.graph) kry_[l]->alloc_ctl();   // device-controlled solves of captured K-cycles
        }
        if (l < L - 1) W_[l].alloc(sizeof(cplx<R>) * (size_t)G.V() * G.nv * G.dof);
    }
    // fused SAP iterations: state for every coarse level that the matrix cores smooth
    fuse_done_.resize(L);
    fuse_ctr_.resize(L);
    fuse_part_.resize(L);
    fuse_epoch_.assign(L, 0u);
    fuse_ok_.assign(L, 0);
    if (sizeof(R) == 4 && p.sap_fused) {
        int dev = 0, ncu = 0;
        DD_HIP(hipGetDevice(&dev));
        DD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        fuse_slices_ = std::max(1, std::min(ncu, 1024));
        fuse_fail_.alloc(sizeof(unsigned));
        DD_HIP(hipMemsetAsync(fuse_fail_.as<void>(), 0, sizeof(unsigned), s));
        for (int l = 1; l < L - 1; l++) {
            const LevelGeom& G = g_[l];
            const int nbx = G.Nx / G.bx, nbt = G.Nt / G.bt;
            if ((nbx & 1) || (nbt & 1) || !sapc_supported<R>(32, G.dof, G.bx * G.bt)) continue;
            fuse_done_[l].alloc(sizeof(unsigned) * (size_t)nbx * nbt * 4);
            DD_HIP(hipMemsetAsync(fuse_done_[l].as<void>(), 0, fuse_done_[l].bytes(), s));
            fuse_ctr_[l].alloc(sizeof(unsigned) * 2);
            fuse_part_[l].alloc(sizeof(float) * 64 * (size_t)fuse_slices_);
            fuse_ok_[l] = 1;
        }
        DD_HIP(hipStreamSynchronize(s));
    }
    if (p.coarse_poly < 0 || p.coarse_poly > 32) throw BadShape("coarse_poly: degree outside [0, 32]");
    if (p.coarse_poly > 0) {
        const LevelGeom& Gc = g_[L - 1];
        poly_coef_.alloc(sizeof(cplx<double>) * (size_t)p.coarse_poly * 3 * 64 * 3);
        poly_r_.alloc(sizeof(cplx<R>) * Gc.n() * cap_[L - 1]);
        poly_t_.alloc(sizeof(cplx<R>) * Gc.n() * cap_[L - 1]);
    }
    if (p.ortho_coarse >= 2) {
        const LevelGeom& Gc = g_[L - 1];
        coarse_dev_.reset(new CoarsestGmres<R>(s, Gc.Nx, Gc.Nt, Gc.dof, p.coarse_m, p.coarse_restarts, p.coarse_tol,
                                               cap, p.coarse_basis == 1));
        if (!coarse_dev_->supported()) coarse_dev_.reset();
    }
    // launch-per-step engine: always for 0/1, as the alternative the setup times for 3
    if (!coarse_dev_ || p.ortho_coarse == 3)
        coarse_.reset(new Krylov<R>(s, g_[L - 1].n(), p.coarse_m, p.coarse_restarts, p.coarse_tol, cap,
                                    p.coarse_poly > 0, p.ortho_coarse == 0 ? 0 : 1));
    use_dev_ = (bool)coarse_dev_;
    iters_scratch_.resize(64);
}

template <typename R>
void Multigrid<R>::apply(int l, const cplx<R>* in, cplx<R>* out, const cplx<R>* b, uint64_t mask) {
    const LevelGeom& G = g_[l];
    const int nr = mask_count(mask);
    if (l == 0) wilson<R>(s_, U_, in, out, b, G.Nx, G.Nt, p_.m0, false, nr, mask, G.n());
    else coarse_apply<R>(s_, links(l), in, out, b, G.Nx, G.Nt, G.dof, p_.m0, nr, mask, G.n());
}

template <typename R>
void Multigrid<R>::restrict_(int l, const cplx<R>* in, cplx<R>* out, uint64_t mask) {
    const LevelGeom& G = g_[l];
    dd::restrict_<R>(s_, W(l), in, out, G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv, mask_count(mask), mask, G.n(),
                     g_[l + 1].n());
}

template <typename R>
void Multigrid<R>::prolong(int l, const cplx<R>* vc, const cplx<R>* y, cplx<R>* out, uint64_t mask) {
    const LevelGeom& G = g_[l];
    dd::prolong<R>(s_, W(l), vc, y, out, G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv, mask_count(mask), mask, G.n(),
                   g_[l + 1].n());
}

template <typename R>
void Multigrid<R>::sap(int l, const cplx<R>* v, cplx<R>* x, int nu, double tol, uint64_t mask) {
    const LevelGeom& G = g_[l];
    const int nr = mask_count(mask);
    if (nr > cap_[l]) throw BadShape("SAP: batch larger than capacity");
    cplx<R>* r = sapr_[l].as<cplx<R>>();
    int* gate = gate_[l].as<int>();
    cplx<double>* nrm = nrm_[l].as<cplx<double>>();
    const size_t n = G.n();
    gate_init(s_, gate);   // open for the systems of the mask (and of the device mask, if any)
    dot_partial<R>(s_, v, v, n, n, nr, mask, rb_[l]);
    dot_final<R>(s_, rb_[l], nrm + 0, 2, nr, mask);
    // one launch per iteration where the level allows it (not inside a graph capture: the
    // launch epoch is an argument that a replay would repeat)
    if (l > 0 && fuse_ok_[l] && !capturing_) {
        SapFuse fz;
        fz.done = fuse_done_[l].as<unsigned>();
        fz.ctr = fuse_ctr_[l].as<unsigned>();
        fz.fail = fuse_fail_.as<unsigned>();
        fz.part = fuse_part_[l].as<float>();
        fz.gate = gate;
        fz.vnorm = nrm;
        fz.tol = tol;
        const int var = l < (int)sapc_var_.size() ? sapc_var_[l] : -1;
        bool ran = true;
        for (int i = 0; i < nu && ran; i++) {
            k_fill_int<<<1, 64, 0, s_>>>(reinterpret_cast<int*>(fz.ctr), 0, 2);
            DD_HIP(hipGetLastError());
            fz.epoch = ++fuse_epoch_[l];
            // the last iteration's residual norm would feed no stopping test: no norm stage
            ran = sap_fused_iteration<R>(s_, links(l), hlinks(l), v, x, G.Nx, G.Nt, G.dof, G.bx, G.bt, p_.m0, p_.sap_m,
                                         p_.sap_restarts, p_.sap_tol, nr, mask, n, var, fz,
                                         i + 1 < nu ? fuse_slices_ : 0);
            if (!ran && i > 0) throw BadShape("fused SAP became unavailable between iterations");
        }
        if (ran) return;
        fuse_ok_[l] = 0;   // this level's variant has no fused form: separate launches from now on
    }
    // Level 0 with a proper red-black colouring (even block counts): each colour's blocks
    // form v - D x at their own sites, so the full residual is needed only after the black
    // blocks, for the stopping test (src/sap.cpp:115-156 forms it after both colours).
    // After the last iteration's black blocks it would feed only a stopping test that no
    // iteration follows, so it is not formed there (nor its norm).
    const bool fr = l == 0 && (G.Nx / G.bx) % 2 == 0 && (G.Nt / G.bt) % 2 == 0;
    if (!fr) apply(l, x, r, v, mask);
    for (int i = 0; i < nu; i++) {
        const bool last = i + 1 == nu;
        for (int colour = 0; colour < 2; colour++) {
            if (l == 0)
                sap_blocks0<R>(s_, U_, fr ? v : r, x, G.Nx, G.Nt, G.bx, G.bt, colour, p_.m0, p_.sap_m,
                               p_.sap_restarts, p_.sap_tol, nr, mask, n, gate, fr);
            else
                sap_blocksc<R>(s_, links(l), r, x, G.Nx, G.Nt, G.dof, G.bx, G.bt, colour, p_.m0, p_.sap_m,
                               p_.sap_restarts, p_.sap_tol, nr, mask, n, gate,
                               l < (int)sapc_var_.size() ? sapc_var_[l] : -1, hlinks(l));
            if (colour == 0 ? !fr : !last) apply(l, x, r, v, mask);
        }
        if (last) break;
        dot_partial<R>(s_, r, r, n, n, nr, mask, rb_[l]);
        dot_final<R>(s_, rb_[l], nrm + 1, 2, nr, mask);
        gate_update(s_, gate, nrm + 1, nrm + 0, 2, tol, nr, mask);
    }
}

// Start vector of the Arnoldi run that fixes the coarsest polynomial preconditioner: the
// hash fill of this seed (k_fill_hash), the same on every run.
constexpr uint64_t kPolySeed = 4242;

// The roots of the GMRES polynomial preconditioner (poly.h) from p.coarse_poly Arnoldi
// steps of the coarsest operator, one system: the operator is applied on the device in
// the hierarchy's precision to the basis vector rounded to it, and the Gram-Schmidt
// projections (two passes) run on the host in FP64.
template <typename R>
void Multigrid<R>::poly_build() {
    const int d = p_.coarse_poly;
    poly_theta_.clear();
    if (d <= 0) return;
    const int L = levels();
    const size_t n = g_[L - 1].n();
    cplx<R>* v = poly_r_.as<cplx<R>>();
    cplx<R>* w = poly_t_.as<cplx<R>>();
    k_fill_hash<R><<<(unsigned)((n + 255) / 256), 256, 0, s_>>>(v, n, kPolySeed);
    DD_HIP(hipGetLastError());
    std::vector<cplx<R>> hb(n);
    auto down = [&](const cplx<R>* src, std::vector<Cd>& dst) {
        DD_HIP(hipMemcpyAsync(hb.data(), src, sizeof(cplx<R>) * n, hipMemcpyDeviceToHost, s_));
        DD_HIP(hipStreamSynchronize(s_));
        for (size_t i = 0; i < n; i++) dst[i] = Cd((double)hb[i].x, (double)hb[i].y);
    };
    auto norm = [&](const std::vector<Cd>& x) {
        double a = 0;
        for (const Cd& z : x) a += std::norm(z);
        return std::sqrt(a);
    };
    std::vector<std::vector<Cd>> Q(d + 1, std::vector<Cd>(n));
    std::vector<Cd> H((size_t)(d + 1) * d, 0.0), y(n);
    down(v, Q[0]);
    const double n0 = norm(Q[0]);
    for (Cd& z : Q[0]) z /= n0;
    for (int j = 0; j < d; j++) {
        for (size_t i = 0; i < n; i++) hb[i] = mk<R>((R)Q[j][i].real(), (R)Q[j][i].imag());
        DD_HIP(hipMemcpyAsync(v, hb.data(), sizeof(cplx<R>) * n, hipMemcpyHostToDevice, s_));
        apply(L - 1, v, w, nullptr, 1ull);
        down(w, y);
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i <= j; i++) {
                Cd h = 0.0;
                for (size_t k = 0; k < n; k++) h += std::conj(Q[i][k]) * y[k];
                for (size_t k = 0; k < n; k++) y[k] -= h * Q[i][k];
                H[(size_t)i * d + j] += h;
            }
        const double beta = norm(y);
        if (!(beta > 0.0)) throw BadShape("coarse_poly: Arnoldi breakdown (degree above the operator's grade)");
        H[(size_t)(j + 1) * d + j] = beta;
        for (size_t k = 0; k < n; k++) Q[j + 1][k] = y[k] / beta;
    }
    try {
        poly_theta_ = gmres_poly_roots(H, d);
    } catch (const std::runtime_error&) {
        throw BadShape("coarse_poly: no harmonic Ritz values (singular Hessenberg matrix or no QR convergence)");
    }
    // lincomb coefficients per root i and system r: form 0 (1/theta_i, 0, 0),
    // form 1 (1, 1/theta_i, 0), form 2 (1, -1/theta_i, 0)
    std::vector<cplx<double>> c((size_t)d * 3 * 64 * 3, mk<double>(0.0, 0.0));
    for (int i = 0; i < d; i++) {
        const Cd inv = 1.0 / poly_theta_[i];
        for (int r = 0; r < 64; r++) {
            cplx<double>* f0 = &c[((size_t)(i * 3 + 0) * 64 + r) * 3];
            cplx<double>* f1 = &c[((size_t)(i * 3 + 1) * 64 + r) * 3];
            cplx<double>* f2 = &c[((size_t)(i * 3 + 2) * 64 + r) * 3];
            f0[0] = mk<double>(inv.real(), inv.imag());
            f1[0] = mk<double>(1.0, 0.0);
            f1[1] = mk<double>(inv.real(), inv.imag());
            f2[0] = mk<double>(1.0, 0.0);
            f2[1] = mk<double>(-inv.real(), -inv.imag());
        }
    }
    DD_HIP(hipMemcpyAsync(poly_coef_.as<void>(), c.data(), sizeof(cplx<double>) * c.size(), hipMemcpyHostToDevice,
                          s_));
    DD_HIP(hipStreamSynchronize(s_));
    if (coarse_dev_) coarse_dev_->set_poly(poly_theta_);
}

// out = p(D_c) in by the product form: out = in/theta_0, r = in; for i < d - 1:
// r -= D_c r / theta_i, out += r / theta_{i+1}
template <typename R>
void Multigrid<R>::poly_apply(const cplx<R>* in, cplx<R>* out, uint64_t mask) {
    const int L = levels(), d = (int)poly_theta_.size();
    const size_t n = g_[L - 1].n();
    const int nr = mask_count(mask);
    if (nr > cap_[L - 1]) throw BadShape("coarse_poly: batch larger than capacity");
    cplx<R>* r = poly_r_.as<cplx<R>>();
    cplx<R>* t = poly_t_.as<cplx<R>>();
    const cplx<double>* c = poly_coef_.as<cplx<double>>();
    auto cf = [&](int i, int f) { return c + (size_t)(i * 3 + f) * 64 * 3; };
    lincomb<R>(s_, out, in, nullptr, nullptr, cf(0, 0), n, n, nr, mask);
    const cplx<R>* cur = in;
    for (int i = 0; i + 1 < d; i++) {
        apply(L - 1, cur, t, nullptr, mask);
        lincomb<R>(s_, r, cur, t, nullptr, cf(i, 2), n, n, nr, mask);
        lincomb<R>(s_, out, out, r, nullptr, cf(i + 1, 1), n, n, nr, mask);
        cur = r;
    }
}

template <typename R>
void Multigrid<R>::coarsest(const cplx<R>* eta, cplx<R>* psi, uint64_t mask) {
    const int L = levels();
    if (use_dev_) {
        coarse_dev_->deferred = capturing_;
        try {
            coarse_dev_->solve(links(L - 1), p_.m0, eta, psi, mask, iters_scratch_.data());
        } catch (const HipFailure& e) {
            if (capturing_) throw;   // the capture is abandoned; the caller runs eagerly
            // the runtime refused the cooperative launch: continue with the
            // launch-per-step engine for the rest of this hierarchy's life
            (void)hipGetLastError();
            if (std::getenv("DDAMG_VERBOSE"))
                fprintf(stderr, "coarsest: persistent kernel unavailable (%s), launch-per-step engine\n",
                        hipGetErrorString(e.err));
            use_dev_ = false;
            if (!coarse_)
                coarse_.reset(new Krylov<R>(s_, g_[L - 1].n(), p_.coarse_m, p_.coarse_restarts, p_.coarse_tol,
                                            std::max(1, p_.max_rhs), p_.coarse_poly > 0, 1));
        }
    }
    if (capturing_) return;   // no host read-back inside a capture
    if (!use_dev_) {
        auto A = [this, L](const cplx<R>* in, cplx<R>* out, uint64_t m) { apply(L - 1, in, out, nullptr, m); };
        typename Krylov<R>::Func fA = A;
        typename Krylov<R>::Prec fM;
        if (p_.coarse_poly > 0) {
            const size_t n = g_[L - 1].n();
            if (!poly_theta_.empty())
                fM = [this](const cplx<R>* in, cplx<R>* out, uint64_t m) { poly_apply(in, out, m); };
            else   // no roots yet (the coarsest operator was never built): the identity
                fM = [this, n](const cplx<R>* in, cplx<R>* out, uint64_t m) {
                    copy<R>(s_, out, in, n, n, n, mask_count(m), m);
                };
        }
        coarse_->solve(eta, eta, psi, mask, fA, p_.coarse_poly > 0 ? &fM : nullptr, iters_scratch_.data(), nullptr);
    }
    int mx = 0;
    for (int q = 0; q < mask_count(mask); q++)
        if ((mask >> q) & 1) mx = std::max(mx, iters_scratch_[q]);
    coarse_iters.push_back(mx);
}

template <typename R>
Multigrid<R>::~Multigrid() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
}

template <typename R>
bool Multigrid<R>::vcycle_graph(const cplx<R>* eta, cplx<R>* psi, uint64_t mask) {
    return capture_(eta, psi, mask, false);
}

template <typename R>
bool Multigrid<R>::kcycle_graph(const cplx<R>* eta, cplx<R>* psi, uint64_t mask) {
    return capture_(eta, psi, mask, true);
}

// psi = cycle(eta) from level 0 with zero initial guess, replayed from a HIP graph that is
// captured on the first call for these buffers, this mask and this cycle.  A K-cycle is
// captured with its FGMRES solves device-controlled (every step issued, the systems masked
// on the device), so that the graph holds no host decision.
template <typename R>
bool Multigrid<R>::capture_(const cplx<R>* eta, cplx<R>* psi, uint64_t mask, bool kcyc) {
    if (!graphs_ok_ || !use_dev_ || timer_any()) return false;
    const auto key = std::make_tuple(static_cast<const void*>(eta), static_cast<const void*>(psi), mask, kcyc);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        DD_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
        dc_ = kcyc;
        bool ok = true;
        try {
            if (kcyc) kcycle(0, eta, psi, mask, true);   // zero initial guess
            else vcycle(0, eta, psi, mask, true);
        } catch (...) {
            ok = false;
        }
        capturing_ = false;
        dc_ = false;
        const hipError_t e = hipStreamEndCapture(s_, &g);
        ok = ok && e == hipSuccess && g != nullptr && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        if (!ok) {
            (void)hipGetLastError();
            if (ex) (void)hipGraphExecDestroy(ex);
            graphs_ok_ = false;
            if (std::getenv("DDAMG_VERBOSE")) fprintf(stderr, "cycle graph capture refused; eager cycles\n");
            return false;
        }
        it = graphs_.emplace(key, ex).first;
    }
    DD_HIP(hipGraphLaunch(it->second, s_));
    return true;
}

template <typename R>
void Multigrid<R>::check_graph_failures() {
    if (coarse_dev_) coarse_dev_->check_failure();
}

template <typename R>
void Multigrid<R>::check_sap_failures() {
    if (!fuse_fail_.bytes()) return;
    unsigned f = 0;
    DD_HIP(hipMemcpyAsync(&f, fuse_fail_.as<void>(), sizeof(unsigned), hipMemcpyDeviceToHost, s_));
    DD_HIP(hipStreamSynchronize(s_));
    if (f) {
        DD_HIP(hipMemsetAsync(fuse_fail_.as<void>(), 0, sizeof(unsigned), s_));
        throw BadShape("SAP: a wait inside a fused iteration timed out; results of this call are invalid");
    }
}

template <typename R>
void Multigrid<R>::kcycle(int l, const cplx<R>* eta, cplx<R>* psi, uint64_t mask, bool zero_guess) {
    const int L = levels();
    // a zero initial guess is stored only where it is read before psi is written: by the
    // coarsest solve and by the pre-smoothing; otherwise the prolongation writes psi
    const bool fresh = zero_guess && p_.nu1 == 0;
    if (zero_guess && (!fresh || l == L - 1))
        zero<R>(s_, psi, g_[l].n(), g_[l].n(), mask_count(mask), mask);
    if (l == L - 1) {
        coarsest(eta, psi, mask);
        return;
    }
    if (p_.nu1 > 0) sap(l, eta, psi, p_.nu1, 1e-10, mask);
    cplx<R>* r = res_[l].as<cplx<R>>();
    cplx<R>* e1 = eta_[l + 1].template as<cplx<R>>();
    cplx<R>* p1 = psi_[l + 1].template as<cplx<R>>();
    if (fresh) {
        restrict_(l, eta, e1, mask);   // eta - D 0 = eta, bit for bit
    } else {
        apply(l, psi, r, eta, mask);
        restrict_(l, r, e1, mask);
    }
    typename Krylov<R>::Func fA = [this, l](const cplx<R>* in, cplx<R>* out, uint64_t m) {
        apply(l + 1, in, out, nullptr, m);
    };
    typename Krylov<R>::Prec fM = [this, l](const cplx<R>* in, cplx<R>* io, uint64_t m) { kcycle(l + 1, in, io, m); };
    if (dc_) {
        kry_[l + 1]->solve_dc(e1, p1, mask, fA, &fM);
    } else {
        std::vector<int> its(64);
        kry_[l + 1]->solve(e1, e1, p1, mask, fA, &fM, its.data(), nullptr);
    }
    prolong(l, p1, fresh ? nullptr : psi, psi, mask);   // psi = P p1 (+ psi)
    if (p_.nu2 > 0) sap(l, eta, psi, p_.nu2, 1e-10, mask);
}

template <typename R>
void Multigrid<R>::vcycle(int l, const cplx<R>* eta, cplx<R>* psi, uint64_t mask, bool zero_guess) {
    const int L = levels();
    // as in kcycle: a zero initial guess is stored only where it is read first
    const bool fresh = zero_guess && p_.nu1 == 0;
    if (zero_guess && (!fresh || l == L - 1))
        zero<R>(s_, psi, g_[l].n(), g_[l].n(), mask_count(mask), mask);
    if (l == L - 1) {
        coarsest(eta, psi, mask);
        return;
    }
    if (p_.nu1 > 0) sap(l, eta, psi, p_.nu1, 1e-10, mask);
    cplx<R>* r = res_[l].as<cplx<R>>();
    cplx<R>* e1 = eta_[l + 1].template as<cplx<R>>();
    cplx<R>* p1 = psi_[l + 1].template as<cplx<R>>();
    if (fresh) {
        restrict_(l, eta, e1, mask);   // eta - D 0 = eta, bit for bit
    } else {
        apply(l, psi, r, eta, mask);
        restrict_(l, r, e1, mask);
    }
    vcycle(l + 1, e1, p1, mask, true);
    prolong(l, p1, fresh ? nullptr : psi, psi, mask);   // psi = P p1 (+ psi)
    if (p_.nu2 > 0) sap(l, eta, psi, p_.nu2, 1e-10, mask);
}

template <typename R>
void Multigrid<R>::setup(const cplx<R>* tv0, RankStreams& rng) {
    const int L = levels();
    std::vector<DevMem> raw(L);
    for (int l = 0; l < L - 1; l++) raw[l].alloc(sizeof(cplx<R>) * g_[l].nv * g_[l].n());
    DD_HIP(hipMemcpyAsync(raw[0].as<void>(), tv0, sizeof(cplx<R>) * g_[0].nv * g_[0].n(), hipMemcpyDeviceToDevice, s_));
    // v_l = P^dagger_{l-1} v_{l-1} of the raw (unsmoothed) vectors, extra vectors random
    for (int l = 1; l < L - 1; l++) {
        const LevelGeom& Gp = g_[l - 1];
        const LevelGeom& G = g_[l];
        DevMem Wraw(sizeof(cplx<R>) * Gp.V() * Gp.nv * Gp.dof);
        transpose_tv<R>(s_, raw[l - 1].as<cplx<R>>(), Wraw.as<cplx<R>>(), Gp.V(), Gp.nv, Gp.dof, true);
        const int c = std::min(G.nv, Gp.nv);
        dd::restrict_<R>(s_, Wraw.as<cplx<R>>(), raw[l - 1].as<cplx<R>>(), raw[l].as<cplx<R>>(), Gp.Nx, Gp.Nt,
                         Gp.dof, Gp.ax, Gp.at, Gp.nv, c, full_mask(c), Gp.n(), G.n());
        if (G.nv > Gp.nv) {
            const int extra = G.nv - Gp.nv;
            std::vector<std::complex<double>> h((size_t)G.nv * G.n());
            rng.fill(h.data(), Gp.nv, extra, G.Nx, G.Nt, G.dof);
            DevMem tmp(sizeof(cplx<double>) * extra * G.n());
            DD_HIP(hipMemcpyAsync(tmp.as<void>(), h.data() + (size_t)Gp.nv * G.n(),
                                  sizeof(cplx<double>) * extra * G.n(), hipMemcpyHostToDevice, s_));
            convert<R, double>(s_, raw[l].as<cplx<R>>() + (size_t)Gp.nv * G.n(), tmp.as<cplx<double>>(),
                               G.n() * extra, 0, 0, 1, 1);
            DD_HIP(hipStreamSynchronize(s_));
        }
        DD_HIP(hipStreamSynchronize(s_));
    }
    DevMem L0(sizeof(cplx<R>) * (size_t)g_[0].V() * 5 * 4);
    dirac_links<R>(s_, U_, L0.as<cplx<R>>(), g_[0].Nx, g_[0].Nt);
    for (int l = 0; l < L - 1; l++) {
        const LevelGeom& G = g_[l];
        DevMem zero_rhs(sizeof(cplx<R>) * G.nv * G.n());
        zero<R>(s_, zero_rhs.as<cplx<R>>(), G.n() * G.nv, 0, 1, 1);
        // SAP(0, w, setup iterations) on all test vectors of the level at once
        sap(l, zero_rhs.as<cplx<R>>(), raw[l].as<cplx<R>>(), p_.setup_sap_iters, 1e-10, full_mask(G.nv));
        if (p_.bootstrap > 0) {   // the bootstrap passes start from these vectors
            tvraw_[l].alloc(sizeof(cplx<R>) * G.nv * G.n());
            DD_HIP(hipMemcpyAsync(tvraw_[l].as<void>(), raw[l].as<void>(), sizeof(cplx<R>) * G.nv * G.n(),
                                  hipMemcpyDeviceToDevice, s_));
        }
        transpose_tv<R>(s_, raw[l].as<cplx<R>>(), W_[l].as<cplx<R>>(), G.V(), G.nv, G.dof, true);
        orthonormalize<R>(s_, W_[l].as<cplx<R>>(), G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv);
        const cplx<R>* Lf = (l == 0) ? L0.as<cplx<R>>() : links(l);
        galerkin<R>(s_, Lf, W(l), links_[l + 1].as<cplx<R>>(), G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv, l > 0);
        if (hlinks_[l + 1].bytes())
            to_half(s_, reinterpret_cast<const float*>(links(l + 1)), hlinks_[l + 1].as<_Float16>(),
                    hlinks_[l + 1].bytes() / sizeof(_Float16));
        DD_HIP(hipStreamSynchronize(s_));
    }
    check_sap_failures();
    poly_build();
}

template <typename R>
void Multigrid<R>::clear_storage() {
    for (int l = 0; l < levels(); l++) {
        if (W_[l].bytes()) zero<R>(s_, W_[l].as<cplx<R>>(), W_[l].bytes() / sizeof(cplx<R>), 0, 1, 1);
        if (links_[l].bytes()) zero<R>(s_, links_[l].as<cplx<R>>(), links_[l].bytes() / sizeof(cplx<R>), 0, 1, 1);
        if (hlinks_[l].bytes()) DD_HIP(hipMemsetAsync(hlinks_[l].as<void>(), 0, hlinks_[l].bytes(), s_));
    }
    DD_HIP(hipStreamSynchronize(s_));
}

template <typename R>
void Multigrid<R>::set_test_vectors(int l, const cplx<R>* tv) {
    if (l < 0 || l >= levels() - 1) throw BadShape("test vectors exist on levels below the coarsest");
    const LevelGeom& G = g_[l];
    transpose_tv<R>(s_, tv, W_[l].as<cplx<R>>(), G.V(), G.nv, G.dof, true);
    DD_HIP(hipStreamSynchronize(s_));
}

template <typename R>
void Multigrid<R>::orthonormalize_level(int l) {
    if (l < 0 || l >= levels() - 1) throw BadShape("test vectors exist on levels below the coarsest");
    const LevelGeom& G = g_[l];
    orthonormalize<R>(s_, W_[l].as<cplx<R>>(), G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv);
    DD_HIP(hipStreamSynchronize(s_));
}

template <typename R>
void Multigrid<R>::dirac_links_level0(cplx<R>* out) {
    dirac_links<R>(s_, U_, out, g_[0].Nx, g_[0].Nt);
}

template <typename R>
void Multigrid<R>::galerkin_level(int l) {
    if (l < 0 || l >= levels() - 1) throw BadShape("no coarser level to build");
    const LevelGeom& G = g_[l];
    DevMem L0;
    if (l == 0) {
        L0.alloc(sizeof(cplx<R>) * (size_t)G.V() * 5 * 4);
        dirac_links_level0(L0.as<cplx<R>>());
    }
    const cplx<R>* Lf = (l == 0) ? L0.as<cplx<R>>() : links(l);
    galerkin<R>(s_, Lf, W(l), links_[l + 1].as<cplx<R>>(), G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv, l > 0);
    if (hlinks_[l + 1].bytes())
        to_half(s_, reinterpret_cast<const float*>(links(l + 1)), hlinks_[l + 1].as<_Float16>(),
                hlinks_[l + 1].bytes() / sizeof(_Float16));
    DD_HIP(hipStreamSynchronize(s_));
    if (l + 1 == levels() - 1) poly_build();
}

// Bootstrap adaptivity as the reference's documentation describes it (AMG_Documentation,
// section 3: "use the current multilevel method defined by P to approximately solve
// D x = (w_i - D w_i) and update w_i <- x; the resultant test vectors are orthonormalized
// again"; the reference declares the knob, AMGV::Nit include/variables.h:152, but its
// AlgebraicMG::setUpPhase src/amg.cpp:4-62 does not use it).  The approximate solve is
// one cycle of the configured kind from level l; test vectors go through it in batches
// of the solver's capacity.  w_i are the test vectors as they were before their
// orthonormalisation: the aggregate-wise orthonormalised ones are not smooth across
// aggregates, and passes started from them degrade the prolongator.
template <typename R>
void Multigrid<R>::bootstrap(int passes) {
    const int L = levels();
    if (passes > 0 && !tvraw_[0].bytes()) throw BadShape("bootstrap needs a setup with bootstrap > 0");
    const int cb = std::max(1, p_.max_rhs);
    DevMem L0(sizeof(cplx<R>) * (size_t)g_[0].V() * 5 * 4);
    dirac_links<R>(s_, U_, L0.as<cplx<R>>(), g_[0].Nx, g_[0].Nt);
    auto rebuild = [&](int l0) {   // coarse operators of levels l0+1 .. L-1 from the current W
        for (int l = l0; l < L - 1; l++) {
            const LevelGeom& G = g_[l];
            const cplx<R>* Lf = (l == 0) ? L0.as<cplx<R>>() : links(l);
            galerkin<R>(s_, Lf, W(l), links_[l + 1].as<cplx<R>>(), G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv, l > 0);
            if (hlinks_[l + 1].bytes())
                to_half(s_, reinterpret_cast<const float*>(links(l + 1)), hlinks_[l + 1].as<_Float16>(),
                        hlinks_[l + 1].bytes() / sizeof(_Float16));
        }
    };
    for (int it = 0; it < passes; it++)
        for (int l = 0; l < L - 1; l++) {
            const LevelGeom& G = g_[l];
            const size_t n = G.n();
            DevMem bb(sizeof(cplx<R>) * G.nv * n), xx(sizeof(cplx<R>) * G.nv * n);
            cplx<R>* w = tvraw_[l].as<cplx<R>>();
            cplx<R>* b = bb.as<cplx<R>>();
            cplx<R>* x = xx.as<cplx<R>>();
            for (int i0 = 0; i0 < G.nv; i0 += cb) {
                const int c = std::min(cb, G.nv - i0);
                const uint64_t mk = full_mask(c);
                apply(l, w + i0 * n, b + i0 * n, w + i0 * n, mk);   // b = w - D w
                if (p_.cycle == 1) kcycle(l, b + i0 * n, x + i0 * n, mk, true);   // zero initial guess
                else vcycle(l, b + i0 * n, x + i0 * n, mk, true);
            }
            DD_HIP(hipMemcpyAsync(w, x, sizeof(cplx<R>) * G.nv * n, hipMemcpyDeviceToDevice, s_));
            transpose_tv<R>(s_, x, W_[l].as<cplx<R>>(), G.V(), G.nv, G.dof, true);
            orthonormalize<R>(s_, W_[l].as<cplx<R>>(), G.Nx, G.Nt, G.dof, G.ax, G.at, G.nv);
            rebuild(l);
            DD_HIP(hipStreamSynchronize(s_));
            poly_build();   // the coarsest operator changed
        }
}

template <typename R>
void Multigrid<R>::autotune() {
    const int L = levels();
    sapc_var_.assign(L, -1);
    sapc_ms_.assign(L, 0.0);
    const int nr = std::max(1, p_.max_rhs);
    hipEvent_t e0, e1;
    DD_HIP(hipEventCreate(&e0));
    DD_HIP(hipEventCreate(&e1));
    for (int l = 1; l < L - 1; l++) {
        const LevelGeom& G = g_[l];
        const size_t n = G.n();
        cplx<R>* r = res_[l].as<cplx<R>>();
        cplx<R>* x = sapr_[l].as<cplx<R>>();
        k_fill_hash<R><<<(unsigned)((n * nr + 255) / 256), 256, 0, s_>>>(r, n * nr, 12345 + l);
        DD_HIP(hipGetLastError());
        // Each variant is timed once on the blocks of the first quarter of the level's x
        // extent (a prefix of the arrays: every access stays inside them), after one launch
        // on a single pair of blocks that loads its code.  The quarter keeps a quarter of the
        // work-groups of a full colour, so relative speeds are those of the full sweep.
        int nxq = G.Nx;
        for (int q = 4; q >= 2; q /= 2)
            if ((G.Nx / q) % G.bx == 0 && ((G.Nx / q / G.bx) * (G.Nt / G.bt)) % 2 == 0 && G.Nx / q >= 2 * G.bx) {
                nxq = G.Nx / q;
                break;
            }
        const int nbt = G.Nt / G.bt;
        const bool pair_ok = (nbt % 2 == 0);   // one block row of the first x block: two colours
        double best = 1e300;
        int bestv = -1;
        for (int v : sapc_variants<R>()) {
            if (!sapc_supported<R>(v, G.dof, G.bx * G.bt)) continue;
            if (v != 32 && v != 33 && v / 2 > 1 && nr == 1) continue;   // multi-system vector variants only for batches
            if (v == 33 && !hlinks(l)) continue;                          // no binary16 link copy
            zero<R>(s_, x, n * nr, 0, 1, 1);
            if (pair_ok)   // code loading: a lattice of one x block row and two t blocks
                sap_blocksc<R>(s_, links(l), r, x, G.bx, 2 * G.bt, G.dof, G.bx, G.bt, 0, p_.m0, p_.sap_m,
                               p_.sap_restarts, p_.sap_tol, nr, full_mask(nr), n, nullptr, v, hlinks(l));
            DD_HIP(hipEventRecord(e0, s_));
            sap_blocksc<R>(s_, links(l), r, x, nxq, G.Nt, G.dof, G.bx, G.bt, 0, p_.m0, p_.sap_m, p_.sap_restarts,
                           p_.sap_tol, nr, full_mask(nr), n, nullptr, v, hlinks(l));
            DD_HIP(hipEventRecord(e1, s_));
            DD_HIP(hipEventSynchronize(e1));
            float ms = 0;
            DD_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double t = ms * (double)G.Nx / nxq;   // scaled to the full half-sweep
            if (t < best) {
                best = t;
                bestv = v;
            }
            if (std::getenv("DDAMG_VERBOSE"))
                fprintf(stderr, "autotune level %d: coarse SAP variant %d (%d systems/work-group%s): %.3f ms\n", l, v,
                        v >= 32 ? 16 : v / 2,
                        v == 32 ? ", matrix cores" : v == 33 ? ", matrix cores, binary16 links"
                                                   : ((v & 1) ? ", batched link loads" : ""),
                        t);
        }
        sapc_var_[l] = bestv;
        sapc_ms_[l] = best;
    }
    // coarsest engine (ortho_coarse = 3): the same 96 Arnoldi steps (one cycle, no
    // convergence test) on max_rhs pseudo-random systems with each engine; keep the faster
    if (coarse_dev_ && coarse_) {
        const LevelGeom& G = g_[L - 1];
        const size_t n = G.n();
        cplx<R>* e = eta_[L - 1].template as<cplx<R>>();
        cplx<R>* x = psi_[L - 1].template as<cplx<R>>();
        k_fill_hash<R><<<(unsigned)((n * nr + 255) / 256), 256, 0, s_>>>(e, n * nr, 777);
        DD_HIP(hipGetLastError());
        double t[2] = {0, 0};
        const int msteps = std::min(96, p_.coarse_m);
        for (int eng = 0; eng < 2; eng++) {
            use_dev_ = eng == 1;
            // a short untimed run first (loads the kernels' code), then the timed one
            coarse_->set_limits(2, 1, 0.0);
            coarse_dev_->set_limits(2, 1, 0.0);
            coarsest(e, x, full_mask(nr));
            coarse_->set_limits(msteps, 1, 0.0);
            coarse_dev_->set_limits(msteps, 1, 0.0);
            DD_HIP(hipEventRecord(e0, s_));
            coarsest(e, x, full_mask(nr));
            DD_HIP(hipEventRecord(e1, s_));
            DD_HIP(hipEventSynchronize(e1));
            float ms = 0;
            DD_HIP(hipEventElapsedTime(&ms, e0, e1));
            t[eng] = ms;
        }
        coarse_->set_limits(p_.coarse_m, p_.coarse_restarts, p_.coarse_tol);
        coarse_dev_->set_limits(p_.coarse_m, p_.coarse_restarts, p_.coarse_tol);
        coarse_->steps = 0;
        coarse_dev_->steps = 0;
        use_dev_ = t[1] <= t[0];
        coarse_ms_[0] = t[0];
        coarse_ms_[1] = t[1];
        coarse_iters.clear();
        if (std::getenv("DDAMG_VERBOSE"))
            fprintf(stderr, "autotune coarsest: launch-per-step %.3f ms, persistent %.3f ms -> %s\n", t[0], t[1],
                    use_dev_ ? "persistent" : "launch-per-step");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

template <typename R>
void Multigrid<R>::set_sap_variant(int v) {
    const int L = levels();
    sapc_var_.assign(L, -1);
    sapc_ms_.assign(L, 0.0);
    for (int l = 1; l < L - 1; l++) {
        if (!sapc_supported<R>(v, g_[l].dof, g_[l].bx * g_[l].bt) || (v == 33 && !hlinks(l)))
            throw BadShape("sap_variant not available for this level's precision, dof, block or link copy");
        sapc_var_[l] = v;
    }
}

template class Krylov<double>;
template class Krylov<float>;
template class Multigrid<double>;
template class Multigrid<float>;

// ---------------------------------------------------------------------------------------
// Top-level solver
// ---------------------------------------------------------------------------------------
Solver::Solver(int nx, int nt, int dev) : Nx(nx), Nt(nt), device(dev) {
    if (Nx < 2 || Nt < 2) throw BadShape("lattice too small");
    DD_HIP(hipSetDevice(device));
    DD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    U64.alloc(sizeof(cplx<double>) * (size_t)Nx * Nt * 2);
    U32.alloc(sizeof(cplx<float>) * (size_t)Nx * Nt * 2);
    theta.alloc(sizeof(double) * (size_t)Nx * Nt * 2);
}

Solver::~Solver() {
    mg64.reset();
    mg32.reset();
    outer_.reset();
    if (s) (void)hipStreamDestroy(s);
}

void Solver::set_gauge(const cplx<double>* U_dev) {
    const size_t n = (size_t)Nx * Nt * 2;
    DD_HIP(hipMemcpyAsync(U64.as<void>(), U_dev, sizeof(cplx<double>) * n, hipMemcpyDeviceToDevice, s));
    convert<float, double>(s, U32.as<cplx<float>>(), U64.as<cplx<double>>(), n, 0, 0, 1, 1);
    DD_HIP(hipStreamSynchronize(s));
    have_gauge = true;
    have_setup = false;
}

void Solver::apply_D(const cplx<double>* in, cplx<double>* out, int nrhs, double m0, bool dagger, int precision) {
    const size_t n = (size_t)Nx * Nt * 2;
    if (precision == 0) {
        wilson<double>(s, U64.as<cplx<double>>(), in, out, nullptr, Nx, Nt, m0, dagger, nrhs, full_mask(nrhs), n);
    } else {
        DevMem a(sizeof(cplx<float>) * n * nrhs), b(sizeof(cplx<float>) * n * nrhs);
        convert<float, double>(s, a.as<cplx<float>>(), in, n, n, n, nrhs, full_mask(nrhs));
        wilson<float>(s, U32.as<cplx<float>>(), a.as<cplx<float>>(), b.as<cplx<float>>(), nullptr, Nx, Nt, m0, dagger,
                      nrhs, full_mask(nrhs), n);
        convert<double, float>(s, out, b.as<cplx<float>>(), n, n, n, nrhs, full_mask(nrhs));
        DD_HIP(hipStreamSynchronize(s));
    }
}

void Solver::setup(const dd_params& p) {
    if (!have_gauge) throw BadShape("setup before the gauge field");
    auto t0 = std::chrono::steady_clock::now();
    prm = p;
    mg64.reset();
    mg32.reset();
    outer_.reset();
    std::vector<LevelGeom> g = make_geometry(Nx, Nt, p);
    RankStreams rng(p.rng_ranks_x, p.rng_ranks_t);
    std::vector<std::complex<double>> h((size_t)g[0].nv * g[0].n());
    rng.fill(h.data(), 0, g[0].nv, Nx, Nt, 2);
    DevMem tv(sizeof(cplx<double>) * h.size());
    DD_HIP(hipMemcpyAsync(tv.as<void>(), h.data(), sizeof(cplx<double>) * h.size(), hipMemcpyHostToDevice, s));
    if (p.precision == 0) {
        mg64.reset(new Multigrid<double>(s, Nx, Nt, p, U64.as<cplx<double>>()));
        mg64->setup(tv.as<cplx<double>>(), rng);
        if (p.bootstrap > 0) mg64->bootstrap(p.bootstrap);
        if (p.sap_variant >= 0) mg64->set_sap_variant(p.sap_variant);
        else if (p.autotune) mg64->autotune();
    } else {
        DevMem tv32(sizeof(cplx<float>) * h.size());
        convert<float, double>(s, tv32.as<cplx<float>>(), tv.as<cplx<double>>(), h.size(), 0, 0, 1, 1);
        mg32.reset(new Multigrid<float>(s, Nx, Nt, p, U32.as<cplx<float>>()));
        mg32->setup(tv32.as<cplx<float>>(), rng);
        if (p.bootstrap > 0) mg32->bootstrap(p.bootstrap);
        if (p.sap_variant >= 0) mg32->set_sap_variant(p.sap_variant);
        else if (p.autotune) mg32->autotune();
    }
    DD_HIP(hipStreamSynchronize(s));
    setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    have_setup = true;
}

void Solver::setup_storage(const dd_params& p) {
    if (!have_gauge) throw BadShape("setup before the gauge field");
    prm = p;
    mg64.reset();
    mg32.reset();
    outer_.reset();
    if (p.precision == 0) {
        mg64.reset(new Multigrid<double>(s, Nx, Nt, p, U64.as<cplx<double>>()));
        mg64->clear_storage();
    } else {
        mg32.reset(new Multigrid<float>(s, Nx, Nt, p, U32.as<cplx<float>>()));
        mg32->clear_storage();
    }
    setup_seconds = 0;
    have_setup = true;
}

int Solver::solve(int nrhs, const cplx<double>* b, cplx<double>* x, int* iters, dd_solve_info* info, bool use_x0) {
    if (!have_setup) throw BadShape("solve before setup");
    if (nrhs < 1 || nrhs > std::max(1, prm.max_rhs)) throw BadShape("nrhs exceeds max_rhs of the setup");
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = (size_t)Nx * Nt * 2;
    const int cap = std::max(1, prm.max_rhs);
    if (!outer_ || outer_cap_ != cap) {
        outer_.reset(new Krylov<double>(s, n, prm.fgmres_m, prm.fgmres_restarts, prm.tol, cap, true, prm.ortho_outer));
        outer_cap_ = cap;
        if (prm.precision == 1) {
            x32_.alloc(sizeof(cplx<float>) * n * cap);
            y32_.alloc(sizeof(cplx<float>) * n * cap);
        } else if (prm.graph) {   // fixed buffers for the captured cycle
            x64_.alloc(sizeof(cplx<double>) * n * cap);
            y64_.alloc(sizeof(cplx<double>) * n * cap);
        }
    }
    bool graphs_used = false;
    const uint64_t mask = full_mask(nrhs);
    Krylov<double>::Func A = [this, n](const cplx<double>* in, cplx<double>* out, uint64_t m) {
        wilson<double>(s, U64.as<cplx<double>>(), in, out, nullptr, Nx, Nt, prm.m0, false, mask_count(m), m, n);
    };
    // graph = 1: the cycle on fixed buffers is captured once per mask and replayed; the
    // copies in and out stay outside the graph
    Krylov<double>::Prec M = [this, n, &graphs_used](const cplx<double>* in, cplx<double>* io, uint64_t m) {
        const int nr = mask_count(m);
        const bool graph = prm.graph != 0;
        const bool kc = prm.cycle == 1;
        if (prm.precision == 0) {
            if (graph) {
                cplx<double>* a = x64_.as<cplx<double>>();
                cplx<double>* y = y64_.as<cplx<double>>();
                copy<double>(s, a, in, n, n, n, nr, m);
                if (kc ? mg64->kcycle_graph(a, y, m) : mg64->vcycle_graph(a, y, m)) {
                    graphs_used = true;
                    copy<double>(s, io, y, n, n, n, nr, m);
                    return;
                }
            }
            if (prm.cycle == 1) mg64->kcycle(0, in, io, m, true);   // zero initial guess
            else mg64->vcycle(0, in, io, m, true);
        } else {
            cplx<float>* a = x32_.as<cplx<float>>();
            cplx<float>* y = y32_.as<cplx<float>>();
            convert<float, double>(s, a, in, n, n, n, nr, m);
            if (graph && (kc ? mg32->kcycle_graph(a, y, m) : mg32->vcycle_graph(a, y, m))) {
                graphs_used = true;
            } else {
                if (prm.cycle == 1) mg32->kcycle(0, a, y, m, true);   // zero initial guess
                else mg32->vcycle(0, a, y, m, true);
            }
            convert<double, float>(s, io, y, n, n, n, nr, m);
        }
    };
    if (!use_x0) zero<double>(s, x, n, n, nrhs, mask);
    std::vector<int> its(64, 0);
    std::vector<double> rr(64, 0.0);
    const uint64_t conv = outer_->solve(b, x, x, mask, A, &M, its.data(), rr.data());
    if (graphs_used) {
        if (mg64) mg64->check_graph_failures();
        if (mg32) mg32->check_graph_failures();
    }
    if (mg64) mg64->check_sap_failures();
    if (mg32) mg32->check_sap_failures();
    DD_HIP(hipStreamSynchronize(s));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int mx = 0;
    double mr = 0;
    for (int q = 0; q < nrhs; q++) {
        if (iters) iters[q] = its[q];
        mx = std::max(mx, its[q]);
        mr = std::max(mr, rr[q]);
    }
    if (info) {
        info->iterations = mx;
        info->relres = mr;
        info->seconds = secs;
    }
    return conv == mask ? DD_OK : DD_ERR_NOT_CONVERGED;
}

}  // namespace dd
