// Comparison solvers of the reference on the level-0 FP64 Wilson-Dirac operator, batched
// over independent systems in lock-step (see Solver::solve_cg & co. in solver.h).
//
// Each system follows the reference's recurrence and stopping test exactly; the vector
// operations run as device kernels (Wilson-Dirac stencil, two-pass dots, per-system
// complex linear combinations) and the handful of scalars per iteration are formed on
// the host in std::complex<double>, as the reference forms them.  A system leaves the
// batch when it converges; the others continue.
#include <chrono>
#include <complex>
#include <vector>

#include "solver.h"

namespace dd {

namespace {

using C = std::complex<double>;

// Per-solve scratch: nvec vectors of nrhs systems, reduction partials, the coefficient
// array of lincomb and pinned mirrors of the per-system dots.
struct Scratch {
    Scratch(size_t n, int nrhs, int nvec) : n(n), nrhs(nrhs) {
        vec.alloc(sizeof(cplx<double>) * n * nrhs * nvec);
        part.alloc(sizeof(cplx<double>) * (size_t)nrhs * 256);
        rb.partial = part.as<void>();
        dots.alloc(sizeof(cplx<double>) * nrhs * 4);
        doth.alloc(sizeof(cplx<double>) * nrhs * 4);
        coef.alloc(sizeof(cplx<double>) * nrhs * 3);
        coefh.alloc(sizeof(cplx<double>) * nrhs * 3);
    }
    cplx<double>* v(int k) const { return vec.as<cplx<double>>() + (size_t)k * n * nrhs; }
    size_t n;
    int nrhs;
    DevMem vec, part, dots, coef;
    PinMem doth, coefh;
    RedBuf rb;
};

// d[slot][r] = sum_i a_r[i] conj(b_r[i]) for the systems in mask (the reference's dot)
void dots(hipStream_t s, Scratch& w, const cplx<double>* a, const cplx<double>* b, int slot, uint64_t mask) {
    dot_partial<double>(s, a, b, w.n, w.n, w.nrhs, mask, w.rb);
    dot_final<double>(s, w.rb, w.dots.as<cplx<double>>() + slot, 4, w.nrhs, mask);
}

// copies the dot slots to the host (one synchronisation per call)
std::vector<std::vector<C>> fetch(hipStream_t s, Scratch& w) {
    DD_HIP(hipMemcpyAsync(w.doth.as<void>(), w.dots.as<void>(), sizeof(cplx<double>) * 4 * w.nrhs,
                          hipMemcpyDeviceToHost, s));
    DD_HIP(hipStreamSynchronize(s));
    std::vector<std::vector<C>> d(w.nrhs, std::vector<C>(4));
    const cplx<double>* h = w.doth.as<cplx<double>>();
    for (int r = 0; r < w.nrhs; r++)
        for (int k = 0; k < 4; k++) d[r][k] = C(h[4 * r + k].x, h[4 * r + k].y);
    return d;
}

// out = c0 x + c1 y + c2 z with per-system coefficients from the host
void combine(hipStream_t s, Scratch& w, cplx<double>* out, const cplx<double>* x, const cplx<double>* y,
             const cplx<double>* z, const std::vector<C>& c0, const std::vector<C>& c1, const std::vector<C>& c2,
             uint64_t mask) {
    cplx<double>* h = w.coefh.as<cplx<double>>();
    // the previous combine's coefficients may still be in flight to the device
    DD_HIP(hipStreamSynchronize(s));
    for (int r = 0; r < w.nrhs; r++) {
        h[3 * r] = mk<double>(c0[r].real(), c0[r].imag());
        h[3 * r + 1] = mk<double>(c1[r].real(), c1[r].imag());
        h[3 * r + 2] = mk<double>(c2[r].real(), c2[r].imag());
    }
    DD_HIP(hipMemcpyAsync(w.coef.as<void>(), h, sizeof(cplx<double>) * 3 * w.nrhs, hipMemcpyHostToDevice, s));
    lincomb<double>(s, out, x, y, z, w.coef.as<cplx<double>>(), w.n, w.n, w.nrhs, mask);
}

void finish(int nrhs, const std::vector<int>& its, const std::vector<double>& rel, int* iters, dd_solve_info* info,
            std::chrono::steady_clock::time_point t0) {
    int mx = 0;
    double mr = 0;
    for (int q = 0; q < nrhs; q++) {
        if (iters) iters[q] = its[q];
        mx = std::max(mx, its[q]);
        mr = std::max(mr, rel[q]);
    }
    if (info) {
        info->iterations = mx;
        info->relres = mr;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

}  // namespace

// conjugate_gradient (src/conjugate_gradient.cpp:4-84) on D^dagger D x = phi (Methods::CG,
// src/methods.cpp:43-59, passes phi = D^dagger b).  sol = phi initially; converged when
// ||r|| < tol ||phi||, returning k+1; the reference returns 0 on max_iter, here the count
// is max_iter and the status DD_ERR_NOT_CONVERGED.  relres is ||phi - D^dagger D x|| / ||phi||.
int Solver::solve_cg(int nrhs, const cplx<double>* phi, cplx<double>* x, double m0, int max_iter, double tol,
                     int* iters, dd_solve_info* info) {
    if (!have_gauge) throw BadShape("CG before the gauge field");
    if (nrhs < 1 || nrhs > 64) throw BadShape("CG: 1..64 systems");
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = (size_t)Nx * Nt * 2;
    const uint64_t all = full_mask(nrhs);
    Scratch w(n, nrhs, 4);
    cplx<double> *r = w.v(0), *d = w.v(1), *Ad = w.v(2), *tmp = w.v(3);
    const cplx<double>* U = U64.as<cplx<double>>();
    auto DdD = [&](const cplx<double>* in, cplx<double>* out, uint64_t m) {
        wilson<double>(s, U, in, tmp, nullptr, Nx, Nt, m0, false, nrhs, m, n);
        wilson<double>(s, U, tmp, out, nullptr, Nx, Nt, m0, true, nrhs, m, n);
    };
    if (x != phi) copy<double>(s, x, phi, n, n, n, nrhs, all);                // sol = phi
    DdD(x, Ad, all);
    const std::vector<C> one(nrhs, C(1, 0)), mone(nrhs, C(-1, 0)), zero(nrhs, C(0, 0));
    combine(s, w, r, phi, Ad, nullptr, one, mone, zero, all);                 // r = phi - D^+D x
    copy<double>(s, d, r, n, n, n, nrhs, all);                                // d = r
    dots(s, w, r, r, 0, all);
    dots(s, w, phi, phi, 1, all);
    auto h = fetch(s, w);
    std::vector<C> rn2(nrhs);
    std::vector<double> nphi(nrhs), err(nrhs, 0.0);
    std::vector<int> its(nrhs, max_iter);
    for (int q = 0; q < nrhs; q++) {
        rn2[q] = h[q][0];
        nphi[q] = std::sqrt(h[q][1].real());
        err[q] = std::sqrt(h[q][0].real());
    }
    uint64_t run = all;
    for (int k = 0; k < max_iter && run; k++) {
        DdD(d, Ad, run);
        dots(s, w, d, Ad, 0, run);
        h = fetch(s, w);
        std::vector<C> alpha(nrhs, C(0, 0)), malpha(nrhs, C(0, 0));
        for (int q = 0; q < nrhs; q++)
            if ((run >> q) & 1) {
                alpha[q] = rn2[q] / h[q][0];   // alpha = (r,r)/(d,Ad)
                malpha[q] = -alpha[q];
            }
        combine(s, w, x, x, d, nullptr, one, alpha, zero, run);     // x += alpha d
        combine(s, w, r, r, Ad, nullptr, one, malpha, zero, run);   // r -= alpha Ad
        dots(s, w, r, r, 0, run);
        h = fetch(s, w);
        std::vector<C> beta(nrhs, C(0, 0));
        uint64_t next = run;
        for (int q = 0; q < nrhs; q++) {
            if (!((run >> q) & 1)) continue;
            const double e2 = h[q][0].real();
            err[q] = std::sqrt(e2);
            if (err[q] < tol * nphi[q]) {
                its[q] = k + 1;
                next &= ~(1ull << q);
                continue;
            }
            beta[q] = C(e2, 0) / rn2[q];   // beta = (r_{i+1},r_{i+1})/(r_i,r_i)
            rn2[q] = C(e2, 0);
        }
        run = next;
        if (run) combine(s, w, d, r, d, nullptr, one, beta, zero, run);   // d = r + beta d
    }
    std::vector<double> rel(nrhs);
    for (int q = 0; q < nrhs; q++) rel[q] = nphi[q] > 0 ? err[q] / nphi[q] : 0.0;
    finish(nrhs, its, rel, iters, info, t0);
    return run ? DD_ERR_NOT_CONVERGED : DD_OK;
}

// bi_cgstab (src/conjugate_gradient.cpp:87-176) on D x = b from x0, r_tilde = r_0:
//   rho = (r, r~); d = r (k = 0) or r + beta (d - omega Ad) with
//   beta = alpha rho / (omega rho_prev); alpha = rho / (Ad, r~); s = r - alpha Ad;
//   stop when ||s|| < tol ||b|| with x += alpha d (count k+1); omega = (s, t)/(t, t)
//   with t = D s; r = s - omega t; x += alpha d + omega s.
// The count is max_iter when a system does not converge (as the reference returns).
int Solver::solve_bicgstab(int nrhs, const cplx<double>* b, const cplx<double>* x0, cplx<double>* x, double m0,
                           int max_iter, double tol, int* iters, dd_solve_info* info) {
    if (!have_gauge) throw BadShape("BiCGStab before the gauge field");
    if (nrhs < 1 || nrhs > 64) throw BadShape("BiCGStab: 1..64 systems");
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = (size_t)Nx * Nt * 2;
    const uint64_t all = full_mask(nrhs);
    Scratch w(n, nrhs, 6);
    cplx<double> *r = w.v(0), *rt = w.v(1), *d = w.v(2), *sv = w.v(3), *t = w.v(4), *Ad = w.v(5);
    const cplx<double>* U = U64.as<cplx<double>>();
    if (x != x0) copy<double>(s, x, x0, n, n, n, nrhs, all);
    wilson<double>(s, U, x, r, b, Nx, Nt, m0, false, nrhs, all, n);   // r = b - D x
    copy<double>(s, rt, r, n, n, n, nrhs, all);                      // r~ = r
    dots(s, w, b, b, 1, all);
    auto h = fetch(s, w);
    std::vector<double> nphi(nrhs), err(nrhs, 0.0);
    for (int q = 0; q < nrhs; q++) nphi[q] = std::sqrt(h[q][1].real());
    const std::vector<C> one(nrhs, C(1, 0)), zero(nrhs, C(0, 0));
    std::vector<C> alpha(nrhs, C(0, 0)), omega(nrhs, C(0, 0)), rho(nrhs, C(0, 0)), rho_prev(nrhs, C(0, 0));
    std::vector<int> its(nrhs, max_iter);
    uint64_t run = all;
    for (int k = 0; k < max_iter && run; k++) {
        dots(s, w, r, rt, 0, run);
        h = fetch(s, w);
        for (int q = 0; q < nrhs; q++)
            if ((run >> q) & 1) rho[q] = h[q][0];
        if (k == 0) {
            copy<double>(s, d, r, n, n, n, nrhs, run);
        } else {
            std::vector<C> beta(nrhs, C(0, 0)), mbo(nrhs, C(0, 0));
            for (int q = 0; q < nrhs; q++)
                if ((run >> q) & 1) {
                    beta[q] = alpha[q] * rho[q] / (omega[q] * rho_prev[q]);
                    mbo[q] = -beta[q] * omega[q];
                }
            combine(s, w, d, r, d, Ad, one, beta, mbo, run);   // d = r + beta (d - omega Ad)
        }
        wilson<double>(s, U, d, Ad, nullptr, Nx, Nt, m0, false, nrhs, run, n);
        dots(s, w, Ad, rt, 0, run);
        h = fetch(s, w);
        std::vector<C> malpha(nrhs, C(0, 0));
        for (int q = 0; q < nrhs; q++)
            if ((run >> q) & 1) {
                alpha[q] = rho[q] / h[q][0];
                malpha[q] = -alpha[q];
            }
        combine(s, w, sv, r, Ad, nullptr, one, malpha, zero, run);   // s = r - alpha Ad
        dots(s, w, sv, sv, 0, run);
        h = fetch(s, w);
        uint64_t conv = 0;
        for (int q = 0; q < nrhs; q++) {
            if (!((run >> q) & 1)) continue;
            err[q] = std::sqrt(h[q][0].real());
            if (err[q] < tol * nphi[q]) {
                conv |= 1ull << q;
                its[q] = k + 1;
            }
        }
        if (conv) combine(s, w, x, x, d, nullptr, one, alpha, zero, conv);   // x += alpha d
        run &= ~conv;
        if (!run) break;
        wilson<double>(s, U, sv, t, nullptr, Nx, Nt, m0, false, nrhs, run, n);
        dots(s, w, sv, t, 0, run);
        dots(s, w, t, t, 2, run);
        h = fetch(s, w);
        std::vector<C> momega(nrhs, C(0, 0));
        for (int q = 0; q < nrhs; q++)
            if ((run >> q) & 1) {
                omega[q] = h[q][0] / h[q][2];
                momega[q] = -omega[q];
            }
        combine(s, w, r, sv, t, nullptr, one, momega, zero, run);   // r = s - omega t
        combine(s, w, x, x, d, sv, one, alpha, omega, run);         // x += alpha d + omega s
        for (int q = 0; q < nrhs; q++)
            if ((run >> q) & 1) rho_prev[q] = rho[q];
    }
    wilson<double>(s, U, x, t, b, Nx, Nt, m0, false, nrhs, all, n);
    dots(s, w, t, t, 0, all);
    h = fetch(s, w);
    std::vector<double> rel(nrhs);
    for (int q = 0; q < nrhs; q++) rel[q] = nphi[q] > 0 ? std::sqrt(h[q][0].real()) / nphi[q] : 0.0;
    finish(nrhs, its, rel, iters, info, t0);
    return run ? DD_ERR_NOT_CONVERGED : DD_OK;
}

// FGMRES_fine_level (identity preconditioner: GMRES) or FGMRES_SAP, whose preconditioner
// is SAP_fine_level::SAP(in, out = 0, nu = 1, tolerance 1e-10) (include/sap.h:274-284):
// one red-black sweep over sap_bx x sap_bt site blocks, each block solved by GMRES(5) x 5
// to 1e-3 from the block right-hand side (the SAPV defaults, src/variables.cpp:100-105).
// Iteration count k*m + maxIt as FGMRES::fgmres.
int Solver::solve_gmres(int nrhs, const cplx<double>* b, const cplx<double>* x0, cplx<double>* x, double m0, int m,
                        int restarts, double tol, int sap_bx, int sap_bt, int* iters, dd_solve_info* info) {
    if (!have_gauge) throw BadShape("GMRES before the gauge field");
    if (nrhs < 1 || nrhs > 64) throw BadShape("GMRES: 1..64 systems");
    const bool sap_prec = sap_bx > 0;
    if (sap_prec) {
        if (sap_bt < 1 || Nx % sap_bx || Nt % sap_bt) throw BadShape("FGMRES+SAP: block size must divide the lattice");
        const int nbx = Nx / sap_bx, nbt = Nt / sap_bt, S = sap_bx * sap_bt;
        if ((nbx * nbt) % 2 || ((nbt & 1) && (nbx & 1))) throw BadShape("FGMRES+SAP: SAP needs an even number of blocks");
        if (S > 64 || (S & (S - 1))) throw BadShape("FGMRES+SAP: blocks of a power-of-two number of sites, at most 64");
    }
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = (size_t)Nx * Nt * 2;
    const uint64_t all = full_mask(nrhs);
    const cplx<double>* U = U64.as<cplx<double>>();
    Krylov<double> kry(s, n, m, restarts, tol, nrhs, sap_prec, 0);
    Krylov<double>::Func A = [this, n, m0, U](const cplx<double>* in, cplx<double>* out, uint64_t mk) {
        wilson<double>(s, U, in, out, nullptr, Nx, Nt, m0, false, mask_count(mk), mk, n);
    };
    DevMem rbuf;
    if (sap_prec) rbuf.alloc(sizeof(cplx<double>) * n * nrhs);
    // out = SAP(v, 0, nu = 1): the residual before each colour is v - D out; with one
    // iteration the tolerance test of SAP_C::SAP (src/sap.cpp:150-156) stops nothing
    Krylov<double>::Prec M = [&](const cplx<double>* v, cplx<double>* out, uint64_t mk) {
        const int nr = mask_count(mk);
        cplx<double>* r = rbuf.as<cplx<double>>();
        zero<double>(s, out, n, n, nr, mk);
        copy<double>(s, r, v, n, n, n, nr, mk);
        for (int colour = 0; colour < 2; colour++) {
            sap_blocks0<double>(s, U, r, out, Nx, Nt, sap_bx, sap_bt, colour, m0, 5, 5, 1e-3, nr, mk, n, nullptr);
            wilson<double>(s, U, out, r, v, Nx, Nt, m0, false, nr, mk, n);
        }
    };
    std::vector<int> its(64, 0);
    std::vector<double> rr(64, 0.0);
    const uint64_t conv = kry.solve(b, x0, x, all, A, sap_prec ? &M : nullptr, its.data(), rr.data());
    DD_HIP(hipStreamSynchronize(s));
    finish(nrhs, its, rr, iters, info, t0);
    return conv == all ? DD_OK : DD_ERR_NOT_CONVERGED;
}

// SAP_fine_level::SAP(v, x, nu, tol) as a stand-alone solver (src/sap.cpp:81-165;
// Methods::SAP, src/methods.cpp:61-78): r = v - D x, then per iteration the red blocks,
// r = v - D x, the black blocks, r = v - D x, and the test ||r|| < tol ||v||, which ends
// the system's iterations with the count i (0-based), as the reference returns it; nu
// when it never holds.  Blocks of sap_bx x sap_bt sites, solved by GMRES(5) x 5 to 1e-3
// from the block right-hand side (the SAPV defaults, src/variables.cpp:100-105).
int Solver::sap_fine(int nrhs, const cplx<double>* v, cplx<double>* x, int nu, double tol, int sap_bx, int sap_bt,
                     double m0, int* iters, double* relres) {
    if (!have_gauge) throw BadShape("SAP before the gauge field");
    if (nrhs < 1 || nrhs > 64) throw BadShape("SAP: 1..64 systems");
    if (sap_bx < 1 || sap_bt < 1 || Nx % sap_bx || Nt % sap_bt) throw BadShape("SAP: block size must divide the lattice");
    const int nbx = Nx / sap_bx, nbt = Nt / sap_bt, S = sap_bx * sap_bt;
    if ((nbx * nbt) % 2 || ((nbt & 1) && (nbx & 1))) throw BadShape("SAP needs an even number of blocks");
    if (S > 64 || (S & (S - 1))) throw BadShape("SAP: blocks of a power-of-two number of sites, at most 64");
    const size_t n = (size_t)Nx * Nt * 2;
    const cplx<double>* U = U64.as<cplx<double>>();
    DevMem rbuf(sizeof(cplx<double>) * n * nrhs), nrm(sizeof(cplx<double>) * 2 * nrhs),
        part(dot_scratch_bytes(nrhs));
    PinMem nh;
    nh.alloc(sizeof(cplx<double>) * 2 * nrhs);
    RedBuf rb;
    rb.partial = part.as<void>();
    cplx<double>* r = rbuf.as<cplx<double>>();
    cplx<double>* nd = nrm.as<cplx<double>>();
    cplx<double>* hn = nh.as<cplx<double>>();
    uint64_t run = full_mask(nrhs);
    dot_partial<double>(s, v, v, n, n, nrhs, run, rb);
    dot_final<double>(s, rb, nd, 2, nrhs, run);
    wilson<double>(s, U, x, r, v, Nx, Nt, m0, false, nrhs, run, n);
    DD_HIP(hipMemcpyAsync(hn, nd, sizeof(cplx<double>) * 2 * nrhs, hipMemcpyDeviceToHost, s));
    DD_HIP(hipStreamSynchronize(s));
    std::vector<double> vnorm(nrhs), err(nrhs, 0.0);
    for (int q = 0; q < nrhs; q++) {
        vnorm[q] = std::sqrt(hn[2 * q].x);
        iters[q] = nu;
    }
    for (int i = 0; i < nu && run; i++) {
        for (int colour = 0; colour < 2; colour++) {
            sap_blocks0<double>(s, U, r, x, Nx, Nt, sap_bx, sap_bt, colour, m0, 5, 5, 1e-3, nrhs, run, n, nullptr);
            wilson<double>(s, U, x, r, v, Nx, Nt, m0, false, nrhs, run, n);
        }
        dot_partial<double>(s, r, r, n, n, nrhs, run, rb);
        dot_final<double>(s, rb, nd + 1, 2, nrhs, run);
        DD_HIP(hipMemcpyAsync(hn, nd, sizeof(cplx<double>) * 2 * nrhs, hipMemcpyDeviceToHost, s));
        DD_HIP(hipStreamSynchronize(s));
        for (int q = 0; q < nrhs; q++) {
            if (!((run >> q) & 1)) continue;
            err[q] = std::sqrt(hn[2 * q + 1].x);
            if (err[q] < tol * vnorm[q]) {
                iters[q] = i;
                run &= ~(1ull << q);
            }
        }
    }
    if (relres)
        for (int q = 0; q < nrhs; q++) relres[q] = vnorm[q] > 0 ? err[q] / vnorm[q] : 0.0;
    return run ? DD_ERR_NOT_CONVERGED : DD_OK;
}

}  // namespace dd
