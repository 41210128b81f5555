// Roots of the GMRES polynomial preconditioner of the coarsest solve (host code).
//
// d Arnoldi steps on the operator A from a start vector give the (d+1) x d Hessenberg
// matrix H.  The residual polynomial of d GMRES steps from that vector is
// pi(z) = prod_i (1 - z / theta_i) over the harmonic Ritz values theta_i, the eigenvalues
// of H_d + |h_{d+1,d}|^2 f e_d^T with H_d^H f = e_d (Loe and Morgan, "Toward efficient
// polynomial preconditioning for GMRES", Numer. Linear Algebra Appl. 29 (2022) e2427).
// The preconditioner is p(z) = (1 - pi(z)) / z, so that A p(A) = 1 - pi(A) has its
// spectrum pulled towards 1.  It is applied in the product form
//     y = 0, r = v;  for i < d: y += r / theta_i, r -= A r / theta_i  (the last A skipped)
// with the roots in Leja order (each next root the one farthest, in the product of
// distances, from those before it), which keeps the partial products bounded.
#pragma once
#include <cmath>
#include <complex>
#include <stdexcept>
#include <vector>

namespace dd {

using Cd = std::complex<double>;

// Eigenvalues of the n x n upper Hessenberg matrix a (row-major; entries below the
// subdiagonal are ignored) by the shifted QR algorithm: Wilkinson shifts, Givens
// rotations on the active block, deflation where a subdiagonal entry falls below
// 1e-15 of its diagonal neighbours.  Throws if an eigenvalue has not converged after
// 60 iterations.
inline std::vector<Cd> hessenberg_eigenvalues(std::vector<Cd> a, int n) {
    std::vector<Cd> ev(n);
    auto A = [&](int i, int j) -> Cd& { return a[(size_t)i * n + j]; };
    int hi = n - 1, iter = 0;
    std::vector<Cd> cs(n), sn(n);
    while (hi >= 0) {
        if (hi == 0) {
            ev[0] = A(0, 0);
            break;
        }
        int lo = hi;
        while (lo > 0) {
            const double s = std::abs(A(lo - 1, lo - 1)) + std::abs(A(lo, lo));
            if (std::abs(A(lo, lo - 1)) <= 1e-15 * s) {
                A(lo, lo - 1) = 0.0;
                break;
            }
            lo--;
        }
        if (lo == hi) {
            ev[hi] = A(hi, hi);
            hi--;
            iter = 0;
            continue;
        }
        if (++iter > 60) throw std::runtime_error("hessenberg_eigenvalues: no convergence");
        // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block closer to its last entry
        const Cd p = A(hi - 1, hi - 1), q = A(hi - 1, hi), r = A(hi, hi - 1), t = A(hi, hi);
        const Cd half = 0.5 * (p + t), disc = std::sqrt(0.25 * (p - t) * (p - t) + q * r);
        const Cd m1 = half + disc, m2 = half - disc;
        Cd mu = std::abs(m1 - t) < std::abs(m2 - t) ? m1 : m2;
        if (iter % 10 == 0) mu = t + Cd(std::abs(r), 0.0);   // exceptional shift
        for (int i = lo; i <= hi; i++) A(i, i) -= mu;
        // A - mu = Q R with Q = G_lo ... G_{hi-1}; G_k^H zeroes (k+1, k)
        for (int k = lo; k < hi; k++) {
            const Cd x = A(k, k), y = A(k + 1, k);
            const double nr = std::hypot(std::abs(x), std::abs(y));
            const Cd c = nr > 0 ? x / nr : Cd(1.0), s = nr > 0 ? y / nr : Cd(0.0);
            for (int j = k; j <= hi; j++) {
                const Cd u = A(k, j), v = A(k + 1, j);
                A(k, j) = std::conj(c) * u + std::conj(s) * v;
                A(k + 1, j) = -s * u + c * v;
            }
            cs[k] = c;
            sn[k] = s;
        }
        // R Q + mu
        for (int k = lo; k < hi; k++) {
            const Cd c = cs[k], s = sn[k];
            for (int i = lo; i <= hi; i++) {
                const Cd u = A(i, k), v = A(i, k + 1);
                A(i, k) = u * c + v * s;
                A(i, k + 1) = -u * std::conj(s) + v * std::conj(c);
            }
        }
        for (int i = lo; i <= hi; i++) A(i, i) += mu;
    }
    return ev;
}

// Harmonic Ritz values of H ((d+1) x d, row-major): eigenvalues of
// H_d + |h_{d+1,d}|^2 f e_d^T with H_d^H f = e_d (Gaussian elimination with partial
// pivoting).  Throws if H_d is singular.
inline std::vector<Cd> harmonic_ritz(const std::vector<Cd>& H, int d) {
    // M = H_d^H, augmented with e_d
    std::vector<Cd> M((size_t)d * (d + 1));
    for (int i = 0; i < d; i++) {
        for (int j = 0; j < d; j++) M[(size_t)i * (d + 1) + j] = std::conj(H[(size_t)j * d + i]);
        M[(size_t)i * (d + 1) + d] = (i == d - 1) ? 1.0 : 0.0;
    }
    for (int c = 0; c < d; c++) {
        int piv = c;
        for (int i = c + 1; i < d; i++)
            if (std::abs(M[(size_t)i * (d + 1) + c]) > std::abs(M[(size_t)piv * (d + 1) + c])) piv = i;
        if (std::abs(M[(size_t)piv * (d + 1) + c]) == 0.0) throw std::runtime_error("harmonic_ritz: singular H_d");
        if (piv != c)
            for (int j = 0; j <= d; j++) std::swap(M[(size_t)c * (d + 1) + j], M[(size_t)piv * (d + 1) + j]);
        for (int i = c + 1; i < d; i++) {
            const Cd f = M[(size_t)i * (d + 1) + c] / M[(size_t)c * (d + 1) + c];
            for (int j = c; j <= d; j++) M[(size_t)i * (d + 1) + j] -= f * M[(size_t)c * (d + 1) + j];
        }
    }
    std::vector<Cd> f(d);
    for (int i = d - 1; i >= 0; i--) {
        Cd s = M[(size_t)i * (d + 1) + d];
        for (int j = i + 1; j < d; j++) s -= M[(size_t)i * (d + 1) + j] * f[j];
        f[i] = s / M[(size_t)i * (d + 1) + i];
    }
    const double h2 = std::norm(H[(size_t)d * d + d - 1]);
    std::vector<Cd> a((size_t)d * d);
    for (int i = 0; i < d; i++)
        for (int j = 0; j < d; j++) a[(size_t)i * d + j] = H[(size_t)i * d + j];
    for (int i = 0; i < d; i++) a[(size_t)i * d + d - 1] += h2 * f[i];
    return hessenberg_eigenvalues(a, d);
}

// Leja order: the root of largest modulus first, then repeatedly the remaining root
// with the largest sum of log-distances to those already taken (the lowest index on a
// tie).
inline std::vector<Cd> leja_order(std::vector<Cd> th) {
    std::vector<Cd> out;
    const int n = (int)th.size();
    std::vector<char> used(n, 0);
    std::vector<double> ls(n, 0.0);
    for (int k = 0; k < n; k++) {
        int best = -1;
        for (int i = 0; i < n; i++) {
            if (used[i]) continue;
            const double key = k == 0 ? std::abs(th[i]) : ls[i];
            const double bk = best < 0 ? 0.0 : (k == 0 ? std::abs(th[best]) : ls[best]);
            if (best < 0 || key > bk) best = i;
        }
        used[best] = 1;
        out.push_back(th[best]);
        for (int i = 0; i < n; i++)
            if (!used[i]) ls[i] += std::log(std::abs(th[i] - th[best]));
    }
    return out;
}

// The preconditioner's roots from H: harmonic Ritz values in Leja order.
inline std::vector<Cd> gmres_poly_roots(const std::vector<Cd>& H, int d) { return leja_order(harmonic_ritz(H, d)); }

}  // namespace dd
