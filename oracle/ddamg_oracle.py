"""CPU restatement (numpy, complex128) of the reference DD-alphaAMG solve path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: only tests/, the smoke
check and bench.py's cpu_baseline leg may import it, and never as the thing that
is measured or shipped.  The product path is the HIP library; it never falls back
to this file.

Every function cites the reference file:line it restates (paths relative to the
reference root).  Vectors are numpy arrays without halos, shaped
``(Nx, Nt, dof)`` (x outer, t inner, dof = colour*2 + spin), exactly the interior
ordering of the reference's halo-padded arrays.  mu = 0 is the t direction
(t+1 neighbour), mu = 1 the x direction; periodic in x, antiperiodic in t.

Pinning (what checks each function, see tests/test_oracle.py):
  D_phi / D_dagger_phi   golden vectors from the compiled reference, free-field
                         momentum-space closed form, sigma3-hermiticity, dense
                         brute force on 4x4.
  stencil (Dirac links)  golden Level::D_operator at level 0 and equality with D_phi.
  P, P^dagger            golden P_vc/Pdagg_v, P^dagger P = 1 after orthonormalisation.
  orthonormalise         golden test vectors (hierarchy dump of the reference).
  coarse links           golden G1/G2/G3, and P^dagger D P = D_c by dense brute force.
  SAP                    golden SAP(nu=1) at levels 0 and 1; exact block inverses
                         reproduce the closed-form red-black iteration.
  fgmres                 dense solve (textbook); the reference's solution and
                         iteration count for the whole FGMRES-AMG solve.
  setup_phase RNG        golden test vectors (std::mt19937 per virtual rank).
  gauge_metropolis       <cos theta_P> against I1(beta)/I0(beta).
  cg_normal, bicgstab    golden iteration counts and solutions of the compiled
                         reference (g32_methods), dense solve, and CG's finite
                         termination in at most n steps on a 4x4 lattice.
  AMG.bootstrap          parity unpinned: the reference does not implement it.
                         Checked for what the mathematics fixes: zero passes change
                         nothing, the new prolongators are orthonormal and the
                         coarse operators their Galerkin products (dense brute
                         force), and the test vectors move toward the low modes of
                         D (smaller residual of the two-grid error propagator on
                         them).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

# --------------------------------------------------------------------------------------
# Geometry helpers


def fwd(v: np.ndarray, mu: int) -> np.ndarray:
    """v(x + mu) with periodic wrap (no sign).  mu=0: t, mu=1: x."""
    return np.roll(v, -1, axis=1 if mu == 0 else 0)


def bwd(v: np.ndarray, mu: int) -> np.ndarray:
    """v(x - mu) with periodic wrap (no sign)."""
    return np.roll(v, 1, axis=1 if mu == 0 else 0)


def sign_fwd(Nt: int) -> np.ndarray:
    """rsign: -1 on the last time slice for the forward t hop (include/boundary.h:41-43,
    include/level.h:438-443)."""
    s = np.ones((1, Nt, 1))
    s[0, Nt - 1, 0] = -1.0
    return s


def sign_bwd(Nt: int) -> np.ndarray:
    """lsign: -1 on the first time slice for the backward t hop (include/boundary.h:44-46,
    include/level.h:445-450)."""
    s = np.ones((1, Nt, 1))
    s[0, 0, 0] = -1.0
    return s


# --------------------------------------------------------------------------------------
# Level-0 Wilson-Dirac operator


def D_phi(U: np.ndarray, phi: np.ndarray, m0: float) -> np.ndarray:
    """Wilson-Dirac operator, src/dirac_operator.cpp:24-54 (doc eq. 6).

    U: (Nx, Nt, 2) links U_mu(x); phi: (Nx, Nt, 2).  Term order follows the reference.
    """
    Nt = phi.shape[1]
    sp, sm = sign_fwd(Nt)[..., 0], sign_bwd(Nt)[..., 0]
    p_t, p_x = fwd(phi, 0), fwd(phi, 1)
    m_t, m_x = bwd(phi, 0), bwd(phi, 1)
    U0, U1 = U[..., 0], U[..., 1]
    U0m, U1m = np.conj(bwd(U, 0)[..., 0]), np.conj(bwd(U, 1)[..., 1])
    out = np.empty_like(phi)
    out[..., 0] = (m0 + 2) * phi[..., 0] - 0.5 * (
        U0 * sp * (p_t[..., 0] - p_t[..., 1])
        + U1 * (p_x[..., 0] + 1j * p_x[..., 1])
        + U0m * sm * (m_t[..., 0] + m_t[..., 1])
        + U1m * (m_x[..., 0] - 1j * m_x[..., 1]))
    out[..., 1] = (m0 + 2) * phi[..., 1] - 0.5 * (
        U0 * sp * (-p_t[..., 0] + p_t[..., 1])
        + U1 * (-1j * p_x[..., 0] + p_x[..., 1])
        + U0m * sm * (m_t[..., 0] + m_t[..., 1])
        + U1m * (1j * m_x[..., 0] + m_x[..., 1]))
    return out


def D_dagger_phi(U: np.ndarray, phi: np.ndarray, m0: float) -> np.ndarray:
    """Adjoint Wilson-Dirac operator, src/dirac_operator.cpp:56-85."""
    Nt = phi.shape[1]
    sp, sm = sign_fwd(Nt)[..., 0], sign_bwd(Nt)[..., 0]
    p_t, p_x = fwd(phi, 0), fwd(phi, 1)
    m_t, m_x = bwd(phi, 0), bwd(phi, 1)
    U0, U1 = U[..., 0], U[..., 1]
    U0m, U1m = np.conj(bwd(U, 0)[..., 0]), np.conj(bwd(U, 1)[..., 1])
    out = np.empty_like(phi)
    out[..., 0] = (m0 + 2) * phi[..., 0] - 0.5 * (
        U0m * sm * (m_t[..., 0] - m_t[..., 1])
        + U1m * (m_x[..., 0] + 1j * m_x[..., 1])
        + U0 * sp * (p_t[..., 0] + p_t[..., 1])
        + U1 * (p_x[..., 0] - 1j * p_x[..., 1]))
    out[..., 1] = (m0 + 2) * phi[..., 1] - 0.5 * (
        U0m * sm * (-m_t[..., 0] + m_t[..., 1])
        + U1m * (-1j * m_x[..., 0] + m_x[..., 1])
        + U0 * sp * (p_t[..., 0] + p_t[..., 1])
        + U1 * (1j * p_x[..., 0] + p_x[..., 1]))
    return out


def sigma3(v: np.ndarray) -> np.ndarray:
    """sigma_3 acting on the spin index (dof = colour*2 + spin)."""
    out = v.copy()
    out[..., 1::2] *= -1
    return out


# --------------------------------------------------------------------------------------
# Generic level stencil


@dataclass
class Stencil:
    """out = (m0+2) v - A v - sum_mu (B_mu s+ v(x+mu) + C_mu s- v(x-mu)).

    Restates Level::D_operator, src/level.cpp:394-426 (doc eq. 31/39).  A: (Nx,Nt,d,d)
    or None (level 0, G1 = 0); B, C: (2,Nx,Nt,d,d), row = colour*2+alpha, col = b*2+beta
    (the reference stores G1[n][alf][bet][c][b], include/level.h:349-365).
    """
    A: np.ndarray | None
    B: np.ndarray
    C: np.ndarray
    m0: float

    @property
    def shape(self):
        return self.B.shape[1:3]

    @property
    def dof(self):
        return self.B.shape[-1]

    def apply(self, v: np.ndarray) -> np.ndarray:
        Nt = v.shape[1]
        out = (self.m0 + 2) * v
        if self.A is not None:
            out = out - np.einsum('xtij,xtj->xti', self.A, v)
        sp, sm = sign_fwd(Nt), sign_bwd(Nt)
        for mu in (0, 1):
            f = fwd(v, mu)
            b = bwd(v, mu)
            if mu == 0:
                f = f * sp
                b = b * sm
            out = out - (np.einsum('xtij,xtj->xti', self.B[mu], f)
                         + np.einsum('xtij,xtj->xti', self.C[mu], b))
        return out


_M = np.array([[[1, -1], [-1, 1]], [[1, 1j], [-1j, 1]]], dtype=complex)   # 1 - sigma_mu
_P = np.array([[[1, 1], [1, 1]], [[1, -1j], [1j, 1]]], dtype=complex)     # 1 + sigma_mu


def dirac_stencil(U: np.ndarray, m0: float) -> Stencil:
    """Level-0 link matrices, Level::makeDirac src/level.cpp:286-331 (doc eq. 40):
    B_mu = 1/2 (1 - sigma_mu) U_mu(x),  C_mu = 1/2 (1 + sigma_mu) U_mu(x - mu)^*,  A = 0."""
    Nx, Nt = U.shape[:2]
    B = np.empty((2, Nx, Nt, 2, 2), complex)
    C = np.empty((2, Nx, Nt, 2, 2), complex)
    for mu in (0, 1):
        B[mu] = 0.5 * _M[mu][None, None] * U[..., mu][..., None, None]
        C[mu] = 0.5 * _P[mu][None, None] * np.conj(bwd(U, mu)[..., mu])[..., None, None]
    return Stencil(None, B, C, m0)


# --------------------------------------------------------------------------------------
# Aggregation: P, P^dagger, orthonormalisation, Galerkin links


def restrict(W: np.ndarray, v: np.ndarray, agg: tuple[int, int]) -> np.ndarray:
    """P^dagger v, Level::Pdagg_v src/level.cpp:145-189 (doc eq. 22).

    W: (Nv, Nx, Nt, dof) test vectors; returns (Nx/ax, Nt/at, 2*Nv), dof = cc*2 + s.
    """
    Nv, Nx, Nt, d = W.shape
    ax, at = agg
    col = d // 2
    Wb = W.reshape(Nv, Nx // ax, ax, Nt // at, at, col, 2)
    vb = v.reshape(Nx // ax, ax, Nt // at, at, col, 2)
    out = np.einsum('pxaybcs,xaybcs->xyps', np.conj(Wb), vb)
    return out.reshape(Nx // ax, Nt // at, 2 * Nv)


def prolong(W: np.ndarray, vc: np.ndarray, agg: tuple[int, int]) -> np.ndarray:
    """P v_c, Level::P_vc src/level.cpp:92-141 (doc eq. 21)."""
    Nv, Nx, Nt, d = W.shape
    ax, at = agg
    col = d // 2
    Wb = W.reshape(Nv, Nx // ax, ax, Nt // at, at, col, 2)
    vcb = vc.reshape(Nx // ax, Nt // at, Nv, 2)
    out = np.einsum('pxaybcs,xyps->xaybcs', Wb, vcb)
    return out.reshape(Nx, Nt, d)


def orthonormalize(W: np.ndarray, agg: tuple[int, int]) -> np.ndarray:
    """Block-local modified Gram-Schmidt per (aggregate, spin), Level::orthonormalize
    src/level.cpp:193-284: ascending nt, proj = sum v_nt conj(v_ntt), normalise only if
    the norm exceeds 1e-8."""
    Nv, Nx, Nt, d = W.shape
    ax, at = agg
    col = d // 2
    Wb = W.reshape(Nv, Nx // ax, ax, Nt // at, at, col, 2).copy()
    # gather each (block, spin) aggregate into a vector: (Nv, Bx, By, s, ax*at*col)
    A = np.transpose(Wb, (0, 1, 3, 6, 2, 4, 5)).reshape(Nv, Nx // ax, Nt // at, 2, ax * at * col)
    for nt in range(Nv):
        for ntt in range(nt):
            proj = np.sum(A[nt] * np.conj(A[ntt]), axis=-1)
            A[nt] = A[nt] - proj[..., None] * A[ntt]
        norm = np.sqrt(np.sum(np.abs(A[nt]) ** 2, axis=-1))
        safe = norm > 1e-8
        A[nt] = np.where(safe[..., None], A[nt] / np.where(safe, norm, 1.0)[..., None], A[nt])
    Wb = np.transpose(A.reshape(Nv, Nx // ax, Nt // at, 2, ax, at, col), (0, 1, 4, 2, 5, 6, 3))
    return Wb.reshape(Nv, Nx, Nt, d)


def _site_prolongator(W: np.ndarray) -> np.ndarray:
    """P restricted to each fine site: (Nx, Nt, dof_l, 2Nv), P[(c,a),(p,a')] = w_p(x,c,a) d_aa'."""
    Nv, Nx, Nt, d = W.shape
    col = d // 2
    Ps = np.zeros((Nx, Nt, col, 2, Nv, 2), complex)
    Wt = np.transpose(W.reshape(Nv, Nx, Nt, col, 2), (1, 2, 3, 4, 0))  # x t c s p
    for s in (0, 1):
        Ps[:, :, :, s, :, s] = Wt[:, :, :, s, :]
    return Ps.reshape(Nx, Nt, d, 2 * Nv)


def coarse_links(op: Stencil, W: np.ndarray, agg: tuple[int, int], m0: float) -> Stencil:
    """Galerkin coarse operator of the next level, Level::makeCoarseLinks
    src/level.cpp:431-542 (doc eqs. 27-29, 37).

    A_c(X) = sum_{x in X} P_x^+ [A(x) P_x + hops that stay in X];  B_c,mu(X) collects the
    forward hops that leave X through its +mu face, C_c,mu(X) the backward hops through
    its -mu face.  No antiperiodic sign enters here: the coarse stencil applies it
    (include/level.h:438-450).
    """
    Nv, Nx, Nt, d = W.shape
    ax, at = agg
    Xc, Tc = Nx // ax, Nt // at
    Ps = _site_prolongator(W)                  # x t d 2Nv
    PsH = np.conj(np.swapaxes(Ps, -1, -2))     # x t 2Nv d
    D2 = 2 * Nv
    Ac = np.zeros((Xc, Tc, D2, D2), complex)
    Bc = np.zeros((2, Xc, Tc, D2, D2), complex)
    Cc = np.zeros((2, Xc, Tc, D2, D2), complex)

    def blocksum(M):
        return M.reshape(Xc, ax, Tc, at, D2, D2).sum(axis=(1, 3))

    if op.A is not None:
        Ac += blocksum(PsH @ op.A @ Ps)
    xs = np.arange(Nx)[:, None]
    ts = np.arange(Nt)[None, :]
    for mu in (0, 1):
        lf = fwd(Ps, mu)
        lb = bwd(Ps, mu)
        termB = PsH @ op.B[mu] @ lf
        termC = PsH @ op.C[mu] @ lb
        if mu == 0:
            cross_f = np.broadcast_to(ts % at == at - 1, (Nx, Nt))
            cross_b = np.broadcast_to(ts % at == 0, (Nx, Nt))
        else:
            cross_f = np.broadcast_to(xs % ax == ax - 1, (Nx, Nt))
            cross_b = np.broadcast_to(xs % ax == 0, (Nx, Nt))
        Ac += blocksum(np.where(cross_f[..., None, None], 0, termB))
        Ac += blocksum(np.where(cross_b[..., None, None], 0, termC))
        Bc[mu] += blocksum(np.where(cross_f[..., None, None], termB, 0))
        Cc[mu] += blocksum(np.where(cross_b[..., None, None], termC, 0))
    return Stencil(Ac, Bc, Cc, m0)


# --------------------------------------------------------------------------------------
# FGMRES (batched over independent systems, reference control flow per system)


def _dot(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """dot(A, B) = sum A conj(B) per system (include/utils.h:68-84)."""
    nb = a.shape[0]
    return np.sum((a * np.conj(b)).reshape(nb, -1), axis=1)


def _bc(s: np.ndarray, like: np.ndarray) -> np.ndarray:
    return s.reshape((-1,) + (1,) * (like.ndim - 1))


def cg_normal(U: np.ndarray, b: np.ndarray, m0: float, max_iter: int, tol: float):
    """Methods::CG (src/methods.cpp:43-59): conjugate_gradient
    (src/conjugate_gradient.cpp:4-84) on D^dagger D x = D^dagger b for a batch b[nb].

    x0 = D^dagger b (`sol = phi`); r = phi - D^dagger D x; alpha = (r,r)/(d,Ad);
    stop when ||r|| < tol ||phi|| with count k+1; beta = (r',r')/(r,r); d = beta d + r.
    Returns (x, iterations) with iterations 0 where the reference reports a failure.
    """
    DdD = lambda v: np.stack([D_dagger_phi(U, D_phi(U, vi, m0), m0) for vi in v])
    phi = np.stack([D_dagger_phi(U, bi, m0) for bi in b])
    x = phi.copy()
    r = phi - DdD(x)
    d = r.copy()
    rn2 = _dot(r, r)
    nphi = np.sqrt(np.real(_dot(phi, phi)))
    its = np.zeros(b.shape[0], int)
    run = np.ones(b.shape[0], bool)
    for k in range(max_iter):
        Ad = DdD(d)
        alpha = np.where(run, rn2 / np.where(run, _dot(d, Ad), 1), 0)
        x = x + _bc(alpha, x) * d
        r = r - _bc(alpha, r) * Ad
        e2 = np.real(_dot(r, r))
        conv = run & (np.sqrt(e2) < tol * nphi)
        its[conv] = k + 1
        run &= ~conv
        if not run.any():
            break
        beta = np.where(run, e2 / np.where(run, rn2, 1), 0)
        d = np.where(_bc(run, d), _bc(beta, d) * d + r, d)
        rn2 = np.where(run, e2, rn2)
    return x, its


def bicgstab(U: np.ndarray, b: np.ndarray, x0: np.ndarray, m0: float, max_iter: int, tol: float):
    """bi_cgstab (src/conjugate_gradient.cpp:87-176) on D x = b for a batch b[nb].

    r~ = r_0 = b - D x0; rho = (r, r~); d = r (k = 0) or r + beta (d - omega Ad) with
    beta = alpha rho / (omega rho_prev); alpha = rho / (Ad, r~); s = r - alpha Ad; stop
    when ||s|| < tol ||b|| after x += alpha d (count k+1); omega = (s,t)/(t,t), t = D s;
    r = s - omega t; x += alpha d + omega s.  Count max_iter on failure.
    """
    Dm = lambda v: np.stack([D_phi(U, vi, m0) for vi in v])
    nb = b.shape[0]
    x = x0.copy()
    r = b - Dm(x)
    rt = r.copy()
    nphi = np.sqrt(np.real(_dot(b, b)))
    its = np.full(nb, max_iter)
    run = np.ones(nb, bool)
    alpha = np.zeros(nb, complex)
    omega = np.zeros(nb, complex)
    rho_prev = np.zeros(nb, complex)
    d = np.zeros_like(r)
    Ad = np.zeros_like(r)
    for k in range(max_iter):
        rho = _dot(r, rt)
        if k == 0:
            d = r.copy()
        else:
            beta = np.where(run, alpha * rho / np.where(run, omega * rho_prev, 1), 0)
            d = np.where(_bc(run, d), r + _bc(beta, d) * (d - _bc(omega, d) * Ad), d)
        Ad = np.where(_bc(run, d), Dm(d), Ad)
        alpha = np.where(run, rho / np.where(run, _dot(Ad, rt), 1), alpha)
        sv = r - _bc(alpha, r) * Ad
        conv = run & (np.sqrt(np.real(_dot(sv, sv))) < tol * nphi)
        x = np.where(_bc(conv, x), x + _bc(alpha, x) * d, x)
        its[conv] = k + 1
        run &= ~conv
        if not run.any():
            break
        t = Dm(sv)
        om = _dot(sv, t) / np.where(run, _dot(t, t), 1)
        omega = np.where(run, om, omega)
        r = np.where(_bc(run, r), sv - _bc(omega, sv) * t, r)
        x = np.where(_bc(run, x), x + _bc(alpha, x) * d + _bc(omega, x) * sv, x)
        rho_prev = np.where(run, rho, rho_prev)
    return x, its


@dataclass
class FgmresResult:
    x: np.ndarray
    iterations: np.ndarray
    steps: int = 0        # total Arnoldi steps taken (max over systems per step)


def fgmres(func, precond, phi: np.ndarray, x0: np.ndarray, m: int, restarts: int, tol: float
           ) -> FgmresResult:
    """Flexible GMRES(m), FGMRES::fgmres src/fgmres.cpp:5-97, rotation :99-115,
    solve_upper_triangular :118-127, for a batch of independent systems.

    Reproduces the reference's control flow per system: modified Gram-Schmidt, the
    early break on |g_{j+1}| < tol ||phi||, maxIt that is not reset between restarts
    (:14, :56), the true-residual test after each cycle, the k*m + maxIt count, and the
    Z vectors that persist across restarts within a call (preconditioners that use their
    output as an initial guess see the previous cycle's Z_j).
    precond(v, z_old) -> z_new; None means identity (Z_j = V_j).
    """
    nb = phi.shape[0]
    V = np.zeros((m + 1,) + phi.shape, complex)
    Z = np.zeros((m,) + phi.shape, complex)
    H = np.zeros((nb, m + 1, m), complex)
    g = np.zeros((nb, m + 1), complex)
    sn = np.zeros((nb, m), complex)
    cn = np.zeros((nb, m), complex)
    x = x0.copy()
    r = phi - func(x)
    norm_phi = np.sqrt(np.real(_dot(phi, phi)))
    err = np.sqrt(np.real(_dot(r, r)))
    maxIt = np.full(nb, m)
    done = np.zeros(nb, bool)
    iters = np.full(nb, restarts * m)
    steps = 0
    for k in range(restarts):
        live = ~done
        beta = err.astype(complex)
        V[0] = np.where(_bc(live, r), r / _bc(np.where(live, beta, 1.0), r), V[0])
        g[live, 0] = beta[live]
        running = live.copy()
        for j in range(m):
            if not running.any():
                break
            steps += 1
            zj = V[j].copy() if precond is None else precond(V[j], Z[j])
            Z[j] = np.where(_bc(running, zj), zj, Z[j])
            w = func(Z[j])
            for i in range(j + 1):
                h = _dot(w, V[i])
                H[running, i, j] = h[running]
                w = w - _bc(H[:, i, j], w) * V[i]
            hn = np.sqrt(np.real(_dot(w, w)))
            H[running, j + 1, j] = hn[running]
            upd = running & (hn > 0)
            V[j + 1] = np.where(_bc(upd, w), w / _bc(np.where(upd, hn, 1.0), w), V[j + 1])
            for b in np.nonzero(running)[0]:
                _rotation(H[b], sn[b], cn[b], j)
                g[b, j + 1] = -sn[b, j] * g[b, j]
                g[b, j] = np.conj(cn[b, j]) * g[b, j]
            brk = running & (np.abs(g[:, j + 1]) < tol * norm_phi)
            maxIt[brk] = j + 1
            running &= ~brk
        for b in np.nonzero(live)[0]:
            n = maxIt[b]
            eta = _solve_upper(H[b], g[b], n)
            xb = x[b]
            for j in range(n):
                xb = xb + eta[j] * Z[j][b]
            x[b] = xb
        r_new = phi - func(x)
        err_new = np.sqrt(np.real(_dot(r_new, r_new)))
        r = np.where(_bc(live, r), r_new, r)
        err = np.where(live, err_new, err)
        conv = live & (err < tol * norm_phi)
        iters[conv] = k * m + maxIt[conv]
        done |= conv
        if done.all():
            break
    return FgmresResult(x, iters, steps)


def _rotation(H: np.ndarray, sn: np.ndarray, cn: np.ndarray, j: int) -> None:
    """FGMRES::rotation src/fgmres.cpp:99-115 (complex Givens)."""
    for i in range(j):
        temp = np.conj(cn[i]) * H[i, j] + np.conj(sn[i]) * H[i + 1, j]
        H[i + 1, j] = -sn[i] * H[i, j] + cn[i] * H[i + 1, j]
        H[i, j] = temp
    den = math.sqrt(np.real(np.conj(H[j, j]) * H[j, j] + np.conj(H[j + 1, j]) * H[j + 1, j]))
    sn[j] = H[j + 1, j] / den
    cn[j] = H[j, j] / den
    H[j, j] = np.conj(cn[j]) * H[j, j] + np.conj(sn[j]) * H[j + 1, j]
    H[j + 1, j] = 0.0


def _solve_upper(H: np.ndarray, g: np.ndarray, n: int) -> np.ndarray:
    """FGMRES::solve_upper_triangular src/fgmres.cpp:118-127."""
    out = np.zeros(n, complex)
    for i in range(n - 1, -1, -1):
        out[i] = g[i]
        for j in range(i + 1, n):
            out[i] -= H[i, j] * out[j]
        out[i] /= H[i, i]
    return out


# --------------------------------------------------------------------------------------
# SAP smoother


@dataclass
class SapBlocks:
    """Red-black Schwarz blocks of bx x bt sites, SAP_C::SchwarzBlocks src/sap.cpp:3-41.

    Colour: red iff (block_x + block_t) is even.  The reference colours the blocks of each
    MPI rank locally (x*Block_t + t parity, :28-38); with an even number of blocks per rank
    in each direction that is the same as this global rule.
    """
    Nx: int
    Nt: int
    bx: int
    bt: int

    @property
    def nbx(self):
        return self.Nx // self.bx

    @property
    def nbt(self):
        return self.Nt // self.bt

    def colour_ids(self, colour: int) -> np.ndarray:
        ids = [(X, T) for X in range(self.nbx) for T in range(self.nbt) if (X + T) % 2 == colour]
        return np.array(ids, dtype=int).reshape(-1, 2)


def _gather_blocks(v: np.ndarray, blocks: SapBlocks, ids: np.ndarray) -> np.ndarray:
    """(Nx, Nt, ...) -> (nblk, bx, bt, ...) for the listed blocks."""
    bx, bt = blocks.bx, blocks.bt
    r = v.reshape((blocks.nbx, bx, blocks.nbt, bt) + v.shape[2:])
    r = np.moveaxis(r, 2, 1)   # X T bx bt ...
    return r[ids[:, 0], ids[:, 1]]


def _scatter_add_blocks(x: np.ndarray, e: np.ndarray, blocks: SapBlocks, ids: np.ndarray) -> np.ndarray:
    bx, bt = blocks.bx, blocks.bt
    r = x.reshape((blocks.nbx, bx, blocks.nbt, bt) + x.shape[2:]).copy()
    r = np.moveaxis(r, 2, 1)
    r[ids[:, 0], ids[:, 1]] += e
    r = np.moveaxis(r, 1, 2)
    return r.reshape(x.shape)


def block_operator(op: Stencil, blocks: SapBlocks, ids: np.ndarray):
    """D restricted to each block with zero data outside and no boundary sign,
    Level::SAP_level_l::D_local src/level.cpp:607-653.  Returns f(vb) for vb of shape
    (nblk, bx, bt, dof)."""
    A = None if op.A is None else _gather_blocks(op.A, blocks, ids)
    B = [_gather_blocks(op.B[mu], blocks, ids) for mu in (0, 1)]
    C = [_gather_blocks(op.C[mu], blocks, ids) for mu in (0, 1)]
    m0 = op.m0

    def shift(vb, mu, direction):
        out = np.zeros_like(vb)
        ax = 2 if mu == 0 else 1
        sl_src = [slice(None)] * vb.ndim
        sl_dst = [slice(None)] * vb.ndim
        if direction > 0:   # v(x+mu)
            sl_src[ax] = slice(1, None)
            sl_dst[ax] = slice(0, -1)
        else:
            sl_src[ax] = slice(0, -1)
            sl_dst[ax] = slice(1, None)
        out[tuple(sl_dst)] = vb[tuple(sl_src)]
        return out

    def f(vb):
        out = (m0 + 2) * vb
        if A is not None:
            out = out - np.einsum('nxtij,nxtj->nxti', A, vb)
        for mu in (0, 1):
            out = out - (np.einsum('nxtij,nxtj->nxti', B[mu], shift(vb, mu, +1))
                         + np.einsum('nxtij,nxtj->nxti', C[mu], shift(vb, mu, -1)))
        return out

    return f


@dataclass
class SapParams:
    """Block-GMRES knobs, src/variables.cpp:100-105."""
    m: int = 5
    restarts: int = 5
    tol: float = 1e-3


def sap(op: Stencil, blocks: SapBlocks, v: np.ndarray, x: np.ndarray, nu: int, tol: float,
        prm: SapParams = SapParams(), stats: dict | None = None,
        block_op: Stencil | None = None) -> tuple[np.ndarray, int]:
    """Red-black multiplicative Schwarz, SAP_C::SAP src/sap.cpp:81-165 with the block solves
    of SAP_C::I_D_B_1_It :48-77 (GMRES(m) x restarts, tol, initial guess = block rhs).

    All red blocks see the same residual; the global residual is recomputed between
    colours and once more at the end of each iteration.  Returns (x, iterations).
    block_op: the operator whose blocks the block solves invert, when it is not op (the
    device's binary16 link copy in the FP32 hierarchy's smoother; the residuals keep op).
    """
    x = x.copy()
    v_norm = math.sqrt(np.real(np.vdot(v, v)))
    r = v - op.apply(x)
    err = 0.0
    for i in range(nu):
        for colour in (0, 1):
            ids = blocks.colour_ids(colour)
            rb = _gather_blocks(r, blocks, ids)
            res = fgmres(block_operator(block_op or op, blocks, ids), None, rb, rb, prm.m, prm.restarts, prm.tol)
            if stats is not None:
                stats.setdefault('block_iters', []).append(res.iterations.copy())
            x = _scatter_add_blocks(x, res.x, blocks, ids)
            r = v - op.apply(x)
        err = math.sqrt(np.real(np.vdot(r, r)))
        if err < tol * v_norm:
            return x, i
    return x, nu


# --------------------------------------------------------------------------------------
# Hierarchy, setup and cycles


@dataclass
class LevelSpec:
    """One row of the reference's parameters file (include/params.h:7-57), expressed in
    sites: aggregate block size and SAP block size at this level, and the number of test
    vectors used to build the next level."""
    agg: tuple[int, int]
    sap_block: tuple[int, int]
    nv: int


@dataclass
class AmgParams:
    """Solver knobs with the reference defaults (src/variables.cpp:90-134)."""
    levels: list[LevelSpec]
    m0: float
    nu1: int = 0
    nu2: int = 2
    setup_sap_iters: int = 4
    sap: SapParams = field(default_factory=SapParams)
    coarse_m: int = 400
    coarse_restarts: int = 10
    coarse_tol: float = 0.1
    kcycle_m: int = 5
    kcycle_restarts: int = 2
    kcycle_tol: float = 0.1
    fgmres_m: int = 20
    fgmres_restarts: int = 50
    tol: float = 1e-10
    rng_ranks: tuple[int, int] = (2, 2)


def mt19937_uniform(seed: int, n: int) -> np.ndarray:
    """n draws of std::uniform_real_distribution<double>(-1, 1) on std::mt19937(seed), as
    libstdc++ computes them (generate_canonical with two 32-bit words: (w1 + w2 2^32) / 2^64)."""
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    raw = bg.random_raw(2 * n).astype(np.float64).reshape(n, 2)
    s = raw[:, 0] + raw[:, 1] * 4294967296.0
    u = s / 18446744073709551616.0
    u = np.where(u >= 1.0, np.nextafter(1.0, 0.0), u)
    return -1.0 + 2.0 * u


class _RankStreams:
    """One std::mt19937(rank2d) per virtual MPI rank, continued across levels as the
    function-local static engine of AlgebraicMG::setUpPhase is (src/amg.cpp:6-7)."""

    def __init__(self, ranks):
        self.ranks = ranks
        self.pos = {r: 0 for r in range(ranks[0] * ranks[1])}
        self.cache = {}

    def draw(self, r, n):
        p = self.pos[r]
        need = p + n
        if r not in self.cache or self.cache[r].size < need:
            self.cache[r] = mt19937_uniform(r, max(need, 2 * need))
        self.pos[r] = need
        return self.cache[r][p:need]


def random_test_vectors(streams: _RankStreams, count: int, Nx: int, Nt: int, dof: int,
                        first: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    """Random test vectors as AlgebraicMG::setUpPhase draws them (src/amg.cpp:10-21, :33-39):
    per rank, loop cc -> local x -> local t -> dof, each entry consuming two uniforms.  The
    reference writes `d(g) + I*d(g)`; g++ evaluates the imaginary operand first, so the
    first draw is the imaginary part and the second the real part."""
    Rx, Rt = streams.ranks
    wx, wt = Nx // Rx, Nt // Rt
    if out is None:
        out = np.zeros((first + count, Nx, Nt, dof), complex)
    for cx in range(Rx):
        for ct in range(Rt):
            r = cx * Rt + ct
            d = streams.draw(r, 2 * count * wx * wt * dof).reshape(count, wx, wt, dof, 2)
            out[first:first + count, cx * wx:(cx + 1) * wx, ct * wt:(ct + 1) * wt] = d[..., 1] + 1j * d[..., 0]
    return out


@dataclass
class Level:
    op: Stencil
    spec: LevelSpec | None       # None on the coarsest level
    W: np.ndarray | None = None  # orthonormal test vectors (prolongator)
    V: np.ndarray | None = None  # the smoothed test vectors before orthonormalisation


class AMG:
    """AlgebraicMG (include/amg.h:96-134, src/amg.cpp) with setup, V- and K-cycle."""

    def __init__(self, U: np.ndarray, prm: AmgParams):
        self.U = U
        self.prm = prm
        self.levels: list[Level] = []
        self.stats: dict = {}

    # -- setup ------------------------------------------------------------------------
    def setup(self, raw_test_vectors: list[np.ndarray] | None = None) -> None:
        """AlgebraicMG::setUpPhase src/amg.cpp:4-62."""
        prm = self.prm
        L = len(prm.levels) + 1
        op0 = dirac_stencil(self.U, prm.m0)
        Nx, Nt = self.U.shape[:2]
        shapes = [(Nx, Nt, 2)]
        for spec in prm.levels:
            px, pt, pd = shapes[-1]
            shapes.append((px // spec.agg[0], pt // spec.agg[1], 2 * spec.nv))
        streams = _RankStreams(prm.rng_ranks)
        tv: list[np.ndarray] = []
        if raw_test_vectors is not None:
            tv = [w.copy() for w in raw_test_vectors]
        else:
            nx, nt, d = shapes[0]
            tv.append(random_test_vectors(streams, prm.levels[0].nv, nx, nt, d))
            for l in range(1, L - 1):
                nx, nt, d = shapes[l]
                nv = prm.levels[l].nv
                nprev = prm.levels[l - 1].nv
                w = np.zeros((nv, nx, nt, d), complex)
                for cc in range(min(nv, nprev)):
                    w[cc] = restrict(tv[l - 1], tv[l - 1][cc], prm.levels[l - 1].agg)
                if nv > nprev:
                    random_test_vectors(streams, nv - nprev, nx, nt, d, first=nprev, out=w)
                tv.append(w)
        op = op0
        self.levels = []
        for l in range(L - 1):
            spec = prm.levels[l]
            blocks = SapBlocks(op.shape[0], op.shape[1], *spec.sap_block)
            zero = np.zeros(op.shape + (op.dof,), complex)
            W = tv[l]
            for cc in range(spec.nv):
                W[cc], _ = sap(op, blocks, zero, W[cc], prm.setup_sap_iters, 1e-10, prm.sap)
            V = W.copy()
            W = orthonormalize(W, spec.agg)
            self.levels.append(Level(op, spec, W, V))
            op = coarse_links(op, W, spec.agg, prm.m0)
        self.levels.append(Level(op, None))

    def bootstrap(self, passes: int, cycle: str = 'K') -> None:
        """Bootstrap adaptivity as AMG_Documentation.pdf section 3 describes it ("use the
        current multilevel method defined by P to approximately solve D x = (w_i - D w_i)
        and update w_i <- x; the resultant test vectors are orthonormalized again"); the
        reference declares AMGV::Nit (include/variables.h:152) but AlgebraicMG::setUpPhase
        (src/amg.cpp:4-62) never uses it, so no reference output pins this function.

        Per pass and per level l (finest first): w_i are the test vectors as they were
        before their orthonormalisation (the aggregate-wise orthonormalised ones are not
        smooth across aggregates; started from them the pass degrades the prolongator);
        the approximate solve is one cycle from level l with zero initial guess; the new
        vectors replace w_i, are orthonormalised per aggregate into the prolongator, and
        the coarse operators of levels l+1 .. are rebuilt from the current prolongators.
        """
        cyc = self.k_cycle if cycle == 'K' else self.v_cycle
        L = len(self.levels)
        for _ in range(passes):
            for l in range(L - 1):
                lv = self.levels[l]
                X = np.empty_like(lv.V)
                for i in range(lv.V.shape[0]):
                    w = lv.V[i]
                    X[i] = cyc(l, w - lv.op.apply(w), np.zeros_like(w))
                lv.V = X
                lv.W = orthonormalize(X, lv.spec.agg)
                for k in range(l, L - 1):
                    lk = self.levels[k]
                    self.levels[k + 1].op = coarse_links(lk.op, lk.W, lk.spec.agg, self.prm.m0)

    # -- cycles -----------------------------------------------------------------------
    def _count(self, key):
        self.stats[key] = self.stats.get(key, 0) + 1

    def coarsest_solve(self, eta: np.ndarray) -> np.ndarray:
        """GMRES_level_l on the coarsest level (include/level.h:70-140), x0 = rhs
        (src/amg.cpp:69, :124)."""
        prm = self.prm
        op = self.levels[-1].op
        res = fgmres(lambda v: op.apply(v[0])[None], None, eta[None], eta[None],
                     prm.coarse_m, prm.coarse_restarts, prm.coarse_tol)
        self.stats.setdefault('coarsest_iters', []).append(int(res.iterations[0]))
        return res.x[0]

    def _smooth(self, l, eta, psi, nu):
        lv = self.levels[l]
        blocks = SapBlocks(lv.op.shape[0], lv.op.shape[1], *lv.spec.sap_block)
        psi, _ = sap(lv.op, blocks, eta, psi, nu, 1e-10, self.prm.sap)
        return psi

    def k_cycle(self, l: int, eta: np.ndarray, psi: np.ndarray) -> np.ndarray:
        """AlgebraicMG::k_cycle src/amg.cpp:119-174; psi is the in/out initial guess."""
        prm = self.prm
        if l == len(self.levels) - 1:
            return self.coarsest_solve(eta)
        lv = self.levels[l]
        psi = psi.copy()
        if prm.nu1 > 0:
            psi = self._smooth(l, eta, psi, prm.nu1)
        r = eta - lv.op.apply(psi)
        eta1 = restrict(lv.W, r, lv.spec.agg)
        nxt = self.levels[l + 1].op
        res = fgmres(lambda v: nxt.apply(v[0])[None],
                     lambda v, z: self.k_cycle(l + 1, v[0], z[0])[None],
                     eta1[None], eta1[None], prm.kcycle_m, prm.kcycle_restarts, prm.kcycle_tol)
        self.stats.setdefault(f'kcycle_iters_l{l + 1}', []).append(int(res.iterations[0]))
        psi = psi + prolong(lv.W, res.x[0], lv.spec.agg)
        if prm.nu2 > 0:
            psi = self._smooth(l, eta, psi, prm.nu2)
        return psi

    def v_cycle(self, l: int, eta: np.ndarray, psi: np.ndarray) -> np.ndarray:
        """AlgebraicMG::v_cycle src/amg.cpp:64-116."""
        prm = self.prm
        if l == len(self.levels) - 1:
            return self.coarsest_solve(eta)
        lv = self.levels[l]
        psi = psi.copy()
        if prm.nu1 > 0:
            psi = self._smooth(l, eta, psi, prm.nu1)
        r = eta - lv.op.apply(psi)
        eta1 = restrict(lv.W, r, lv.spec.agg)
        psi1 = self.v_cycle(l + 1, eta1, np.zeros_like(eta1))
        psi = psi + prolong(lv.W, psi1, lv.spec.agg)
        if prm.nu2 > 0:
            psi = self._smooth(l, eta, psi, prm.nu2)
        return psi

    def solve(self, b: np.ndarray, cycle: str = 'K') -> FgmresResult:
        """FGMRES_AMG_{k,v}_cycle (include/amg.h:147-270): outer FGMRES with D_phi as the
        operator and a cycle from level 0 with zero initial guess as the preconditioner."""
        prm = self.prm
        cyc = self.k_cycle if cycle == 'K' else self.v_cycle
        U, m0 = self.U, prm.m0
        x0 = np.zeros_like(b)
        return fgmres(lambda v: D_phi(U, v[0], m0)[None],
                      lambda v, z: cyc(0, v[0], np.zeros_like(v[0]))[None],
                      b[None], x0[None], prm.fgmres_m, prm.fgmres_restarts, prm.tol)


# --------------------------------------------------------------------------------------
# Gauge fields and file formats


def plaquette_angle(theta: np.ndarray) -> np.ndarray:
    """theta_P(x) = theta_0(x) + theta_1(x+0) - theta_0(x+1) - theta_1(x); mu=0 is t."""
    t0, t1 = theta[..., 0], theta[..., 1]
    return t0 + np.roll(t1, -1, axis=1) - np.roll(t0, -1, axis=0) - t1


def gauge_metropolis(Nx: int, Nt: int, beta: float, sweeps: int, seed: int,
                     step: float = 0.5) -> np.ndarray:
    """Seeded quenched Metropolis on S = beta sum_x (1 - cos theta_P(x)), cold start.

    The reference ships no generator (its tests read external configurations,
    src/tests.cpp:556-566); this is the generator the workloads name.  Links that share
    no plaquette are updated together: mu=0 links by parity of x, mu=1 links by parity
    of t.  Proposal theta + U(-step, step), accept with min(1, exp(-dS)).  Returns theta
    of shape (Nx, Nt, 2)."""
    rng = np.random.default_rng(seed)
    theta = np.zeros((Nx, Nt, 2))
    for _ in range(sweeps):
        for mu in (0, 1):
            for par in (0, 1):
                prop = rng.uniform(-step, step, size=(Nx, Nt))
                acc = rng.uniform(0.0, 1.0, size=(Nx, Nt))
                idx = np.arange(Nx)[:, None] if mu == 0 else np.arange(Nt)[None, :]
                mask = np.broadcast_to(idx % 2 == par, (Nx, Nt))
                old = theta[..., mu].copy()
                P = plaquette_angle(theta)
                # the two plaquettes containing link (x, mu)
                if mu == 0:
                    s_old = np.cos(P) + np.cos(np.roll(P, 1, axis=0))
                    th = theta.copy()
                    th[..., 0] = np.where(mask, old + prop, old)
                    Pn = plaquette_angle(th)
                    s_new = np.cos(Pn) + np.cos(np.roll(Pn, 1, axis=0))
                else:
                    s_old = np.cos(P) + np.cos(np.roll(P, 1, axis=1))
                    th = theta.copy()
                    th[..., 1] = np.where(mask, old + prop, old)
                    Pn = plaquette_angle(th)
                    s_new = np.cos(Pn) + np.cos(np.roll(Pn, 1, axis=1))
                dS = -beta * (s_new - s_old)
                accept = mask & (acc < np.exp(-np.maximum(dS, -50.0)))
                theta[..., mu] = np.where(accept, old + prop, old)
    return np.mod(theta + np.pi, 2 * np.pi) - np.pi


def write_records(path: str, arr: np.ndarray) -> None:
    """Reference binary record format, read_binary src/io.cpp:32-46: per (x, t, mu) one
    record int x, int t, int mu, double re, double im (28 bytes), x outer, t, mu inner."""
    Nx, Nt, n = arr.shape
    rec = np.zeros((Nx, Nt, n), dtype=[('x', '<i4'), ('t', '<i4'), ('mu', '<i4'), ('re', '<f8'), ('im', '<f8')])
    X, T, M = np.meshgrid(np.arange(Nx), np.arange(Nt), np.arange(n), indexing='ij')
    rec['x'], rec['t'], rec['mu'] = X, T, M
    rec['re'], rec['im'] = arr.real, arr.imag
    rec.tofile(path)


def read_records(path: str, Nx: int, Nt: int, n: int = 2) -> np.ndarray:
    rec = np.fromfile(path, dtype=[('x', '<i4'), ('t', '<i4'), ('mu', '<i4'), ('re', '<f8'), ('im', '<f8')])
    rec = rec.reshape(Nx, Nt, n)
    return rec['re'] + 1j * rec['im']


def read_rank_dump(prefix: str, nranks: int) -> np.ndarray:
    """Assemble the per-rank dumps written by oracle/ref_driver.cpp."""
    parts = []
    for r in range(nranks):
        with open(f'{prefix}.r{r}.bin', 'rb') as f:
            cx, ct, nx, nt, ps = struct.unpack('<5i', f.read(20))
            data = np.frombuffer(f.read(), dtype=np.complex128).reshape(nx, nt, ps)
            parts.append((cx, ct, nx, nt, data))
    Rx = max(p[0] for p in parts) + 1
    Rt = max(p[1] for p in parts) + 1
    nx, nt, ps = parts[0][2], parts[0][3], parts[0][4].shape[-1]
    out = np.zeros((Rx * nx, Rt * nt, ps), complex)
    for cx, ct, _, _, d in parts:
        out[cx * nx:(cx + 1) * nx, ct * nt:(ct + 1) * nt] = d
    return out


def ref_links_to_stencil(G1: np.ndarray, G2: np.ndarray, G3: np.ndarray, m0: float) -> Stencil:
    """Reference coarse-link dumps (G1[n][alf][bet][c][b], G2/G3[...][mu],
    include/level.h:349-365) to Stencil matrices with row = c*2+alf, col = b*2+bet."""
    Nx, Nt, dd = G1.shape
    d = int(round(math.sqrt(dd)))
    col = d // 2
    a = G1.reshape(Nx, Nt, 2, 2, col, col)                  # alf bet c b
    A = np.transpose(a, (0, 1, 4, 2, 5, 3)).reshape(Nx, Nt, d, d)
    B = np.zeros((2, Nx, Nt, d, d), complex)
    C = np.zeros((2, Nx, Nt, d, d), complex)
    for mu in (0, 1):
        b2 = G2.reshape(Nx, Nt, 2, 2, col, col, 2)[..., mu]
        b3 = G3.reshape(Nx, Nt, 2, 2, col, col, 2)[..., mu]
        B[mu] = np.transpose(b2, (0, 1, 4, 2, 5, 3)).reshape(Nx, Nt, d, d)
        C[mu] = np.transpose(b3, (0, 1, 4, 2, 5, 3)).reshape(Nx, Nt, d, d)
    return Stencil(A, B, C, m0)


# --------------------------------------------------------------------------------------
# Differentiation (no reference counterpart; pinned by dense finite differences)


def dL_dtheta(U: np.ndarray, y: np.ndarray, x: np.ndarray) -> np.ndarray:
    """Gradient of L = Re<eta, D^-1 b> with respect to the link phases theta_mu(n)
    (U = e^{i theta}), given x = D^-1 b and y = D^-dagger eta:
        dL/dtheta_mu(n) = -Re[ y^H (dD/dtheta_mu(n)) x ],
    where U_mu(n) enters row n as -1/2 (1 - sigma_mu) s+ U x(n+mu) and row n+mu as
    -1/2 (1 + sigma_mu) s- U^* x(n) (src/dirac_operator.cpp:24-54).  Returns (Nx, Nt, 2)."""
    Nt = x.shape[1]
    g = np.zeros(U.shape[:2] + (2,))
    for mu in (0, 1):
        s = sign_fwd(Nt)[..., 0] if mu == 0 else np.ones(U.shape[:2])
        u = U[..., mu]
        xf = fwd(x, mu)
        yf = fwd(y, mu)
        M, P = _M[mu], _P[mu]
        f = (1j * u * s)[..., None] * np.einsum('ab,xtb->xta', M, xf)
        bk = (-1j * np.conj(u) * s)[..., None] * np.einsum('ab,xtb->xta', P, x)
        d = -0.5 * (np.sum(np.conj(y) * f, axis=-1) + np.sum(np.conj(yf) * bk, axis=-1))
        g[..., mu] = -np.real(d)
    return g


def dense_D(U: np.ndarray, m0: float) -> np.ndarray:
    """Dense matrix of D_phi on a small lattice (columns = D applied to unit vectors)."""
    Nx, Nt = U.shape[:2]
    n = Nx * Nt * 2
    M = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1
        M[:, j] = D_phi(U, e.reshape(Nx, Nt, 2), m0).ravel()
    return M
