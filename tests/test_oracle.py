"""Pins the numpy oracle (oracle/ddamg_oracle.py) before anything is checked against it:
golden vectors written by the compiled reference, plus mathematical checks that a
plausible slip (wrong hop direction, lost sign, transposed link) would fail."""
import math

import numpy as np
import pytest

from oracle import ddamg_oracle as orc


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _setup_from_golden(meta, arr):
    levels = [orc.LevelSpec(tuple(l['agg']), tuple(l['sap_block']), l['nv']) for l in meta['levels']]
    U = np.exp(1j * arr['theta'])
    prm = orc.AmgParams(levels=levels, m0=meta['m0'], rng_ranks=tuple(meta['ranks']))
    amg = orc.AMG(U, prm)
    amg.setup()
    return U, amg


# ---------------------------------------------------------------- Wilson-Dirac operator
def test_dphi_golden(golden):
    meta, a = golden('g16_dirac')
    U = np.exp(1j * a['theta'])
    assert rel(orc.D_phi(U, a['rhs'], meta['m0']), a['Dphi']) < 1e-14
    assert rel(orc.D_dagger_phi(U, a['rhs'], meta['m0']), a['Ddagphi']) < 1e-14
    assert rel(orc.dirac_stencil(U, meta['m0']).apply(a['rhs']), a['Dop0']) < 1e-14


@pytest.mark.parametrize('kx,kt', [(0, 0), (1, 2), (3, 5), (7, 0)])
def test_dphi_free_field(kx, kt):
    """U = 1: D e^{ipx} chi = [m0 + sum(1 - cos p) + i sum sigma_mu sin p_mu] e^{ipx} chi,
    with p_x = 2 pi kx / Nx (periodic) and p_t = (2 kt + 1) pi / Nt (antiperiodic)."""
    Nx, Nt, m0 = 8, 12, -0.3
    px, pt = 2 * np.pi * kx / Nx, (2 * kt + 1) * np.pi / Nt
    X, T = np.meshgrid(np.arange(Nx), np.arange(Nt), indexing='ij')
    ph = np.exp(1j * (px * X + pt * T))
    sx = np.array([[0, 1], [1, 0]], complex)
    sy = np.array([[0, -1j], [1j, 0]], complex)
    Dp = (m0 + (1 - np.cos(px)) + (1 - np.cos(pt))) * np.eye(2) + 1j * (sx * np.sin(pt) + sy * np.sin(px))
    U = np.ones((Nx, Nt, 2), complex)
    for chi in (np.array([1, 0], complex), np.array([0.3 - 1j, 2.0])):
        phi = ph[..., None] * chi
        out = orc.D_phi(U, phi, m0)
        assert rel(out, ph[..., None] * (Dp @ chi)) < 1e-13


def test_dphi_gauge_covariance():
    rng = np.random.default_rng(3)
    Nx, Nt, m0 = 6, 8, -0.2
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (Nx, Nt, 2)))
    g = np.exp(1j * rng.uniform(-np.pi, np.pi, (Nx, Nt)))
    Ug = np.empty_like(U)
    Ug[..., 0] = g * U[..., 0] * np.conj(np.roll(g, -1, axis=1))
    Ug[..., 1] = g * U[..., 1] * np.conj(np.roll(g, -1, axis=0))
    phi = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))
    lhs = orc.D_phi(Ug, g[..., None] * phi, m0)
    assert rel(lhs, g[..., None] * orc.D_phi(U, phi, m0)) < 1e-13
    lhs = orc.D_dagger_phi(Ug, g[..., None] * phi, m0)
    assert rel(lhs, g[..., None] * orc.D_dagger_phi(U, phi, m0)) < 1e-13


def test_gamma5_hermiticity():
    rng = np.random.default_rng(4)
    Nx, Nt, m0 = 6, 8, -0.2
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (Nx, Nt, 2)))
    x = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))
    y = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))
    assert rel(orc.D_dagger_phi(U, x, m0), orc.sigma3(orc.D_phi(U, orc.sigma3(x), m0))) < 1e-14
    assert abs(np.vdot(y, orc.D_phi(U, x, m0)) - np.vdot(orc.D_dagger_phi(U, y, m0), x)) < 1e-12


# ---------------------------------------------------------------- hierarchy
@pytest.mark.parametrize('name', ['g16_setup2', 'g16_setup3'])
def test_setup_golden(golden, name):
    meta, a = golden(name)
    U, amg = _setup_from_golden(meta, a)
    assert rel(amg.levels[0].W, a['tvec_l0']) < 1e-12
    if 'tvec_l1' in a:
        assert rel(amg.levels[1].W, a['tvec_l1']) < 1e-12
    L = len(amg.levels) - 1
    st = orc.ref_links_to_stencil(a[f'G1_l{L}'], a[f'G2_l{L}'], a[f'G3_l{L}'], meta['m0'])
    op = amg.levels[L].op
    for mine, ref in ((op.A, st.A), (op.B, st.B), (op.C, st.C)):
        assert rel(mine, ref) < 1e-12
    lv0 = amg.levels[0]
    vc = orc.restrict(lv0.W, a['rhs'], lv0.spec.agg)
    assert rel(vc, a['Pdagg_rhs']) < 1e-12
    assert rel(amg.levels[1].op.apply(vc), a['Dop1_Pdagg_rhs']) < 1e-12
    assert rel(orc.prolong(lv0.W, vc, lv0.spec.agg), a['P_Pdagg_rhs']) < 1e-12
    b0 = orc.SapBlocks(16, 16, *lv0.spec.sap_block)
    s0, _ = orc.sap(lv0.op, b0, a['rhs'], np.zeros_like(a['rhs']), meta['sap_nu'], 1e-10)
    assert rel(s0, a['SAP0_rhs']) < 1e-12
    if 'SAP1_Pdagg_rhs' in a:
        lv1 = amg.levels[1]
        b1 = orc.SapBlocks(*lv1.op.shape, *lv1.spec.sap_block)
        s1, _ = orc.sap(lv1.op, b1, vc, np.zeros_like(vc), meta['sap_nu'], 1e-10)
        assert rel(s1, a['SAP1_Pdagg_rhs']) < 1e-12


def test_orthonormal_and_galerkin(golden):
    """P^dagger P = 1 and D_c = P^dagger D P at every level (include/tests.h:158-212,
    :382-476), by products of the transfer operators on random coarse vectors."""
    meta, a = golden('g16_setup3')
    U, amg = _setup_from_golden(meta, a)
    rng = np.random.default_rng(0)
    for l in range(len(amg.levels) - 1):
        lv, nxt = amg.levels[l], amg.levels[l + 1]
        shape = nxt.op.shape + (nxt.op.dof,)
        vc = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        Pv = orc.prolong(lv.W, vc, lv.spec.agg)
        assert rel(orc.restrict(lv.W, Pv, lv.spec.agg), vc) < 1e-13
        assert rel(orc.restrict(lv.W, lv.op.apply(Pv), lv.spec.agg), nxt.op.apply(vc)) < 1e-12


def test_sap_exact_blocks_closed_form():
    """With exact block solves SAP(nu=1) from x=0 equals K v, K = A_b (1 - D A_r) + A_r,
    A_c the sum of the inverse block operators of colour c (doc eqs. 10-11)."""
    rng = np.random.default_rng(5)
    Nx = Nt = 8
    U = np.exp(1j * rng.uniform(-1, 1, (Nx, Nt, 2)))
    op = orc.dirac_stencil(U, -0.2)
    blocks = orc.SapBlocks(Nx, Nt, 2, 2)
    v = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))
    exact = orc.SapParams(m=8, restarts=4, tol=1e-15)
    x, _ = orc.sap(op, blocks, v, np.zeros_like(v), 1, 1e-30, exact)

    def A(colour, r):
        ids = blocks.colour_ids(colour)
        rb = orc._gather_blocks(r, blocks, ids)
        f = orc.block_operator(op, blocks, ids)
        n = rb[0].size
        out = np.zeros_like(rb)
        for k in range(len(ids)):        # dense inverse of each block operator
            M = np.zeros((n, n), complex)
            for j in range(n):
                e = np.zeros(rb.shape, complex)
                e[k].flat[j] = 1
                M[:, j] = f(e)[k].ravel()
            out[k] = np.linalg.solve(M, rb[k].ravel()).reshape(rb[k].shape)
        return orc._scatter_add_blocks(np.zeros_like(r), out, blocks, ids)

    ar = A(0, v)
    ref = A(1, v - op.apply(ar)) + ar
    assert rel(x, ref) < 1e-10


def test_fgmres_dense():
    """FGMRES with the identity preconditioner against a dense solve."""
    rng = np.random.default_rng(6)
    n = 40
    M = np.eye(n) * 3 + rng.standard_normal((n, n)) * 0.3 + 1j * rng.standard_normal((n, n)) * 0.3
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    res = orc.fgmres(lambda v: (M @ v[0])[None], None, b[None], np.zeros((1, n), complex), 10, 20, 1e-12)
    assert rel(res.x[0], np.linalg.solve(M, b)) < 1e-10
    assert res.iterations[0] < 200


@pytest.mark.slow
@pytest.mark.parametrize('cycle', ['K', 'V'])
def test_solve_golden(golden, cycle):
    meta, a = golden('g64_solve')
    levels = [orc.LevelSpec(tuple(l['agg']), tuple(l['sap_block']), l['nv']) for l in meta['levels']]
    U = np.exp(1j * a['theta'])
    prm = orc.AmgParams(levels=levels, m0=meta['m0'], nu2=meta['nu2'], fgmres_m=meta['fgmres_m'],
                        rng_ranks=tuple(meta['ranks']))
    amg = orc.AMG(U, prm)
    amg.setup()
    res = amg.solve(a['rhs'], cycle)
    assert int(res.iterations[0]) == meta[f'{cycle}_iterations']
    x = res.x[0]
    rr = np.linalg.norm(a['rhs'] - orc.D_phi(U, x, meta['m0'])) / np.linalg.norm(a['rhs'])
    assert rr < 1e-10
    assert abs(rr - meta[f'{cycle}_relres']) / meta[f'{cycle}_relres'] < 1e-4
    if cycle == 'K':
        assert rel(x, a['x_K']) < 1e-12


# ---------------------------------------------------------------- RNG, gauge, I/O
def test_mt19937_uniform_known_values():
    """First draws of std::uniform_real_distribution<double>(-1,1) on std::mt19937(0),
    printed by a libstdc++ program."""
    u = orc.mt19937_uniform(0, 3)
    assert u.tolist() == [0.18568923303336526, 0.68853148851319657, 0.71589123997965975]


def test_gauge_plaquette():
    from scipy.special import iv
    theta = orc.gauge_metropolis(32, 32, 2.0, 400, seed=9)
    pl = np.mean(np.cos(orc.plaquette_angle(theta)))
    assert abs(pl - iv(1, 2.0) / iv(0, 2.0)) < 0.02


def test_records_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((4, 6, 2)) + 1j * rng.standard_normal((4, 6, 2))
    p = str(tmp_path / 'u.bin')
    orc.write_records(p, a)
    import os
    assert os.path.getsize(p) == 4 * 6 * 2 * 28
    assert np.array_equal(orc.read_records(p, 4, 6), a)


def test_gradient_formula_dense_fd():
    """dL/dtheta and dL/dm0 of L = Re<eta, D^-1 b> by central differences of dense
    solves on a 4x6 lattice (h = 1e-6, exact linear algebra)."""
    rng = np.random.default_rng(8)
    Nx, Nt, m0 = 4, 6, -0.2
    th = rng.uniform(-np.pi, np.pi, (Nx, Nt, 2))
    b = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))
    eta = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))

    def L(theta, m):
        D = orc.dense_D(np.exp(1j * theta), m)
        return np.real(np.vdot(eta.ravel(), np.linalg.solve(D, b.ravel())))

    U = np.exp(1j * th)
    D = orc.dense_D(U, m0)
    x = np.linalg.solve(D, b.ravel()).reshape(b.shape)
    y = np.linalg.solve(D.conj().T, eta.ravel()).reshape(b.shape)
    g = orc.dL_dtheta(U, y, x)
    h = 1e-6
    for (ix, it, mu) in [(0, 0, 0), (1, 5, 0), (3, 2, 1), (2, 5, 1), (0, 3, 1)]:
        tp, tm = th.copy(), th.copy()
        tp[ix, it, mu] += h
        tm[ix, it, mu] -= h
        fd = (L(tp, m0) - L(tm, m0)) / (2 * h)
        assert abs(fd - g[ix, it, mu]) < 1e-7 * max(1, abs(fd))
    fd_m = (L(th, m0 + h) - L(th, m0 - h)) / (2 * h)
    assert abs(fd_m - (-np.real(np.vdot(y, x)))) < 1e-7 * max(1, abs(fd_m))


def test_comparison_solvers_golden(golden):
    """CG on the normal equations, BiCGStab, GMRES(20) and FGMRES(20)+SAP: the oracle's
    restatements reproduce the compiled reference's iteration counts and solutions on
    the g32_methods fixture."""
    meta, a = golden('g32_methods')
    U = np.exp(1j * a['theta'])
    b = a['rhs'][None]
    m0, tol = meta['m0'], meta['tol']
    x, its = orc.cg_normal(U, b, m0, meta['max_iter'], tol)
    assert int(its[0]) == meta['cg_iterations'] and rel(x[0], a['x_cg']) < 1e-10
    x, its = orc.bicgstab(U, b, np.zeros_like(b), m0, meta['max_iter'], tol)
    assert int(its[0]) == meta['bicgstab_iterations'] and rel(x[0], a['x_bicgstab']) < 1e-10
    func = lambda v: np.stack([orc.D_phi(U, vi, m0) for vi in v])
    res = orc.fgmres(func, None, b, np.zeros_like(b), meta['gmres_m'], meta['gmres_restarts'], tol)
    assert int(res.iterations[0]) == meta['gmres_iterations'] and rel(res.x[0], a['x_gmres']) < 1e-10
    op = orc.dirac_stencil(U, m0)
    blocks = orc.SapBlocks(meta['N'], meta['N'], *meta['sap_block'])
    pre = lambda v, z: np.stack([orc.sap(op, blocks, vi, np.zeros_like(vi), 1, 1e-10)[0] for vi in v])
    res = orc.fgmres(func, pre, b, np.zeros_like(b), meta['gmres_m'], meta['gmres_restarts'], tol)
    assert int(res.iterations[0]) == meta['fgmres_sap_iterations'] and rel(res.x[0], a['x_fgmres_sap']) < 1e-10


def test_cg_bicgstab_dense():
    """Against a dense solve on 4x4: both reach the solution; CG on the (Hermitian
    positive definite) normal equations stops within the dimension, 32, up to rounding."""
    rng = np.random.default_rng(11)
    N, m0 = 4, -0.3
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)))
    b = (rng.standard_normal((2, N, N, 2)) + 1j * rng.standard_normal((2, N, N, 2)))
    D = orc.dense_D(U, m0)
    xd = np.stack([np.linalg.solve(D, bi.ravel()).reshape(N, N, 2) for bi in b])
    x, its = orc.cg_normal(U, b, m0, 500, 1e-12)
    assert rel(x, xd) < 1e-9 and (its > 0).all() and (its <= 48).all()
    x, its = orc.bicgstab(U, b, np.zeros_like(b), m0, 500, 1e-12)
    assert rel(x, xd) < 1e-9 and (its < 500).all()


def _low_mode_defect(U, m0, W, agg, k=6):
    """Mean of ||(1 - P P^dagger) v|| over the k right singular vectors of D_W with the
    smallest singular values (unit norm): how well the prolongator's range holds the
    near-kernel."""
    D = orc.dense_D(U, m0)
    _, _, vh = np.linalg.svd(D)
    Nx, Nt = U.shape[:2]
    out = []
    for v in vh[-k:].conj():
        vv = v.reshape(Nx, Nt, 2)
        r = vv - orc.prolong(W, orc.restrict(W, vv, agg), agg)
        out.append(np.linalg.norm(r))
    return float(np.mean(out))


def test_bootstrap_properties(golden):
    """Bootstrap adaptivity (AMG.bootstrap, documented but not implemented by the
    reference): zero passes change nothing; after two passes every prolongator is
    orthonormal and every coarse operator is the Galerkin product of the new ones (dense
    brute force), and the prolongator holds the low modes of D_W better (inverse-iteration
    effect: the part of the six lowest right singular vectors outside its range shrinks by
    more than 20 %)."""
    meta, a = golden('g16_setup3')
    U = np.exp(1j * a['theta'])
    # aggregates of 4x4 sites with 4 test vectors: P has rank 4 of 16 per aggregate and spin
    levels = [orc.LevelSpec((4, 4), (4, 4), 4), orc.LevelSpec((2, 2), (2, 2), 4)]
    amg = orc.AMG(U, orc.AmgParams(levels=levels, m0=meta['m0'], rng_ranks=tuple(meta['ranks'])))
    amg.setup()
    W0 = [lv.W.copy() for lv in amg.levels[:-1]]
    A0 = amg.levels[-1].op.A.copy()
    amg.bootstrap(0)
    assert all(np.array_equal(lv.W, w) for lv, w in zip(amg.levels[:-1], W0))
    assert np.array_equal(amg.levels[-1].op.A, A0)
    lv0 = amg.levels[0]
    before = _low_mode_defect(U, meta['m0'], lv0.W, lv0.spec.agg)
    amg.bootstrap(2)
    after = _low_mode_defect(U, meta['m0'], lv0.W, lv0.spec.agg)
    print(f'low-mode defect of P: {before:.3e} -> {after:.3e}')
    assert after < 0.8 * before
    assert rel(amg.levels[0].W, W0[0]) > 1e-3
    rng = np.random.default_rng(4)
    for l in range(len(amg.levels) - 1):
        lv, nxt = amg.levels[l], amg.levels[l + 1]
        shape = nxt.op.shape + (nxt.op.dof,)
        vc = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        Pv = orc.prolong(lv.W, vc, lv.spec.agg)
        assert rel(orc.restrict(lv.W, Pv, lv.spec.agg), vc) < 1e-13
        assert rel(orc.restrict(lv.W, lv.op.apply(Pv), lv.spec.agg), nxt.op.apply(vc)) < 1e-12
