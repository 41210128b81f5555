"""GPU tests of the coarsest GMRES's polynomial preconditioner (dd_params.coarse_poly).

With coarse_poly = d the coarsest GMRES runs on D_c p(D_c), p(z) = (1 - prod_i (1 -
z/theta_i)) / z, whose roots are the harmonic Ritz values of d Arnoldi steps on D_c from
the library's fixed start vector, and returns x0 + p(D_c) Q y.  Checked here:

- the roots equal the harmonic Ritz values found from their definition on the downloaded
  coarsest operator and the same start vector (tests/poly_ref.py; 1e-9 relative in the
  FP64 hierarchy, 1e-3 in FP32, whose device applies the operator in FP32), in Leja order;
- run to a tight tolerance, the coarsest solve of a 17-system batch equals the dense solve
  of the coarsest operator (FP64 1e-9, FP32 1e-4, the binary16 basis included) on both
  engines (persistent kernel and launch-per-step), and every system equals its single
  solve; at the reference tolerance 0.1 numpy's residual ||eta - D x|| is below
  0.1 ||eta|| for every system, with fewer Arnoldi steps than without the preconditioner;
- 3-level K-cycle solves (FP64 and mixed, 4 systems) reach 1e-10 in the FP64 residual and
  agree with the solves without the preconditioner to 1e-8;
- at 512^2 on the fixture's field, the mixed 32-system batch (K-cycle with coarse_poly = 8,
  V-cycle with coarse_poly = 12) reaches 1e-10 for every system, and system 0 equals the
  FP64 solve to 1e-8;
- at 512^2 (the shape whose polynomial sweeps run their in-slice hops before the
  neighbour wait) tight-tolerance coarsest solves reach the same FP64 residual and the
  same distance to a sparse LU solve of the downloaded coarsest operator as the plain
  kernel's (within 3x).
"""
import hashlib

import numpy as np
import pytest

from oracle import ddamg_oracle as orc
from poly_ref import arnoldi, assert_leja_order, fill_hash, harmonic_ritz_from_images, match_distance

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')

POLY_SEED = 4242   # the library's start vector of the roots' Arnoldi run


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.ravel(b))


def small_field(seed=5, N=64):
    rng = np.random.default_rng(seed)
    return np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)) * 0.5)


BASE = dict(agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[8, 8], m0=-0.2, rng_ranks=(2, 2))


@pytest.mark.parametrize('precision', [0, 1])
def test_roots_match_definition(precision):
    N, d = 64, 8
    with ddamg.Context(N, N) as ctx:
        ctx.set_gauge(small_field())
        ctx.setup(ddamg.params(3, precision=precision, max_rhs=4, coarse_poly=d, **BASE))
        th = ctx.coarse_poly_roots()
        assert len(th) == d
        shape = ctx.level_shape(2)
        A, B, Cm = ctx.links(2, precision=precision)
    op = orc.Stencil(A, B, Cm, BASE['m0'])
    n = int(np.prod(shape))
    V, W, _ = arnoldi(lambda x: op.apply(x.reshape(shape)).ravel(), fill_hash(n, POLY_SEED), d,
                      round32=precision == 1)
    ref = harmonic_ritz_from_images(V, W)
    dist = max(match_distance(th, ref), match_distance(ref, th))
    print(f'precision {precision}: roots vs definition {dist:.2e}, |theta| in [{np.abs(th).min():.3f}, '
          f'{np.abs(th).max():.3f}]')
    assert dist < (1e-9 if precision == 0 else 1e-3)
    assert_leja_order(th)


@pytest.mark.parametrize('precision,basis', [(0, 0), (1, 0), (1, 1)])
def test_coarsest_solve_with_poly(precision, basis):
    N, d = 64, 8
    U = small_field()
    rng = np.random.default_rng(17)
    base = dict(BASE, precision=precision, max_rhs=17, coarse_basis=basis)
    tight = 1e-12 if precision == 0 else 1e-6
    bound = 1e-9 if precision == 0 else 1e-4
    v = None
    for ortho in (2, 1):   # persistent kernel, launch-per-step engine
        with ddamg.Context(N, N) as ctx:
            ctx.set_gauge(U)
            ctx.setup(ddamg.params(3, coarse_tol=tight, ortho_coarse=ortho, coarse_poly=d, **base))
            assert ctx.coarsest_stats()['persistent'] == (1 if ortho == 2 else 0)
            shape = ctx.level_shape(2)
            if v is None:
                v = rng.standard_normal((17,) + shape) + 1j * rng.standard_normal((17,) + shape)
                A, B, Cm = ctx.links(2, precision=precision)
                op = orc.Stencil(A, B, Cm, BASE['m0'])
                nn = int(np.prod(shape))
                D = np.stack([op.apply(e.reshape(shape)).ravel() for e in np.eye(nn)], axis=1)
            x = ctx.cycle(2, v, np.zeros_like(v), kcycle=True, precision=precision)
            for r in (0, 7, 16):
                xd = np.linalg.solve(D, v[r].ravel()).reshape(shape)
                err = rel(x[r], xd)
                print(f'ortho {ortho} precision {precision} basis {basis} system {r}: error vs dense solve {err:.2e}')
                assert err < bound
                assert rel(ctx.cycle(2, v[r], np.zeros_like(v[r]), kcycle=True, precision=precision), x[r]) < 1e-12
    steps = {}
    for poly in (0, d):
        with ddamg.Context(N, N) as ctx:
            ctx.set_gauge(U)
            ctx.setup(ddamg.params(3, ortho_coarse=2, coarse_poly=poly, **base))
            x = ctx.cycle(2, v, np.zeros_like(v), kcycle=True, precision=precision)
            steps[poly] = ctx.coarsest_stats()['steps']
            for r in range(17):
                res = v[r] - op.apply(x[r])
                assert np.linalg.norm(res) < 0.1 * np.linalg.norm(v[r]) * (1 + 1e-5)
    print(f'Arnoldi steps to 0.1: plain {steps[0]:.0f}, degree {d} {steps[d]:.0f}')
    assert steps[d] < steps[0]


@pytest.mark.parametrize('precision', [0, 1])
def test_kcycle_solve_with_poly(precision):
    N = 64
    U = small_field(9)
    rng = np.random.default_rng(23)
    B = rng.standard_normal((4, N, N, 2)) + 1j * rng.standard_normal((4, N, N, 2))
    out = {}
    for poly in (0, 8):
        with ddamg.Context(N, N) as ctx:
            ctx.set_gauge(U)
            ctx.setup(ddamg.params(3, precision=precision, max_rhs=4, coarse_poly=poly, **BASE))
            x, its, info = ctx.solve(B)
            Dx = ctx.apply_D(x, BASE['m0'])
            rr = np.linalg.norm((B - Dx).reshape(4, -1), axis=1) / np.linalg.norm(B.reshape(4, -1), axis=1)
            ro = np.linalg.norm(B[0] - orc.D_phi(U, x[0], BASE['m0'])) / np.linalg.norm(B[0])
            print(f'precision {precision} coarse_poly {poly}: iterations {list(its)}, max relres {rr.max():.2e} '
                  f'(oracle, system 0: {ro:.2e})')
            assert rr.max() < 1e-10 and info.relres < 1e-10 and ro < 1e-10
            out[poly] = x
    d = rel(out[8], out[0])
    print(f'  solutions with and without the preconditioner differ by {d:.2e}')
    assert d < 1e-8


@pytest.mark.parametrize('cycle,poly', [(1, 8), (0, 12)])
def test_512_mixed_batch_with_poly(golden, cycle, poly):
    """K-cycle with degree 8, and the V-cycle with degree 12 (the benchmark's solve)."""
    meta, a = golden('g512_solve')
    N, m0 = meta['N'], meta['m0']
    lv = meta['levels']
    prm = dict(agg=[tuple(l['agg']) for l in lv], sap_block=[tuple(l['sap_block']) for l in lv],
               nv=[l['nv'] for l in lv], m0=m0, nu2=meta['nu2'], fgmres_m=meta['fgmres_m'], cycle=1,
               rng_ranks=tuple(meta['ranks']))
    rng = np.random.default_rng(meta['rhs_seed'])
    b0 = rng.standard_normal((N, N, 2)) + 1j * rng.standard_normal((N, N, 2))
    rng2 = np.random.default_rng(9001)
    B = np.concatenate([b0[None], rng2.standard_normal((31, N, N, 2)) + 1j * rng2.standard_normal((31, N, N, 2))])
    with ddamg.Context(N, N) as ctx:
        th = ctx.generate_gauge(meta['beta'], meta['sweeps'], meta['seed'], meta['step'])
        assert hashlib.sha256(np.ascontiguousarray(th).tobytes()).hexdigest() == meta['theta_sha256']
        ctx.setup(ddamg.params(len(lv) + 1, precision=0, max_rhs=1, **prm))
        x64, its64, _ = ctx.solve(b0)
        prm['cycle'] = cycle
        ctx.setup(ddamg.params(len(lv) + 1, precision=1, max_rhs=32, coarse_poly=poly, **prm))
        assert ctx.coarsest_stats()['persistent'] == 1
        xm, itm, info = ctx.solve(B)
        Dx = ctx.apply_D(xm, m0)
        rr = np.linalg.norm((B - Dx).reshape(32, -1), axis=1) / np.linalg.norm(B.reshape(32, -1), axis=1)
        st = ctx.coarsest_stats()
        err = rel(xm[0], x64)
        print(f'512^2 mixed, 32 RHS, {"K" if cycle else "V"}-cycle, coarse_poly {poly}: iterations '
              f'{itm.min()}..{itm.max()}, max relres {rr.max():.2e}, '
              f'system 0 vs FP64 {err:.2e}, {st["steps"] / max(st["solves"], 1):.1f} Arnoldi steps per coarsest solve')
        assert rr.max() < 1e-10 and info.relres < 1e-10
        assert err < 1e-8


def coarse_sparse(A, B, Cm, m0):
    """The coarse stencil of the oracle (orc.Stencil.apply) as a scipy sparse matrix over
    [x][t][dof]: (m0 + 2 - A) on the diagonal blocks, -B_mu / -C_mu towards x+mu / x-mu
    (mu = 0: t with the antiperiodic sign on the last/first time slice, mu = 1: x)."""
    import scipy.sparse as sp
    Nx, Nt, d = A.shape[:3]
    site = np.arange(Nx * Nt).reshape(Nx, Nt)
    xs, ts = np.meshgrid(np.arange(Nx), np.arange(Nt), indexing='ij')
    rows, cols, vals = [], [], []
    ii, jj = np.meshgrid(np.arange(d), np.arange(d), indexing='ij')

    def add(target, blocks, sign):
        r = (site[..., None, None] * d + ii).ravel()
        c = (target[..., None, None] * d + jj).ravel()
        rows.append(r)
        cols.append(c)
        vals.append((sign[..., None, None] * blocks).ravel())
    one = np.ones((Nx, Nt))
    add(site, (m0 + 2) * np.eye(d) - A, one)
    sf = np.where(ts == Nt - 1, -1.0, 1.0)
    sb = np.where(ts == 0, -1.0, 1.0)
    add(site[:, (np.arange(Nt) + 1) % Nt], B[0], -sf)
    add(site[:, (np.arange(Nt) - 1) % Nt], Cm[0], -sb)
    add(site[(np.arange(Nx) + 1) % Nx, :], B[1], -one)
    add(site[(np.arange(Nx) - 1) % Nx, :], Cm[1], -one)
    n = Nx * Nt * d
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def test_coarsest_poly_512_against_sparse_solve(golden):
    """The persistent kernel's polynomial stages at the shape of the benchmark (32^2
    coarsest sites, 4-site slices: each sweep's in-slice hops run before the neighbour
    wait), FP32 hierarchy with the binary16 basis, run to the coarsest tolerance 1e-6 on 3
    systems: the FP64 residual of each solution with the downloaded operator (a sparse
    matrix that applies as the oracle's stencil does) is below 2e-5 and within 3x of the
    plain kernel's at the same tolerance, and so is its distance to a sparse LU solve.
    (The FP32-stored solutions of both kernels sit near 6e-6 in FP64 residual and 2e-4 from
    the LU solve: the floor that FP32 storage and this operator's conditioning set.)"""
    spla = pytest.importorskip('scipy.sparse.linalg')
    meta, _ = golden('g512_solve')
    N, m0 = meta['N'], meta['m0']
    lv = meta['levels']
    rng = np.random.default_rng(512)
    xs = {}
    for poly in (12, 0):
        with ddamg.Context(N, N) as ctx:
            ctx.generate_gauge(meta['beta'], meta['sweeps'], meta['seed'], meta['step'])
            ctx.setup(ddamg.params(len(lv) + 1, agg=[tuple(l['agg']) for l in lv],
                                   sap_block=[tuple(l['sap_block']) for l in lv], nv=[l['nv'] for l in lv], m0=m0,
                                   precision=1, max_rhs=3, rng_ranks=tuple(meta['ranks']), coarse_poly=poly,
                                   coarse_tol=1e-6, ortho_coarse=2))
            assert ctx.coarsest_stats()['persistent'] == 1
            shape = ctx.level_shape(2)
            if poly:
                v = rng.standard_normal((3,) + shape) + 1j * rng.standard_normal((3,) + shape)
                A, B, Cm = ctx.links(2, precision=1)
            xs[poly] = ctx.cycle(2, v, np.zeros_like(v), kcycle=True, precision=1)
    M = coarse_sparse(A, B, Cm, m0)
    op = orc.Stencil(A, B, Cm, m0)
    assert rel(M @ v[0].ravel(), op.apply(v[0]).ravel()) < 1e-13
    lu = spla.splu(M)
    for r in range(3):
        xd = lu.solve(v[r].ravel()).reshape(shape)
        res = {p: np.linalg.norm(v[r].ravel() - M @ xs[p][r].ravel()) / np.linalg.norm(v[r]) for p in xs}
        err = {p: rel(xs[p][r], xd) for p in xs}
        print(f'512^2 coarsest, system {r}: residual poly {res[12]:.2e} plain {res[0]:.2e}; '
              f'error vs sparse solve poly {err[12]:.2e} plain {err[0]:.2e}')
        assert res[12] < 2e-5 and res[12] < 3 * res[0]
        assert err[12] < 3 * err[0]
