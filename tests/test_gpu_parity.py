"""GPU parity of the HIP path (through the C ABI) against the reference's golden vectors
and the numpy oracle.  Tolerances: FP64 operators 1e-14 relative (the north-star bar for
D_W), hierarchy objects 1e-11 (they pass through 4 SAP iterations of inexact block
GMRES), solves 1e-10 relative residual with the FP64 D_W and the reference's iteration
counts."""
import numpy as np
import pytest

from oracle import ddamg_oracle as orc

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _ctx_from_golden(meta, a):
    N = meta['N']
    ctx = ddamg.Context(N, N)
    ctx.set_gauge(np.exp(1j * a['theta']))
    return ctx


def _prm(meta, **kw):
    lv = meta['levels']
    return ddamg.params(len(lv) + 1, agg=[tuple(l['agg']) for l in lv],
                        sap_block=[tuple(l['sap_block']) for l in lv], nv=[l['nv'] for l in lv],
                        m0=meta['m0'], rng_ranks=tuple(meta['ranks']), **kw)


# ------------------------------------------------------------------------ operator
def test_dirac_golden(golden):
    meta, a = golden('g16_dirac')
    with _ctx_from_golden(meta, a) as ctx:
        assert rel(ctx.apply_D(a['rhs'], meta['m0']), a['Dphi']) < 1e-14
        assert rel(ctx.apply_D(a['rhs'], meta['m0'], dagger=True), a['Ddagphi']) < 1e-14
        assert rel(ctx.apply_D(a['rhs'], meta['m0'], precision=1), a['Dphi']) < 1e-6


def test_dirac_oracle_nonsquare_batch():
    rng = np.random.default_rng(11)
    Nx, Nt, m0 = 24, 40, -0.37
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (Nx, Nt, 2)))
    v = rng.standard_normal((3, Nx, Nt, 2)) + 1j * rng.standard_normal((3, Nx, Nt, 2))
    with ddamg.Context(Nx, Nt) as ctx:
        ctx.set_gauge(U)
        out = ctx.apply_D(v, m0)
        outd = ctx.apply_D(v, m0, dagger=True)
    for r in range(3):
        assert rel(out[r], orc.D_phi(U, v[r], m0)) < 1e-14
        assert rel(outd[r], orc.D_dagger_phi(U, v[r], m0)) < 1e-14


# ------------------------------------------------------------------------ hierarchy
@pytest.mark.parametrize('name', ['g16_setup2', 'g16_setup3'])
def test_setup_golden(golden, name):
    meta, a = golden(name)
    with _ctx_from_golden(meta, a) as ctx:
        ctx.setup(_prm(meta))
        assert rel(ctx.test_vectors(0), a['tvec_l0']) < 1e-11
        if 'tvec_l1' in a:
            assert rel(ctx.test_vectors(1), a['tvec_l1']) < 1e-11
        L = len(meta['levels'])
        st = orc.ref_links_to_stencil(a[f'G1_l{L}'], a[f'G2_l{L}'], a[f'G3_l{L}'], meta['m0'])
        A, B, Cm = ctx.links(L)
        assert rel(A, st.A) < 1e-11 and rel(B, st.B) < 1e-11 and rel(Cm, st.C) < 1e-11
        vc = ctx.restrict(0, a['rhs'])
        assert rel(vc, a['Pdagg_rhs']) < 1e-11
        assert rel(ctx.level_apply(1, vc), a['Dop1_Pdagg_rhs']) < 1e-11
        assert rel(ctx.prolong(0, vc), a['P_Pdagg_rhs']) < 1e-11
        assert rel(ctx.level_apply(0, a['rhs']), a['Dop0'] if 'Dop0' in a else orc.D_phi(
            np.exp(1j * a['theta']), a['rhs'], meta['m0'])) < 1e-14
        s0 = ctx.sap(0, a['rhs'], np.zeros_like(a['rhs']), meta['sap_nu'])
        assert rel(s0, a['SAP0_rhs']) < 1e-11
        if 'SAP1_Pdagg_rhs' in a:
            s1 = ctx.sap(1, vc, np.zeros_like(vc), meta['sap_nu'])
            assert rel(s1, a['SAP1_Pdagg_rhs']) < 1e-11


def test_galerkin_and_orthonormality(golden):
    """P^dagger P = 1 and P^dagger D_l P = D_{l+1} on the device hierarchy."""
    meta, a = golden('g16_setup3')
    rng = np.random.default_rng(2)
    with _ctx_from_golden(meta, a) as ctx:
        ctx.setup(_prm(meta))
        for l in range(len(meta['levels'])):
            shape = ctx.level_shape(l + 1)
            vc = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            Pv = ctx.prolong(l, vc)
            assert rel(ctx.restrict(l, Pv), vc) < 1e-13
            assert rel(ctx.restrict(l, ctx.level_apply(l, Pv)), ctx.level_apply(l + 1, vc)) < 1e-12


# ------------------------------------------------------------------------ solves
@pytest.mark.parametrize('cycle', ['K', 'V'])
def test_solve_golden(golden, cycle):
    meta, a = golden('g64_solve')
    with _ctx_from_golden(meta, a) as ctx:
        ctx.setup(_prm(meta, nu2=meta['nu2'], fgmres_m=meta['fgmres_m'], cycle=1 if cycle == 'K' else 0))
        x, its, info = ctx.solve(a['rhs'])
        U = np.exp(1j * a['theta'])
        rr = np.linalg.norm(a['rhs'] - orc.D_phi(U, x, meta['m0'])) / np.linalg.norm(a['rhs'])
        assert rr < 1e-10
        assert int(its[0]) == meta[f'{cycle}_iterations']
        assert info.relres < 1e-10
        if cycle == 'K':
            assert rel(x, a['x_K']) < 1e-9


def test_solve_mixed_precision(golden):
    meta, a = golden('g64_solve')
    with _ctx_from_golden(meta, a) as ctx:
        ctx.setup(_prm(meta, nu2=meta['nu2'], fgmres_m=meta['fgmres_m'], cycle=1, precision=1))
        x, its, info = ctx.solve(a['rhs'])
        U = np.exp(1j * a['theta'])
        rr = np.linalg.norm(a['rhs'] - orc.D_phi(U, x, meta['m0'])) / np.linalg.norm(a['rhs'])
        assert rr < 1e-10
        # error against the FP64 solve of the same system (the reference's x)
        assert rel(x, a['x_K']) < 1e-8
        assert its[0] <= meta['K_iterations'] + 3


def test_batched_equals_single(golden):
    meta, a = golden('g64_solve')
    rng = np.random.default_rng(3)
    B = np.stack([a['rhs']] + [rng.standard_normal(a['rhs'].shape) + 1j * rng.standard_normal(a['rhs'].shape)
                               for _ in range(3)])
    with _ctx_from_golden(meta, a) as ctx:
        ctx.setup(_prm(meta, nu2=meta['nu2'], fgmres_m=meta['fgmres_m'], cycle=1, max_rhs=4))
        xb, itb, _ = ctx.solve(B)
        for r in range(4):
            x1, it1, _ = ctx.solve(B[r])
            assert int(it1[0]) == int(itb[r])
            assert rel(xb[r], x1) < 1e-12
        assert int(itb[0]) == meta['K_iterations']


def test_adjoint_solve(golden):
    meta, a = golden('g64_solve')
    U = np.exp(1j * a['theta'])
    with _ctx_from_golden(meta, a) as ctx:
        ctx.setup(_prm(meta, nu2=meta['nu2'], fgmres_m=meta['fgmres_m'], cycle=1))
        y, its, info = ctx.solve(a['rhs'], adjoint=True)
        rr = np.linalg.norm(a['rhs'] - orc.D_dagger_phi(U, y, meta['m0'])) / np.linalg.norm(a['rhs'])
        assert rr < 1e-10


def test_gauge_generator_plaquette():
    from scipy.special import iv
    with ddamg.Context(64, 64) as ctx:
        th = ctx.generate_gauge(2.0, 500, seed=5)
        assert np.allclose(np.exp(1j * th), ctx.gauge())
    pl = np.mean(np.cos(orc.plaquette_angle(th)))
    assert abs(pl - iv(1, 2.0) / iv(0, 2.0)) < 0.01


@pytest.mark.parametrize('precision', [0, 1])
def test_coarse_sap_batched(precision):
    """Coarse-level SAP on batches that fill 4- and 8-system work-groups partially equals
    per-system SAP, and (FP64) the oracle's SAP on the device's own level-1 operator."""
    rng = np.random.default_rng(21)
    N = 32
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)) * 0.4)
    prm = ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[8, 8], m0=-0.2,
                       precision=precision, max_rhs=9, rng_ranks=(2, 2))
    with ddamg.Context(N, N) as ctx:
        ctx.set_gauge(U)
        ctx.setup(prm)
        shape = ctx.level_shape(1)
        v = rng.standard_normal((9,) + shape) + 1j * rng.standard_normal((9,) + shape)
        x0 = np.zeros_like(v)
        tol_b = 1e-12 if precision == 0 else 1e-5
        for nr in (1, 3, 6, 9):
            xb = ctx.sap(1, v[:nr], x0[:nr], 2, precision=precision)
            for r in range(nr):
                x1 = ctx.sap(1, v[r], x0[r], 2, precision=precision)
                assert rel(xb[r], x1) < tol_b
        if precision == 0:
            A, B, Cm = ctx.links(1)
            op = orc.Stencil(A, B, Cm, -0.2)
            blocks = orc.SapBlocks(shape[0], shape[1], 2, 2)
            for r in range(3):
                ref, _ = orc.sap(op, blocks, v[r], x0[r], 2, 1e-10)
                assert rel(ctx.sap(1, v[r], x0[r], 2), ref) < 1e-11


def _half(a):
    """Each real and imaginary part rounded to IEEE binary16 (nearest, ties to even)."""
    return a.real.astype(np.float16).astype(np.float64) + 1j * a.imag.astype(np.float16).astype(np.float64)


@pytest.mark.parametrize('blk', [(4, 4), (2, 2)])
def test_coarse_sap_matrix_cores(blk):
    """FP32 coarse SAP on the matrix cores (variant 32, 16 systems per work-group; variant
    33, the same kernel on the binary16 link copy).
    (a) With a tolerance FP32 cannot reach, block GMRES runs all 5x5 iterations, so the
    control flow is fixed; against the oracle's FP64 SAP on the same (FP32) level-1
    operator the matrix-core result must be as accurate as the FP32 vector kernel
    (variant 2): error at most twice the vector kernel's (or 1e-5) and below 1e-3.
    Variant 33 reads a copy of the links rounded to binary16 (checked bit for bit against
    numpy's rounding of the FP32 links); it meets the same bound against the oracle's SAP
    whose block solves use the rounded links (the residuals keep the FP32 ones), and stays
    within 1e-2 of the FP32-link result.
    (b) Batches of 1, 3 and 17 systems (two 16-system tiles) give each system the same
    result as alone: the product columns and reductions of a system do not depend on
    the others."""
    rng = np.random.default_rng(33)
    N = 64
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)) * 0.4)
    out, ref, ref16 = {}, None, None
    v = x0 = None
    for variant in (32, 33, 2):
        prm = ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), blk], nv=[12, 8], m0=-0.2,
                           precision=1, max_rhs=17, rng_ranks=(2, 2), sap_variant=variant,
                           coarse_sap_links=1)
        with ddamg.Context(N, N) as ctx:
            ctx.set_gauge(U)
            ctx.setup(prm)
            assert ctx.tuning(1)[0] == variant
            shape = ctx.level_shape(1)
            assert shape[2] == 24
            if v is None:
                v = rng.standard_normal((17,) + shape) + 1j * rng.standard_normal((17,) + shape)
                x0 = np.zeros_like(v)
                A, B, Cm = ctx.links(1, precision=1)
                blocks = orc.SapBlocks(shape[0], shape[1], *blk)
                op = orc.Stencil(A, B, Cm, -0.2)
                ref = np.stack([orc.sap(op, blocks, v[r], x0[r], 2, 1e-12)[0] for r in range(3)])
                op16 = orc.Stencil(_half(A), _half(B), _half(Cm), -0.2)
                ref16 = np.stack([orc.sap(op, blocks, v[r], x0[r], 2, 1e-12, block_op=op16)[0] for r in range(3)])
            out[variant] = ctx.sap(1, v[:3], x0[:3], 2, tol=1e-12, precision=1)
            if variant == 33:   # the device's binary16 copy: the FP32 links rounded to nearest
                A16, B16, C16 = ctx.links(1, precision=2)
                A32, B32, C32 = ctx.links(1, precision=1)
                assert np.array_equal(A16, _half(A32)) and np.array_equal(B16, _half(B32))
                assert np.array_equal(C16, _half(C32))
            if variant in (32, 33):
                xb = ctx.sap(1, v, x0, 2, precision=1)
                for r in (0, 5, 16):
                    assert rel(xb[r], ctx.sap(1, v[r], x0[r], 2, precision=1)) < 1e-6
    e32, e2 = rel(out[32], ref), rel(out[2], ref)
    e33, d33 = rel(out[33], ref16), rel(out[33], out[32])
    print(f'blocks {blk}: FP32 error vs FP64 oracle: matrix cores {e32:.2e}, vector kernel {e2:.2e}; '
          f'binary16 links vs the oracle on rounded links {e33:.2e}, vs FP32 links {d33:.2e}')
    assert e32 < max(2 * e2, 1e-5) and e32 < 1e-3
    assert e33 < max(2 * e2, 1e-5) and e33 < 1e-3
    assert d33 < 1e-2


def test_bootstrap_matches_oracle(golden):
    """Bootstrap adaptivity (dd_params.bootstrap = 2 passes; documented by the reference
    but not implemented there) against the oracle's AMG.bootstrap on the same field and
    raw test vectors: test vectors of both levels and the coarsest operator agree (FP64,
    launch-per-step coarsest GMRES with modified Gram-Schmidt like the oracle; batches of
    4 test vectors per cycle).  Then the FP64 solve on the bootstrapped hierarchy meets
    1e-10, and the FP32 hierarchy bootstraps and solves too."""
    meta, a = golden('g16_setup3')
    U = np.exp(1j * a['theta'])
    spec = [((4, 4), (4, 4), 4), ((2, 2), (2, 2), 4)]
    amg = orc.AMG(U, orc.AmgParams(levels=[orc.LevelSpec(*l) for l in spec], m0=meta['m0'],
                                   rng_ranks=tuple(meta['ranks'])))
    amg.setup()
    amg.bootstrap(2)
    kw = dict(agg=[l[0] for l in spec], sap_block=[l[1] for l in spec], nv=[l[2] for l in spec], m0=meta['m0'],
              rng_ranks=tuple(meta['ranks']), ortho_coarse=0, max_rhs=4, bootstrap=2)
    with ddamg.Context(16, 16) as ctx:
        ctx.set_gauge(U)
        ctx.setup(ddamg.params(3, **kw))
        for l in (0, 1):
            e = rel(ctx.test_vectors(l), amg.levels[l].W)
            print(f'level {l} test vectors vs oracle: {e:.2e}')
            assert e < 1e-8
        A, B, Cm = ctx.links(2)
        op = amg.levels[2].op
        assert rel(A, op.A) < 1e-8 and rel(B, op.B) < 1e-8 and rel(Cm, op.C) < 1e-8
        x, its, info = ctx.solve(a['rhs'])
        rr = np.linalg.norm(a['rhs'] - orc.D_phi(U, x, meta['m0'])) / np.linalg.norm(a['rhs'])
        assert rr < 1e-10
        ctx.setup(ddamg.params(3, precision=1, **kw))
        x32, its32, _ = ctx.solve(a['rhs'])
        rr32 = np.linalg.norm(a['rhs'] - orc.D_phi(U, x32, meta['m0'])) / np.linalg.norm(a['rhs'])
        print(f'bootstrapped solves: FP64 {its[0]} iterations, mixed {its32[0]} iterations')
        assert rr32 < 1e-10 and its32[0] <= its[0] + 3


@pytest.mark.parametrize('precision,basis', [(0, 0), (1, 0), (1, 1)])
def test_coarsest_persistent(precision, basis):
    """Coarsest-level GMRES in one persistent kernel (CGS with delayed
    re-orthogonalisation, ortho_coarse = 2), basis stored in the working precision
    (basis 0) or in binary16 (basis 1, FP32 hierarchy): (a) run to a tight tolerance it
    equals the dense solve of the device's coarsest operator (FP64 1e-9, FP32 1e-4
    relative, also with the binary16 basis: the restarts' true-residual test carries the
    solution past the stored basis' rounding); (b) a 17-system batch (two 16-system tiles) gives every
    system the result it gets alone (systems share no arithmetic); (c) at the reference
    tolerance 0.1 every system meets ||eta - D x|| < 0.1 ||eta||, as the launch-per-step
    engine's solution does."""
    rng = np.random.default_rng(5)
    N = 64
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)) * 0.5)
    base = dict(agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[8, 8], m0=-0.2, precision=precision,
                max_rhs=17, rng_ranks=(2, 2), coarse_basis=basis)
    tight = 1e-12 if precision == 0 else 1e-6
    bound = 1e-9 if precision == 0 else 1e-4
    with ddamg.Context(N, N) as ctx:
        ctx.set_gauge(U)
        ctx.setup(ddamg.params(3, coarse_tol=tight, ortho_coarse=2, **base))
        assert ctx.coarsest_stats()['persistent'] == 1
        shape = ctx.level_shape(2)
        v = rng.standard_normal((17,) + shape) + 1j * rng.standard_normal((17,) + shape)
        x = ctx.cycle(2, v, np.zeros_like(v), kcycle=True, precision=precision)
        A, B, Cm = ctx.links(2, precision=precision)
        op = orc.Stencil(A, B, Cm, -0.2)
        nn = int(np.prod(shape))
        D = np.stack([op.apply(e.reshape(shape)).ravel() for e in np.eye(nn)], axis=1)
        for r in (0, 7, 16):
            xd = np.linalg.solve(D, v[r].ravel()).reshape(shape)
            err = rel(x[r], xd)
            print(f'precision {precision} basis {basis} system {r}: error vs dense solve {err:.2e}')
            assert err < bound
            assert rel(ctx.cycle(2, v[r], np.zeros_like(v[r]), kcycle=True, precision=precision), x[r]) < 1e-12
    for ortho in (2, 1):
        with ddamg.Context(N, N) as ctx:
            ctx.set_gauge(U)
            ctx.setup(ddamg.params(3, ortho_coarse=ortho, **base))
            x = ctx.cycle(2, v, np.zeros_like(v), kcycle=True, precision=precision)
            for r in range(17):
                res = v[r] - op.apply(x[r])
                assert np.linalg.norm(res) < 0.1 * np.linalg.norm(v[r]) * (1 + 1e-5)


def test_coarsest_persistent_64_systems():
    """The persistent coarsest GMRES at its largest batch (64 systems: every lane of the
    coefficient-flag wait and both halves of the 64-bit system masks in use), FP32 with the
    binary16 basis, one site per work-group (64 work-groups, one per system): every
    system equals the dense solve to 1e-4 at a tight tolerance, and systems 0, 33 and 63
    equal their single-system solves."""
    rng = np.random.default_rng(64)
    N = 64
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)) * 0.5)
    with ddamg.Context(N, N) as ctx:
        ctx.set_gauge(U)
        ctx.setup(ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[8, 8], m0=-0.2,
                               precision=1, max_rhs=64, rng_ranks=(2, 2), coarse_tol=1e-6, ortho_coarse=2))
        assert ctx.coarsest_stats()['persistent'] == 1
        shape = ctx.level_shape(2)
        v = rng.standard_normal((64,) + shape) + 1j * rng.standard_normal((64,) + shape)
        x = ctx.cycle(2, v, np.zeros_like(v), kcycle=True, precision=1)
        A, B, Cm = ctx.links(2, precision=1)
        op = orc.Stencil(A, B, Cm, -0.2)
        nn = int(np.prod(shape))
        D = np.stack([op.apply(e.reshape(shape)).ravel() for e in np.eye(nn)], axis=1)
        xd = np.linalg.solve(D, v.reshape(64, nn).T).T.reshape(v.shape)
        errs = [rel(x[r], xd[r]) for r in range(64)]
        print(f'64 systems: max error vs dense solve {max(errs):.2e}')
        assert max(errs) < 1e-4
        for r in (0, 33, 63):
            assert rel(ctx.cycle(2, v[r], np.zeros_like(v[r]), kcycle=True, precision=1), x[r]) < 1e-12


@pytest.mark.parametrize('precision', [0, 1])
def test_vcycle_graph_equals_eager(golden, precision):
    """V-cycle preconditioner captured as a HIP graph (graph = 1) and replayed once per outer
    FGMRES iteration, with the solver's own eager kernels and host synchronisations between
    the replays: a 4-system batch gives the iteration counts and solutions of the eager
    cycles exactly (every replay feeds the next iteration), a second solve replays the
    cached graphs with the same result, and at least one graph was captured."""
    meta, a = golden('g64_solve')
    rng = np.random.default_rng(9)
    B = np.stack([a['rhs']] + [rng.standard_normal(a['rhs'].shape) + 1j * rng.standard_normal(a['rhs'].shape)
                               for _ in range(3)])
    U = np.exp(1j * a['theta'])
    out = {}
    for graph in (0, 1):
        with _ctx_from_golden(meta, a) as ctx:
            ctx.setup(_prm(meta, nu2=meta['nu2'], fgmres_m=meta['fgmres_m'], cycle=0, precision=precision,
                           max_rhs=4, graph=graph))
            x, its, info = ctx.solve(B)
            if graph:
                assert ctx.graph_count() >= 1
                x2, its2, _ = ctx.solve(B)
                assert list(its2) == list(its) and np.array_equal(x2, x)
            out[graph] = (x, list(its))
    print(f'precision {precision}: eager iterations {out[0][1]}, graph {out[1][1]}')
    assert out[1][1] == out[0][1]
    assert rel(out[1][0], out[0][0]) < 1e-13
    for r in range(4):
        rr = np.linalg.norm(B[r] - orc.D_phi(U, out[1][0][r], meta['m0'])) / np.linalg.norm(B[r])
        assert rr < 1e-10


@pytest.mark.parametrize('precision,poly', [(0, 0), (1, 0), (1, 8)])
def test_kcycle_graph_equals_eager(golden, precision, poly):
    """K-cycle preconditioner captured as a HIP graph (graph = 1, cycle = 1) and replayed once
    per outer FGMRES iteration.  Inside the capture the K-cycle's FGMRES(5) x 2 solves at
    levels 1 and 2 run device-controlled: the Givens rotations, breaks and restart tests of
    src/fgmres.cpp:5-127 are made on the device and every launch is masked by the systems
    still running, so the graph issues all steps and the stopped systems do nothing in them.
    Three levels (64^2 -> 16^2 -> 8^2) and 4 systems: the iteration counts equal the eager
    K-cycles' (host-controlled FGMRES) and the solutions agree to rounding (the rotations
    are the same formulas evaluated on the device; FP64 1e-10, mixed 1e-6); a second solve
    replays the cached graphs bit for bit; every system meets the outer tolerance with the
    oracle's D_W.  The same with the coarsest polynomial preconditioner (poly = 8: the
    persistent kernel's polynomial variant as a node of the captured graph)."""
    meta, a = golden('g64_solve')
    rng = np.random.default_rng(10)
    B = np.stack([a['rhs']] + [rng.standard_normal(a['rhs'].shape) + 1j * rng.standard_normal(a['rhs'].shape)
                               for _ in range(3)])
    U = np.exp(1j * a['theta'])
    out = {}
    for graph in (0, 1):
        with _ctx_from_golden(meta, a) as ctx:
            ctx.setup(ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[8, 8], m0=meta['m0'],
                                   rng_ranks=tuple(meta['ranks']), nu2=meta['nu2'], fgmres_m=meta['fgmres_m'],
                                   cycle=1, precision=precision, max_rhs=4, graph=graph, coarse_poly=poly))
            x, its, info = ctx.solve(B)
            if graph:
                assert ctx.graph_count() >= 1
                x2, its2, _ = ctx.solve(B)
                assert list(its2) == list(its) and np.array_equal(x2, x)
            out[graph] = (x, list(its))
    d = rel(out[1][0], out[0][0])
    print(f'precision {precision} poly {poly}: eager iterations {out[0][1]}, graph {out[1][1]}, '
          f'solution difference {d:.2e}')
    assert out[1][1] == out[0][1]
    assert d < (1e-10 if precision == 0 else 1e-6)
    for r in range(4):
        rr = np.linalg.norm(B[r] - orc.D_phi(U, out[1][0][r], meta['m0'])) / np.linalg.norm(B[r])
        assert rr < 1e-10


def test_ragged_batch_equals_separate_solves():
    """dd_solve_many_dev: several lattices of different sizes (and masses), each with its
    own hierarchy and batch of right-hand sides, solved at the same time give every system
    the result of solving its context alone (contexts share no arithmetic), and every
    system reaches the outer tolerance with the FP64 operator."""
    torch = pytest.importorskip('torch')
    rng = np.random.default_rng(17)
    sizes, m0s, nr = [32, 48, 64], [-0.12, -0.15, -0.1], 3
    ctxs, bs, xs, U = [], [], [], []
    for N, m0 in zip(sizes, m0s):
        c = ddamg.Context(N, N)
        th = c.generate_gauge(2.0, 200, seed=N)
        U.append(np.exp(1j * th))
        c.setup(ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[8, 8], m0=m0, precision=1,
                             max_rhs=nr, rng_ranks=(2, 2)))
        B = rng.standard_normal((nr, N, N, 2)) + 1j * rng.standard_normal((nr, N, N, 2))
        bs.append(torch.from_numpy(B.view(np.float64).copy()).cuda())
        xs.append(torch.zeros_like(bs[-1]))
        ctxs.append(c)
    infos = ddamg.solve_many_dev(ctxs, [nr] * 3, [b.data_ptr() for b in bs], [x.data_ptr() for x in xs])
    torch.cuda.synchronize()
    for i, (c, N, m0) in enumerate(zip(ctxs, sizes, m0s)):
        xi = xs[i].cpu().numpy().view(np.complex128).reshape(nr, N, N, 2)
        Bi = bs[i].cpu().numpy().view(np.complex128).reshape(nr, N, N, 2)
        xa, its, info = c.solve(Bi)
        assert rel(xi, xa) < 1e-12
        assert infos[i].iterations == info.iterations
        for r in range(nr):
            res = np.linalg.norm(Bi[r] - orc.D_phi(U[i], xi[r], m0)) / np.linalg.norm(Bi[r])
            assert res < 1e-10
        c.close()


@pytest.mark.parametrize('links16', [0, 1])
def test_fused_sap_iteration_equals_separate_launches(links16):
    """The fused SAP iteration (one launch: red blocks, black blocks, residual norm and
    stopping test, each block forming v - D x itself) gives the result of the separate
    launches (block solves with full-lattice residuals between the colours), to FP32
    rounding, for a batch of 20 systems (two 16-system tiles, the second partial; block
    GMRES tolerance 1e-12, which FP32 cannot reach, so that rounding cannot move a block
    solve's break and both forms run the same iterations); with a
    loose tolerance the same systems stop early (src/sap.cpp:150-156), which the
    iterates show: a system that stops after one iteration keeps its one-iteration
    result."""
    rng = np.random.default_rng(44)
    N = 64
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)) * 0.4)
    res = {}
    for fused in (1, 0):
        prm = ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[12, 8], m0=-0.2, precision=1,
                           max_rhs=20, rng_ranks=(2, 2), sap_variant=33 if links16 else 32,
                           coarse_sap_links=links16, sap_fused=fused, sap_tol=1e-12)
        with ddamg.Context(N, N) as ctx:
            ctx.set_gauge(U)
            ctx.setup(prm)
            shape = ctx.level_shape(1)
            if fused:
                v = rng.standard_normal((20,) + shape) + 1j * rng.standard_normal((20,) + shape)
                x0 = 0.3 * (rng.standard_normal((20,) + shape) + 1j * rng.standard_normal((20,) + shape))
                # a system that starts near its solution: its residual test holds after the
                # first iteration
                v[5] = ctx.level_apply(1, x0[5], precision=1) + 1e-3 * (rng.standard_normal(shape) +
                                                                         1j * rng.standard_normal(shape))
            res[fused] = (ctx.sap(1, v, x0, 3, tol=1e-12, precision=1), ctx.sap(1, v, x0, 3, tol=0.05, precision=1),
                          ctx.sap(1, v, x0, 1, tol=1e-12, precision=1))
    for k in range(3):
        d = max(rel(res[1][k][r], res[0][k][r]) for r in range(20))
        print(f'fused vs separate launches (case {k}): {d:.2e}')
        assert d < 1e-5
    # the loose tolerance stops some systems after one iteration in both forms
    one_step = [r for r in range(20) if rel(res[1][1][r], res[1][2][r]) < 1e-6]
    one_step0 = [r for r in range(20) if rel(res[0][1][r], res[0][2][r]) < 1e-6]
    print('systems stopped after one iteration:', one_step)
    assert one_step == one_step0 and 5 in one_step


@pytest.mark.parametrize('nu', [1, 2])
@pytest.mark.parametrize('shape', [(16, 16), (20, 24), (24, 20)])
def test_level0_sap_block_counts(shape, nu):
    """Level-0 SAP against the oracle's SAP (src/sap.cpp:81-165) on the library's own
    level-0 operator, for 3 systems: with even block counts (the blocks form v - D x at
    their own sites) and with an odd count in x or in t (full residuals between the
    colours), after 1 and 2 iterations (no residual after a call's last iteration).  FP64
    to 1e-11; the FP32 hierarchy to FP32 accuracy."""
    Nx, Nt = shape
    rng = np.random.default_rng(Nx * 100 + Nt + nu)
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (Nx, Nt, 2)) * 0.5)
    v = rng.standard_normal((3, Nx, Nt, 2)) + 1j * rng.standard_normal((3, Nx, Nt, 2))
    x0 = 0.1 * (rng.standard_normal((3, Nx, Nt, 2)) + 1j * rng.standard_normal((3, Nx, Nt, 2)))
    m0 = -0.2
    refs = None
    for precision in (0, 1):
        prm = ddamg.params(2, agg=[(4, 4)], sap_block=[(4, 4)], nv=[8], m0=m0, precision=precision, max_rhs=3,
                           rng_ranks=(1, 1))
        with ddamg.Context(Nx, Nt) as ctx:
            ctx.set_gauge(U)
            ctx.setup(prm)
            if refs is None:   # the oracle on the FP64 hierarchy's level-0 operator
                A, B, Cm = ctx.links(0)
                op = orc.Stencil(A, B, Cm, m0)
                blocks = orc.SapBlocks(Nx, Nt, 4, 4)
                refs = [orc.sap(op, blocks, v[r], x0[r], nu, 1e-10)[0] for r in range(3)]
            x = ctx.sap(0, v, x0, nu, precision=precision)
            for r in range(3):
                err = rel(x[r], refs[r])
                print(f'{Nx}x{Nt}, nu {nu}, precision {precision}, system {r}: {err:.2e}')
                assert err < (1e-11 if precision == 0 else 2e-5)
