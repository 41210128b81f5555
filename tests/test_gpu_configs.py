"""GPU tests at the benchmark workloads' full sizes, through the C ABI.

- 256^2, plain FP64 (3 levels, 4x4 aggregates, Nv=12, 4x4-site SAP blocks, nu2=2,
  FGMRES(30), tol 1e-10): K- and V-cycle iteration counts equal the reference's on the
  same field, right-hand side and test-vector streams (fixture g256_solve, the reference
  compiled from its sources and run on a 2x4 rank grid), true FP64 residual < 1e-10, and
  the solution at 512 sampled sites within 1e-8 of the reference's.
- 512^2, the same hierarchy: plain FP64 K-cycle count equal to the reference's
  (g512_solve); the mixed-precision (FP32 hierarchy inside FP64 FGMRES) batch of 32
  right-hand sides reaches 1e-10 in FP64 residual for every system, and its system 0
  is within 1e-8 of the FP64 solve.
- 32 right-hand sides on one 256^2 field in mixed precision: every residual < 1e-10
  and sampled systems equal their separate single-system solves to 1e-8 (the K-cycle
  with the plain coarsest GMRES, and the V-cycle with the degree-12 coarsest
  polynomial, as in the benchmark's rows; also for the ragged batch below).
- The ragged batch of 8 fields {64,96,128,128,192,256,256,256}^2 with m0 drawn from
  U[-0.18, -0.10] and 4 right-hand sides each, mixed precision, solved at the same time
  (dd_solve_many_dev): every residual < 1e-10, and each lattice's result equals its own
  solve of the same right-hand sides to 1e-10.
- D_W and D^dagger at 512^2 for a 32-system batch against the oracle to 1e-14 (FP64).

Fields come from the library's device Metropolis generator (beta=2, cold start, 2000
sweeps of +-0.5 rad); the fixtures hold the SHA-256 of the field the reference solved,
and the tests first check that the generator reproduced it.  Reference check of the
residual: src/fgmres.cpp:77-90; the solves batched: include/amg.h:240-270.
"""
import hashlib

import numpy as np
import pytest

from oracle import ddamg_oracle as orc

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.ravel(b))


def gaussian(N, seed, nrhs=None, dof=2):
    rng = np.random.default_rng(seed)
    shape = (N, N, dof) if nrhs is None else (nrhs, N, N, dof)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def hierarchy(meta, precision, nrhs, cycle='K', m0=None, coarse_poly=0):
    lv = meta['levels']
    return ddamg.params(len(lv) + 1, agg=[tuple(l['agg']) for l in lv],
                        sap_block=[tuple(l['sap_block']) for l in lv], nv=[l['nv'] for l in lv],
                        m0=meta['m0'] if m0 is None else m0, nu2=meta['nu2'], fgmres_m=meta['fgmres_m'],
                        cycle=1 if cycle == 'K' else 0, precision=precision, max_rhs=nrhs,
                        rng_ranks=tuple(meta['ranks']), coarse_poly=coarse_poly)


# the mixed-precision forms of the batched tests: the reference's K-cycle with its plain
# coarsest GMRES, and the benchmark's V-cycle with the degree-12 coarsest polynomial
MIXED_FORMS = [('K', 0), ('V', 12)]


def field_ctx(meta):
    """Context on the fixture's field, regenerated on the device and checked bit for bit."""
    N = meta['N']
    ctx = ddamg.Context(N, N)
    th = ctx.generate_gauge(meta['beta'], meta['sweeps'], meta['seed'], meta['step'])
    assert hashlib.sha256(np.ascontiguousarray(th).tobytes()).hexdigest() == meta['theta_sha256'], \
        'the device generator no longer reproduces the field the reference solved'
    return ctx, np.exp(1j * th)


def true_relres(ctx, U, x, b, m0, oracle_systems=(0,)):
    """||b - D x|| / ||b|| per system with the library's FP64 D_W (pinned to the reference
    at 1e-14), and with the oracle's D_phi for the listed systems."""
    xb = x if x.ndim == 4 else x[None]
    bb = b if b.ndim == 4 else b[None]
    Dx = ctx.apply_D(xb, m0)
    rr = np.linalg.norm((bb - Dx).reshape(len(bb), -1), axis=1) / np.linalg.norm(bb.reshape(len(bb), -1), axis=1)
    for r in oracle_systems:
        ro = np.linalg.norm(bb[r] - orc.D_phi(U, xb[r], m0)) / np.linalg.norm(bb[r])
        assert abs(ro - rr[r]) <= 1e-3 * ro + 1e-15
    return rr


@pytest.mark.parametrize('cycle', ['K', 'V'])
def test_256_fp64_matches_reference(golden, cycle):
    meta, a = golden('g256_solve')
    ctx, U = field_ctx(meta)
    with ctx:
        ctx.setup(hierarchy(meta, 0, 1, cycle))
        b = gaussian(meta['N'], meta['rhs_seed'])
        x, its, info = ctx.solve(b)
        rr = true_relres(ctx, U, x, b, meta['m0'])[0]
        print(f'256^2 FP64 {cycle}-cycle: {its[0]} iterations (reference {meta[cycle + "_iterations"]}), '
              f'relres {rr:.3e} (reference {meta[cycle + "_relres"]:.3e})')
        assert int(its[0]) == meta[f'{cycle}_iterations']
        assert rr < 1e-10 and info.relres < 1e-10
        xs = x.reshape(-1, 2)[a['sites']]
        err = rel(xs, a[f'x_{cycle}'])
        print(f'  solution at sampled sites vs reference: {err:.3e}')
        assert err < 1e-8


def test_512_fp64_and_mixed_batch(golden):
    meta, a = golden('g512_solve')
    N, m0 = meta['N'], meta['m0']
    ctx, U = field_ctx(meta)
    b0 = gaussian(N, meta['rhs_seed'])
    with ctx:
        ctx.setup(hierarchy(meta, 0, 1))
        x64, its, info = ctx.solve(b0)
        rr = true_relres(ctx, U, x64, b0, m0)[0]
        print(f'512^2 FP64 K-cycle: {its[0]} iterations (reference {meta["K_iterations"]}), relres {rr:.3e}')
        assert int(its[0]) == meta['K_iterations']
        assert rr < 1e-10
        assert rel(x64.reshape(-1, 2)[a['sites']], a['x_K']) < 1e-8

        B = np.concatenate([b0[None], gaussian(N, 9001, nrhs=31)])
        ctx.setup(hierarchy(meta, 1, 32))
        xm, itm, infom = ctx.solve(B)
        rrm = true_relres(ctx, U, xm, B, m0, oracle_systems=(0, 31))
        err = rel(xm[0], x64)
        print(f'512^2 mixed, 32 RHS: iterations {itm.min()}..{itm.max()}, max relres {rrm.max():.3e}, '
              f'system 0 vs FP64 solve {err:.3e}')
        assert rrm.max() < 1e-10 and infom.relres < 1e-10
        assert err < 1e-8


@pytest.mark.parametrize('cycle,poly', MIXED_FORMS)
def test_256_mixed_32rhs_equals_separate_solves(golden, cycle, poly):
    meta, _ = golden('g256_solve')
    N, m0 = meta['N'], meta['m0']
    ctx, U = field_ctx(meta)
    B = gaussian(N, 9002, nrhs=32)
    with ctx:
        ctx.setup(hierarchy(meta, 1, 32, cycle, coarse_poly=poly))
        xb, itb, info = ctx.solve(B)
        rr = true_relres(ctx, U, xb, B, m0, oracle_systems=(0, 17))
        print(f'256^2 mixed, 32 RHS, {cycle}-cycle, coarse_poly {poly}: iterations {itb.min()}..{itb.max()}, '
              f'max relres {rr.max():.3e}')
        assert rr.max() < 1e-10
        for r in (0, 13, 31):
            x1, it1, _ = ctx.solve(B[r])
            d = rel(xb[r], x1)
            print(f'  system {r}: batched {itb[r]} / separate {it1[0]} iterations, difference {d:.3e}')
            assert d < 1e-8


@pytest.mark.parametrize('cycle,poly', MIXED_FORMS)
def test_ragged_batch_of_eight_fields(cycle, poly):
    torch = pytest.importorskip('torch')
    sizes = [64, 96, 128, 128, 192, 256, 256, 256]
    rng = np.random.default_rng(4242)
    m0s = rng.uniform(-0.18, -0.10, len(sizes))
    meta = dict(levels=[dict(agg=[4, 4], sap_block=[4, 4], nv=12)] * 2, nu2=2, fgmres_m=30, ranks=[2, 2], m0=-0.18)
    ctxs, Bs, Us = [], [], []
    try:
        for i, (N, m0) in enumerate(zip(sizes, m0s)):
            c = ddamg.Context(N, N)
            th = c.generate_gauge(2.0, 2000, 300 + i)
            c.setup(hierarchy(meta, 1, 4, cycle, m0=float(m0), coarse_poly=poly))
            ctxs.append(c)
            Us.append(np.exp(1j * th))
            Bs.append(gaussian(N, 400 + i, nrhs=4))
        bd = [torch.from_numpy(B.view(np.float64).copy()).cuda() for B in Bs]
        xd = [torch.zeros_like(v) for v in bd]
        infos = ddamg.solve_many_dev(ctxs, [4] * len(sizes), [v.data_ptr() for v in bd], [v.data_ptr() for v in xd])
        torch.cuda.synchronize()
        for i, (c, N, m0) in enumerate(zip(ctxs, sizes, m0s)):
            x = xd[i].cpu().numpy().view(np.complex128).reshape(4, N, N, 2)
            rr = true_relres(c, Us[i], x, Bs[i], float(m0))
            xs, its, _ = c.solve(Bs[i])
            d = rel(x, xs)
            print(f'ragged {N}^2 m0={m0:.4f} ({cycle}-cycle, coarse_poly {poly}): iterations {infos[i].iterations}, '
                  f'max relres {rr.max():.3e}, '
                  f'vs own solve {d:.3e}')
            assert rr.max() < 1e-10 and infos[i].relres < 1e-10
            assert d < 1e-10
    finally:
        for c in ctxs:
            c.close()


def test_dirac_full_size(golden):
    """D_W and D^dagger at 512^2 on the fixture's field for 32 systems, the shape of the
    benchmark's batch (XCD-ordered work-groups, 4 systems per thread): each checked system
    equals the oracle's D_phi / D_dagger_phi (src/dirac_operator.cpp:24-85) to 1e-14
    relative in FP64 and 1e-6 in FP32, and batched systems equal single ones bit for bit."""
    meta, _ = golden('g512_solve')
    N = meta['N']
    ctx, U = field_ctx(meta)
    V = gaussian(N, 77, nrhs=32)
    with ctx:
        D64 = ctx.apply_D(V, meta['m0'])
        Dd = ctx.apply_D(V, meta['m0'], dagger=True)
        D32 = ctx.apply_D(V, meta['m0'], precision=1)
        one = ctx.apply_D(V[13], meta['m0'])
    assert np.array_equal(one, D64[13])
    for r in (0, 13, 31):
        ref = orc.D_phi(U, V[r], meta['m0'])
        refd = orc.D_dagger_phi(U, V[r], meta['m0'])
        e64, ed, e32 = rel(D64[r], ref), rel(Dd[r], refd), rel(D32[r], ref)
        print(f'512^2 system {r}: D {e64:.2e}, D^dagger {ed:.2e}, FP32 D {e32:.2e}')
        assert e64 < 1e-14 and ed < 1e-14 and e32 < 1e-6
