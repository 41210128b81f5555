"""The reference's comparison solvers on the GPU (dd_solve_cg, dd_solve_bicgstab,
dd_solve_gmres): CG on the normal equations, BiCGStab, GMRES(20) and FGMRES(20)+SAP,
against the compiled reference's iteration counts and solutions (g32_methods fixture),
and batched systems against single ones."""
import numpy as np
import pytest

from oracle import ddamg_oracle as orc

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')

# (method, fixture key).  The iteration counts must equal the reference's: the GPU forms
# the dot products in another order than the reference's MPI sums, but on this fixture
# that rounding moves no count.
METHODS = [('cg', 'cg'), ('bicgstab', 'bicgstab'), ('gmres', 'gmres'), ('fgmres_sap', 'fgmres_sap')]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _ctx(meta, a):
    ctx = ddamg.Context(meta['N'], meta['N'])
    ctx.set_gauge(np.exp(1j * a['theta']))
    return ctx


# FGMRES+SAP: the reference's 4x4 blocks per rank of the 2x2 grid are 4x4-site blocks
def _kw(meta):
    return dict(tol=meta['tol'], max_iter=meta['max_iter'], m=meta['gmres_m'], restarts=meta['gmres_restarts'],
                sap_block=tuple(meta['sap_block']))


@pytest.mark.parametrize('method,key', METHODS)
def test_methods_golden(golden, method, key):
    meta, a = golden('g32_methods')
    with _ctx(meta, a) as ctx:
        x, its, info = ctx.solve_method(method, a['rhs'], meta['m0'], **_kw(meta))
    U = np.exp(1j * a['theta'])
    m0 = meta['m0']
    rr = np.linalg.norm(a['rhs'] - orc.D_phi(U, x, m0)) / np.linalg.norm(a['rhs'])
    print(f'{method}: {its[0]} iterations (reference {meta[key + "_iterations"]}), relres {rr:.3e} '
          f'(reference {meta[key + "_relres"]:.3e}), error vs reference x {rel(x, a["x_" + key]):.2e}')
    assert int(its[0]) == meta[key + '_iterations']
    if method == 'cg':   # info.relres: the normal equations' residual
        phi = orc.D_dagger_phi(U, a['rhs'], m0)
        rn = np.linalg.norm(phi - orc.D_dagger_phi(U, orc.D_phi(U, x, m0), m0)) / np.linalg.norm(phi)
        assert abs(info.relres - rn) <= 1e-3 * rn + 1e-14 and info.relres < meta['tol']
    else:
        assert abs(info.relres - rr) <= 1e-3 * rr + 1e-14
    assert rr < 2 * meta[key + '_relres'] + 1e-12
    assert rel(x, a['x_' + key]) < 1e-12


@pytest.mark.parametrize('method', ['cg', 'bicgstab', 'fgmres_sap'])
def test_methods_batched_equal_single(golden, method):
    """Three systems solved together give each the iteration count and the solution it
    gets alone (systems share no arithmetic)."""
    meta, a = golden('g32_methods')
    rng = np.random.default_rng(7)
    B = np.stack([a['rhs']] + [rng.standard_normal(a['rhs'].shape) + 1j * rng.standard_normal(a['rhs'].shape)
                               for _ in range(2)])
    kw = _kw(meta)
    with _ctx(meta, a) as ctx:
        xb, itb, _ = ctx.solve_method(method, B, meta['m0'], **kw)
        for r in range(3):
            x1, it1, _ = ctx.solve_method(method, B[r], meta['m0'], **kw)
            assert int(it1[0]) == int(itb[r])
            assert rel(xb[r], x1) < 1e-12


@pytest.mark.parametrize('nu,tol', [(3, 1e-12), (40, 1e-6)])
def test_sap_standalone_matches_oracle(golden, nu, tol):
    """SAP_fine_level::SAP as a solver (dd_sap_fine, src/sap.cpp:81-165) equals the oracle's
    SAP on the same field: the reference's return value (the iteration at which
    ||v - D x|| < tol ||v|| first held, or nu) and the result; batched equals single."""
    meta, a = golden('g32_methods')
    U = np.exp(1j * a['theta'])
    m0 = meta['m0']
    rng = np.random.default_rng(8)
    x0 = 0.1 * (rng.standard_normal(a['rhs'].shape) + 1j * rng.standard_normal(a['rhs'].shape))
    ref, its_ref = orc.sap(orc.dirac_stencil(U, m0), orc.SapBlocks(32, 32, 4, 4), a['rhs'], x0, nu, tol)
    with _ctx(meta, a) as ctx:
        x, its, rr = ctx.sap_fine(a['rhs'], x0, nu, tol, (4, 4), m0)
        xb, itsb, _ = ctx.sap_fine(np.stack([a['rhs'], a['rhs']]), np.stack([x0, x0]), nu, tol, (4, 4), m0)
    print(f'SAP nu={nu} tol={tol}: {its[0]} (oracle {its_ref}), relres {rr[0]:.3e}, vs oracle {rel(x, ref):.2e}')
    assert int(its[0]) == its_ref
    assert rel(x, ref) < 1e-12
    assert int(itsb[1]) == its_ref and rel(xb[1], x) < 1e-14
