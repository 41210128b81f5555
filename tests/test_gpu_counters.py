"""GPU test of the coarse SAP's device flop counter (dd_timer_flops), the figure behind
the matrix-core roofline that bench.py reports for the coarse SAP.

The kernels count 128 dof^2 flops per (site, hop, 16-system tile) product: the block
operator's in-block hops at every application, and in the fused form also the five hops
of the block right-hand side v - D x and of the norm stencil.  With the block GMRES's
tolerance out of reach (sap_tol = 1e-30; GMRES(5) with 5 restarts, src/fgmres.cpp:5-127)
every block solve makes exactly 1 + 5 x (5 + 1) = 31 applications, so the count of one
SAP iteration is fixed by the geometry alone:

- fused (one launch per iteration): blocks x (31 x in-block pairs + 5 x sites per block)
  + lattice sites x 5 (the norm stencil), the norm stencil in every iteration but the
  last (whose residual norm would feed no stopping test);
- separate launches per colour: blocks x 31 x in-block pairs;

with 64 in-block (site, hop) pairs on a 4x4 block (16 self, 24 in t, 24 in x).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')


@pytest.mark.parametrize('nu', [1, 2])
@pytest.mark.parametrize('fused', [1, 0])
def test_sap_flop_count(fused, nu):
    N, nrhs = 64, 16
    rng = np.random.default_rng(41)
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)) * 0.5)
    with ddamg.Context(N, N) as ctx:
        ctx.set_gauge(U)
        ctx.setup(ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (4, 4)], nv=[8, 8], m0=-0.2,
                               precision=1, max_rhs=nrhs, rng_ranks=(2, 2), sap_tol=1e-30, sap_variant=32,
                               sap_fused=fused))
        Nx, Nt, dof = ctx.level_shape(1)
        v = rng.standard_normal((nrhs, Nx, Nt, dof)) + 1j * rng.standard_normal((nrhs, Nx, Nt, dof))
        ddamg.timer_enable(ddamg.K_SAPC, True)
        try:
            ddamg.timer_reset()
            ctx.sap(1, v, np.zeros_like(v), nu, tol=1e-30, precision=1)
            flops = ddamg.timer_flops(ddamg.K_SAPC)
            _, launches, _ = ddamg.timer_read(ddamg.K_SAPC)
        finally:
            ddamg.timer_enable(ddamg.K_SAPC, False)
    bx = bt = 4
    S = bx * bt
    blocks = (Nx // bx) * (Nt // bt)
    inblock = S + 2 * bx * (bt - 1) + 2 * bt * (bx - 1)
    assert inblock == 64
    if fused:
        pairs = nu * blocks * (31 * inblock + 5 * S) + (nu - 1) * Nx * Nt * 5
    else:
        pairs = nu * blocks * 31 * inblock
    expected = pairs * 128 * dof * dof
    print(f'fused {fused}: {launches} launches, {flops:.0f} flops counted, {expected} expected')
    assert launches == (nu if fused else 2 * nu)
    assert flops == expected
