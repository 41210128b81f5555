"""Differentiation through the solve (cfg5 shape): L = Re<eta, D(theta, m0)^-1 b>.
GPU gradient = forward solve + adjoint solve on the same hierarchy + dD/dtheta kernel,
checked against (1) the oracle formula on exact dense solves (small lattice),
(2) FP64 central finite differences of the GPU solve (h = 1e-6, 64 random links and
m0) at 128^2, and (3) a reference gradient built from two solves of the compiled
reference CPU code (oracle/_ref) where it is available."""
import numpy as np
import pytest

from oracle import ddamg_oracle as orc

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')
torch = pytest.importorskip('torch')
from ddamg.autograd import WilsonSolver, wilson_solve  # noqa: E402


def test_dtheta_kernel_dense():
    rng = np.random.default_rng(4)
    Nx, Nt, m0 = 6, 8, -0.25
    th = rng.uniform(-np.pi, np.pi, (Nx, Nt, 2))
    U = np.exp(1j * th)
    D = orc.dense_D(U, m0)
    b = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))
    eta = rng.standard_normal((Nx, Nt, 2)) + 1j * rng.standard_normal((Nx, Nt, 2))
    x = np.linalg.solve(D, b.ravel()).reshape(b.shape)
    y = np.linalg.solve(D.conj().T, eta.ravel()).reshape(b.shape)
    with ddamg.Context(Nx, Nt) as ctx:
        ctx.set_gauge(U)
        g = ctx.dD_dtheta(y, x, scale=-1.0)
    ref = orc.dL_dtheta(U, y, x)
    assert np.abs(g - ref).max() < 1e-12 * np.abs(ref).max()


def _cfg5(seed=128):
    N, m0 = 128, -0.15
    prm = ddamg.params(3, agg=[(4, 4), (4, 4)], sap_block=[(4, 4), (4, 4)], nv=[8, 8], m0=m0, cycle=1,
                       tol=1e-13, fgmres_m=30, rng_ranks=(2, 2))
    solver = WilsonSolver(N, N, prm)
    th = solver.ctx.generate_gauge(2.0, 2000, seed)
    rng = np.random.default_rng(seed + 1)
    b = rng.standard_normal((N, N, 2)) + 1j * rng.standard_normal((N, N, 2))
    eta = rng.standard_normal((N, N, 2)) + 1j * rng.standard_normal((N, N, 2))
    return N, m0, solver, th, b, eta


def test_gradient_autograd_vs_fd_128():
    N, m0, solver, th, b, eta = _cfg5()
    theta = torch.tensor(th, requires_grad=True)
    m0t = torch.tensor(m0, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b)
    x = wilson_solve(theta, m0t, bt, solver)
    L = torch.real(torch.vdot(torch.tensor(eta).flatten(), x.flatten()))
    L.backward()
    g_th = theta.grad.numpy()
    g_m0 = float(m0t.grad)

    def Lnum(t, m):
        solver.prepare(t, m)
        return float(np.real(np.vdot(eta, solver.solve(b))))

    h = 1e-6
    rng = np.random.default_rng(7)
    picks = [(rng.integers(N), rng.integers(N), rng.integers(2)) for _ in range(64)]
    scale = np.abs(g_th).max()
    worst = 0.0
    for (ix, it, mu) in picks:
        tp, tm = th.copy(), th.copy()
        tp[ix, it, mu] += h
        tm[ix, it, mu] -= h
        fd = (Lnum(tp, m0) - Lnum(tm, m0)) / (2 * h)
        worst = max(worst, abs(fd - g_th[ix, it, mu]) / scale)
    fd_m = (Lnum(th, m0 + h) - Lnum(th, m0 - h)) / (2 * h)
    assert worst < 1e-4
    assert abs(fd_m - g_m0) < 1e-4 * max(1.0, abs(fd_m))


def test_gradient_vs_reference_solves_128():
    from oracle import run_ref
    if not run_ref.available(128):
        pytest.skip('reference binary for 128^2 not built')
    N, m0, solver, th, b, eta = _cfg5()
    solver.prepare(th, m0)
    x = solver.solve(b)
    y = solver.solve_adjoint(eta)
    g = solver.ctx.dD_dtheta(y, x, scale=-1.0)
    U = np.exp(1j * th)
    levels = [orc.LevelSpec((4, 4), (4, 4), 8), orc.LevelSpec((4, 4), (4, 4), 8)]
    xs = []
    for rhs in (b, orc.sigma3(eta)):
        wd, _ = run_ref.run('solve', U, rhs, levels, m0, ranks=(2, 2), nu2=2, fgmres_m=30, cycle='K')
        xs.append(run_ref.load(wd, 'x', (2, 2)))
        run_ref.cleanup(wd)
    x_ref, y_ref = xs[0], orc.sigma3(xs[1])   # D^-dagger eta = sigma3 D^-1 sigma3 eta
    g_ref = orc.dL_dtheta(U, y_ref, x_ref)
    assert np.abs(g - g_ref).max() < 1e-7 * np.abs(g_ref).max()
