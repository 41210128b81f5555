"""torch.autograd.Function over the C ABI: x = D(theta, m0)^-1 b with gradients with
respect to the link phases theta, the bare mass m0 and b.  Used for gradient checks.

Backward reuses the forward's multigrid hierarchy: one adjoint solve
y = D^-dagger g = sigma3 D^-1 sigma3 g (dd_solve_adjoint) and the hand-written
stencil-derivative kernel (dd_dD_dtheta):
    dL/dtheta_mu(n) = -Re[ y^H dD/dtheta_mu(n) x ],  dL/dm0 = -Re(y^H x),  dL/db = y,
with g = dL/dx in torch's convention for complex tensors (the conjugate Wirtinger
derivative times two).  Everything runs on the GPU through libddamg.so; tensors cross
the boundary through host memory.
"""
from __future__ import annotations

import numpy as np
import torch

from . import Context, params


class WilsonSolver:
    """One lattice and hierarchy parameters; the hierarchy is rebuilt when theta or m0
    change (forward), and reused by the adjoint solve (backward)."""

    def __init__(self, Nx: int, Nt: int, prm=None, device: int = 0, **kw):
        self.ctx = Context(Nx, Nt, device)
        self.prm = prm if prm is not None else params(**kw)
        self._key = None

    def prepare(self, theta: np.ndarray, m0: float):
        key = (hash(theta.tobytes()), float(m0))
        if key != self._key:
            self.ctx.set_gauge_theta(theta)
            self.prm.m0 = float(m0)
            self.ctx.setup(self.prm)
            self._key = key

    def solve(self, b: np.ndarray) -> np.ndarray:
        return self.ctx.solve(b)[0]

    def solve_adjoint(self, g: np.ndarray) -> np.ndarray:
        return self.ctx.solve(g, adjoint=True)[0]


class WilsonSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta: torch.Tensor, m0: torch.Tensor, b: torch.Tensor, solver: WilsonSolver):
        th = theta.detach().cpu().numpy().astype(np.float64)
        solver.prepare(th, float(m0))
        x = solver.solve(b.detach().cpu().numpy())
        ctx.solver = solver
        ctx.theta = th
        ctx.m0 = float(m0)
        xt = torch.from_numpy(x).to(b.device)
        ctx.save_for_backward(xt)
        return xt

    @staticmethod
    def backward(ctx, gx: torch.Tensor):
        (x,) = ctx.saved_tensors
        solver: WilsonSolver = ctx.solver
        solver.prepare(ctx.theta, ctx.m0)   # no-op unless another forward intervened
        y = solver.solve_adjoint(gx.detach().cpu().numpy())
        xn = x.detach().cpu().numpy()
        g_theta = solver.ctx.dD_dtheta(y, xn, scale=-1.0)
        g_m0 = -np.real(np.vdot(y, xn))
        dev = x.device
        return (torch.from_numpy(g_theta).to(dev), torch.tensor(g_m0, dtype=torch.float64, device=dev),
                torch.from_numpy(y).to(dev), None)


def wilson_solve(theta: torch.Tensor, m0: torch.Tensor, b: torch.Tensor, solver: WilsonSolver) -> torch.Tensor:
    return WilsonSolve.apply(theta, m0, b, solver)
