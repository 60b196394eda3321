"""Discrete-adjoint force-coefficient sensitivities behind torch.autograd.

The device computes the exact linearisation J = dR/dW of the first-order
residual, inviscid or with the viscous terms (rdst_adjoint_linearise; with a
turbulence model its eddy viscosity frozen) and solves J^T lambda = dC/dW
(rdst_adjoint_solve, FP64 or FP32 transposed ILU inside an FP64 defect
correction).  With the state converged (R(W*, x) = 0) the total derivative of
a force coefficient C with respect to shape parameters a is

    dC/da = dC/da|_W - lambda^T dR/da|_W,

where the two partial derivatives at the frozen state are central
differences of the residual and of C on the displaced meshes (the residual
is recomputed from W on each: no face state is stored).

Two autograd Functions:

* ``ForceCoefficient.apply(a, problem)``: a -> C(a) with the flow converged on
  the displaced mesh; backward runs the adjoint.
* ``StateCoefficient.apply(W, problem)``: the frozen-state map W -> C(W) on the
  problem's current mesh; backward is the exact dC/dW from the device.
"""
from __future__ import annotations

import numpy as np
import torch

from . import cases
from ._capi import DeviceCase


class BumpProblem:
    """Cd (or Cl) of the bottom wall of cases.bump_case as a function of the
    Hicks-Henne amplitudes of the wall, with the flow converged to a residual
    drop of ``drop`` (||R||_2 of timeDST.C:328-331)."""

    def __init__(self, n_modes=4, nx=48, ny=24, mach=0.5, scheme="HLLC", direction=(1.0, 0.0, 0.0),
                 drop=1e-12, max_steps=4000, adjoint_precision="fp64", fd_step=1e-7, **settings):
        self.n_modes = n_modes
        self.kw = dict(nx=nx, ny=ny, mach=mach, scheme=scheme, **settings)
        self.direction = np.asarray(direction, dtype=np.float64)
        self.drop = drop
        self.max_steps = max_steps
        self.adjoint_precision = adjoint_precision
        self.fd_step = fd_step
        gamma, R, T, p = 1.4, 287.0, 300.0, 1.0e5
        rho = p / (R * T)
        U = mach * np.sqrt(gamma * R * T)
        self.p_ref = p
        self.q = 0.5 * rho * U * U * 1.0 * 0.1  # unit chord, one-cell depth 0.1
        self.dc = None
        self.amplitudes = None

    # ---- primal
    def case(self, amplitudes, node_disp=None):
        return cases.bump_case(amplitudes=tuple(amplitudes), node_disp=node_disp, **self.kw)

    def converge(self, dc):
        r0 = None
        steps = 0
        while steps < self.max_steps:
            res = dc.iterate(100)
            steps += 100
            if r0 is None:
                r0 = res[0]
            if not np.all(np.isfinite(res)):
                raise RuntimeError("bump case diverged")
            if res[-1] <= self.drop * r0:
                return steps, res[-1] / r0
            print(f"[bump] step {steps} drop {res[-1] / r0:.3e}", flush=True)
        raise RuntimeError(f"bump case did not reach a {self.drop:g} residual drop in {self.max_steps} steps "
                           f"({res[-1] / r0:.3e})")

    def coefficient(self, dc, pressure_only=True):
        patch = dc.spec.mesh.patch_index("bottom")
        out = dc.forces(patch, self.direction, (0.0, 1.0, 0.0), self.p_ref, 1.0, 1.0, 1.0)
        F = out[2:5]  # the force vector (pressure and, viscous, the viscous part)
        return float(F @ self.direction / self.q)

    def solve(self, amplitudes):
        a = np.asarray(amplitudes, dtype=np.float64)
        dc = DeviceCase(self.case(a))
        steps, drop = self.converge(dc)
        self.dc, self.amplitudes = dc, a
        return self.coefficient(dc), steps, drop

    # ---- frozen-state pieces
    def _frozen(self, amplitudes, W, node_disp=None):
        """A case on the displaced mesh carrying state W with consistent
        boundary values and masks: put W, update, face fluxes, update."""
        dc = DeviceCase(self.case(amplitudes, node_disp=node_disp))
        dc.put(9, W)
        dc.update_fields()
        dc.flux_residual()
        dc.update_fields()
        if self.kw.get("viscous"):
            # the boundary gradient that the viscous force reads, formed from
            # boundary values consistent with W
            dc.flux_residual()
            dc.update_fields()
        return dc

    def partials(self, W, amplitudes, direction_vec, node_disp=None):
        """dR/da . direction_vec and dC/da . direction_vec at frozen W (central
        differences of the residual and of C on the displaced meshes)."""
        h = self.fd_step
        a = np.asarray(amplitudes, dtype=np.float64)
        if node_disp is None:
            dp = self._frozen(a + h * direction_vec, W)
            dm = self._frozen(a - h * direction_vec, W)
        else:
            dp = self._frozen(a, W, node_disp=node_disp + h * direction_vec)
            dm = self._frozen(a, W, node_disp=node_disp - h * direction_vec)
        dR = (dp.adjoint_residual() - dm.adjoint_residual()) / (2 * h)
        dC = (self.coefficient(dp) - self.coefficient(dm)) / (2 * h)
        return dR, dC

    def adjoint(self, **solver):
        dc = self.dc
        patch = dc.spec.mesh.patch_index("bottom")
        opts = dict(precision=self.adjoint_precision, n_directions=30, max_cycles=400, tolerance=1e-13, rel_tol=0.0)
        opts.update(solver)
        lam, perf = dc.adjoint_solve(patch, self.direction, self.q, **opts)
        return lam, perf

    def gradient(self, lam=None):
        """dC/da by the adjoint at the converged state of the last solve()."""
        if lam is None:
            lam, _ = self.adjoint()
        W = self.dc.get(9)
        g = np.zeros(self.n_modes)
        for i in range(self.n_modes):
            e = np.zeros(self.n_modes)
            e[i] = 1.0
            dR, dC = self.partials(W, self.amplitudes, e)
            g[i] = dC - np.sum(lam * dR)
        return g

    def node_gradient(self, nodes, lam=None):
        """dC/d(wall-normal displacement) of the given bottom-wall nodes."""
        if lam is None:
            lam, _ = self.adjoint()
        W = self.dc.get(9)
        nx = self.kw["nx"]
        g = np.zeros(len(nodes))
        for k, node in enumerate(nodes):
            e = np.zeros(nx + 1)
            e[node] = 1.0
            dR, dC = self.partials(W, self.amplitudes, e, node_disp=np.zeros(nx + 1))
            g[k] = dC - np.sum(lam * dR)
        return g


class ForceCoefficient(torch.autograd.Function):
    """a (torch float64 [n_modes]) -> C(a); backward by the discrete adjoint."""

    @staticmethod
    def forward(ctx, amplitudes, problem):
        c, _, _ = problem.solve(amplitudes.detach().cpu().numpy())
        ctx.problem = problem
        return torch.tensor(c, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.problem.gradient()
        return grad_out * torch.from_numpy(g), None


class StateCoefficient(torch.autograd.Function):
    """W (torch float64 [n_cells, 5]) -> C(W) on the problem's current mesh at
    frozen boundary masks; backward is the exact dC/dW of the device."""

    @staticmethod
    def forward(ctx, W, problem):
        dc = problem._frozen(problem.amplitudes, W.detach().cpu().numpy())
        ctx.problem = problem
        ctx.dc = dc
        return torch.tensor(problem.coefficient(dc), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):
        dc = ctx.dc
        g = dc.adjoint_objective_gradient(dc.spec.mesh.patch_index("bottom"), ctx.problem.direction, ctx.problem.q)
        return grad_out * torch.from_numpy(g), None
