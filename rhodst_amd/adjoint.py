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


class AirfoilProblem:
    """Cd of the viscous hybrid airfoil (cases.airfoil_case: M 0.73, AoA 2.3
    deg, Re 6.5e6 on the unit chord, Spalart-Allmaras) as a function of the
    displacements of its wall nodes along their wall-layer lines (the
    surface normal at the wall; meshgen.hybrid_airfoil wall_disp), at a shape
    seeded with smooth Hicks-Henne bumps of ``seed_amplitude`` chords on each
    surface.  The design nodes are the wall loop's surface points, lower and
    upper, without the lower trailing-edge corner (n_surf of them).

    The adjoint freezes the eddy viscosity of the converged flow (mut, alphat
    and mut on the boundary): the gradient is exact for the residual with
    that eddy viscosity held, and leaves out the turbulence model's response
    to the shape.  The partial derivatives at the frozen state are central
    differences of each cell's residual and objective share on displaced
    meshes; nodes ``stride`` apart along the loop are displaced together
    (their residual supports -- the cells of the two wall-layer columns
    beside the node and two face neighbours out -- do not overlap at a
    stride of 8) and each changed cell is credited to the nearest displaced
    node, so the whole gradient costs 2 stride mesh builds."""

    def __init__(self, mesh_kw=None, turbulence="SAneg", seed_amplitude=1e-4, n_seed_modes=4, seed=3,
                 fd_step=1e-7, precond_shift=0.3, **case_kw):
        from . import meshgen
        self.precond_shift = precond_shift
        self.meshgen = meshgen
        self.mesh_kw = dict(mesh_kw or {})
        n_surf = self.mesh_kw.get("n_surf", 400)
        loop, kind = meshgen.airfoil_loop(n_surf, self.mesh_kw.get("thickness", 0.12))
        self.loop, self.kind = loop, kind
        lower = np.nonzero(kind == 0)[0]
        te_lower = lower[np.argmax(loop[lower, 0] - 1e-9 * loop[lower, 1])]
        self.design = np.array([i for i in np.nonzero(kind < 2)[0] if i != te_lower])
        rng = np.random.default_rng(seed)
        self.base_disp = np.zeros(len(loop))
        for k in (0, 1):
            pts = np.nonzero(kind == k)[0]
            amp = seed_amplitude * rng.uniform(-1.0, 1.0, n_seed_modes)
            self.base_disp[pts] = amp @ meshgen.hicks_henne(loop[pts, 0], n_seed_modes)
        self.turbulence = turbulence
        self.case_kw = case_kw
        self.fd_step = fd_step
        gamma, R, Tinf, pinf = 1.4, 287.0, 288.15, 101325.0
        mach = case_kw.get("mach", 0.73)
        a = np.deg2rad(case_kw.get("aoa_deg", 2.3))
        self.U = mach * np.sqrt(gamma * R * Tinf)
        self.rho = pinf / (R * Tinf)
        self.p_ref = pinf
        self.a_ref = 1.0 * self.mesh_kw.get("depth", 0.1)
        self.q = 0.5 * self.rho * self.U ** 2 * self.a_ref
        self.drag = np.array([np.cos(a), np.sin(a), 0.0])
        self.lift = np.array([-np.sin(a), np.cos(a), 0.0])
        self.dc = None

    def mesh(self, d=None):
        """The mesh at the seeded shape, displaced by d at the design nodes."""
        if getattr(self, "_undisplaced", None) is None:
            self._undisplaced = self.meshgen.hybrid_airfoil(**self.mesh_kw)
        disp = self.base_disp.copy()
        if d is not None:
            disp[self.design] += d
        return self.meshgen.displace_airfoil_wall(self._undisplaced, disp, self.mesh_kw.get("n_layers", 40))

    def coefficient(self, dc):
        patch = dc.spec.mesh.patch_index("airfoil")
        return float(dc.forces(patch, self.drag, self.lift, self.p_ref, self.rho, self.U, self.a_ref)[0])

    def solve(self, max_steps=3000, drop=1e-8, chunk=100, log=print):
        """Converge the flow on the seeded shape; keep the state and its eddy viscosity."""
        from ._capi import DeviceMesh
        mesh = self.mesh()
        dm = DeviceMesh(mesh)
        wd = dm.wall_distance() if self.turbulence else None
        spec = cases.airfoil_case(mesh=mesh, turbulence=self.turbulence, wall_dist=wd, **self.case_kw)
        dc = DeviceCase(spec, dm)
        r0, steps, hist = None, 0, []
        while steps < max_steps:
            res = dc.iterate(chunk)
            steps += chunk
            if r0 is None:
                r0 = res[0]
            if not np.all(np.isfinite(res)):
                raise RuntimeError("airfoil case diverged")
            hist.append((steps, res[-1] / r0))
            if log:
                log(f"[airfoil] step {steps} drop {res[-1] / r0:.3e}")
            if res[-1] <= drop * r0:
                break
        self.dc, self.mesh0 = dc, mesh
        self.W = dc.get(9).reshape(-1, 5)
        self.frozen_turb = (dc.get(30), dc.get(31), dc.get(34)) if self.turbulence else None
        return self.coefficient(dc), steps, hist

    def adjoint(self, **solver):
        """J^T lambda = dCd/dW at the converged state.  Default: a defect
        correction whose FP64 residuals use the exact viscous J^T and whose
        FP32 inner GMRES(30) cycles solve with the face-neighbour part of J^T
        plus precond_shift x 1/dtau on the diagonal (that matrix also gives the
        ILU(0)): GMRES on the exact J^T itself, preconditioned by one ILU, stalls
        on this mesh's long upstream transport of the adjoint."""
        dc = self.dc
        opts = dict(precision="mixed", n_directions=30, max_cycles=5000, tolerance=0.0, rel_tol=1e-8,
                    precond_shift=self.precond_shift, inner_operator=1)
        opts.update(solver)
        return dc.adjoint_solve(dc.spec.mesh.patch_index("airfoil"), self.drag, self.q, **opts)

    def frozen(self, d, W):
        """A case on the mesh displaced by d carrying W and the frozen eddy
        viscosity, its boundary values and boundary gradient formed from W."""
        spec = cases.airfoil_case(mesh=self.mesh(d), turbulence=None, **self.case_kw)
        dc = DeviceCase(spec)
        dc.put(9, W)
        if self.frozen_turb is not None:
            mut, alphat, bmut = self.frozen_turb
            dc.put(30, mut)
            dc.put(31, alphat)
            dc.put(34, bmut)
        dc.update_fields()
        dc.flux_residual()
        dc.update_fields()
        dc.flux_residual()
        dc.update_fields()
        return dc

    def _shares(self, dc):
        patch = dc.spec.mesh.patch_index("airfoil")
        return dc.adjoint_residual(), dc.adjoint_objective_cells(patch, self.drag, self.q, self.p_ref)

    def node_gradient(self, lam, stride=8, nodes=None):
        """dCd/d(displacement) of the design nodes (all, or the indices
        ``nodes`` into the design list, each then displaced alone)."""
        h = self.fd_step
        nd = len(self.design)
        cen = self.meshgen.cell_centres_approx(self.mesh0)[:, :2]
        wall = self.loop[self.design]
        g = np.zeros(nd)
        groups = [np.array([k]) for k in nodes] if nodes is not None else \
            [np.arange(col, nd, stride) for col in range(stride)]
        for sel in groups:
            e = np.zeros(nd)
            e[sel] = h
            Rp, op = self._shares(self.frozen(e, self.W))
            Rm, om = self._shares(self.frozen(-e, self.W))
            contrib = ((op - om) - np.sum(lam * (Rp - Rm), axis=1)) / (2 * h)
            hit = np.nonzero(np.any(Rp != Rm, axis=1) | (op != om))[0]
            d2 = np.sum((cen[hit, None, :] - wall[None, sel, :]) ** 2, axis=2)
            np.add.at(g, sel[np.argmin(d2, axis=1)], contrib[hit])
        return g if nodes is None else g[np.asarray(nodes)]


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
