"""GPU checks of the discrete adjoint (adjoint.hip): the transpose identity
<J v, w> = <v, J^T w>, the exact J v against central differences of the
residual, the exact dC/dW against central differences of the force
coefficient, the adjoint solve (FP64 and FP32-preconditioned), and the shape
gradient through torch.autograd against central differences of the
re-converged coefficient (target 1e-5 relative)."""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases

pytestmark = pytest.mark.gpu

SCHEMES = ["AUSM+", "AUSM+up", "HLL", "HLLC", "Kurganov"]


def _state_case(scheme):
    spec = cases.box_farfield_case(nx=16, ny=12, scheme=scheme)
    dc = DeviceCase(spec)
    if scheme == "Kurganov":
        # the seeded 5 % perturbation itself: iterated, the box flow relaxes
        # to near uniform, where the central-upwind speeds of neighbouring
        # cells tie (a kink of max/min) on most faces
        dc.flux_residual()
        dc.update_fields()
        dc.flux_residual()
        dc.update_fields()
    else:
        dc.iterate(10)
    return dc


@pytest.mark.parametrize("scheme", SCHEMES)
def test_transpose_identity(scheme):
    print(f"[adjoint transpose] {scheme}", flush=True)
    dc = _state_case(scheme)
    dc.adjoint_linearise()
    rng = np.random.default_rng(4)
    v = rng.normal(size=(dc.n, 5))
    w = rng.normal(size=(dc.n, 5))
    a = np.sum(dc.adjoint_apply(v) * w)
    b = np.sum(v * dc.adjoint_apply(w, transpose=True))
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def _kurganov_near_ties(dc, W, tol):
    from rhodst_amd._capi import G_SF
    gamma, R = 1.4, 287.0
    Cv = R / (gamma - 1)
    own, nei = dc.mesh.get_int(0), dc.mesh.get_int(1)
    Sf = dc.mesh.get_geom(G_SF)[:len(nei)]
    rho = W[:, 0]
    U = W[:, 1:4] / rho[:, None]
    e = W[:, 4] / rho - 0.5 * np.sum(U * U, axis=1)
    c = np.sqrt(gamma * R * e / Cv)
    ms = np.linalg.norm(Sf, axis=1)
    out = np.zeros(len(W), dtype=bool)
    for sgn in (1.0, -1.0):
        a = np.sum(U[own[:len(nei)]] * Sf, axis=1) + sgn * c[own[:len(nei)]] * ms
        b = np.sum(U[nei] * Sf, axis=1) + sgn * c[nei] * ms
        near = np.abs(a - b) <= tol * np.maximum(np.abs(a), np.abs(b))
        out[own[:len(nei)][near]] = True
        out[nei[near]] = True
    return out


@pytest.mark.parametrize("scheme", SCHEMES)
def test_jacobian_against_central_differences(scheme):
    """J v = (R(W + h v) - R(W - h v)) / 2h + O(h^2), v scaled component-wise by W."""
    print(f"[adjoint J fd] {scheme}", flush=True)
    dc = _state_case(scheme)
    dc.adjoint_linearise()
    W = dc.get(9).reshape(-1, 5)
    rng = np.random.default_rng(5)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)
    Jv = dc.adjoint_apply(v)
    h = 1e-6
    dc.put(9, W + h * v)
    Rp = dc.adjoint_residual()
    dc.put(9, W - h * v)
    Rm = dc.adjoint_residual()
    dc.put(9, W)
    fd = (Rp - Rm) / (2 * h)
    scale = np.abs(fd).max(axis=0)
    keep = np.ones(dc.n, dtype=bool)
    if scheme == "Kurganov":
        # the central-upwind speeds a+ = max(phiv_o + c_o|S|, phiv_n + c_n|S|, 0)
        # and a- = min(...) switch branch where the two sides tie; there the
        # residual is not differentiable (J takes the owner branch, a central
        # difference averages the two).  Rows of cells on a face within 1e-4
        # of a tie are left out (a small fraction of this perturbed state).
        keep = ~_kurganov_near_ties(dc, W, 1e-4)
        print("rows left out near Kurganov ties:", int((~keep).sum()), "of", dc.n, flush=True)
        assert keep.mean() > 0.9
    err = np.abs(Jv - fd)[keep].max(axis=0) / scale
    print("relative error per component", err, flush=True)
    if not np.all(err < 1e-6):
        cells = np.argsort(-np.abs(Jv - fd).max(axis=1) / scale.max())[:5]
        print("worst cells", cells, np.abs(Jv - fd)[cells] / scale, flush=True)
        for hh in (1e-5, 1e-7):
            dc.put(9, W + hh * v)
            Rp = dc.adjoint_residual()
            dc.put(9, W - hh * v)
            Rm = dc.adjoint_residual()
            dc.put(9, W)
            e2 = np.abs(Jv - (Rp - Rm) / (2 * hh)).max(axis=0) / scale
            print(f"h {hh:g}: relative error per component", e2, flush=True)
    assert np.all(err < 1e-6), err


def test_objective_gradient_against_central_differences():
    print("[adjoint dC/dW fd]", flush=True)
    spec = cases.bump_case()
    dc = DeviceCase(spec)
    dc.iterate(200)
    patch = spec.mesh.patch_index("bottom")
    d = np.array([1.0, 0.0, 0.0])
    q = 0.5
    g = dc.adjoint_objective_gradient(patch, d, q)
    W = dc.get(9).reshape(-1, 5)
    rng = np.random.default_rng(6)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)

    def C(Wx):
        dc.put(9, Wx)
        dc.update_fields()
        dc.update_fields()
        out = dc.forces(patch, d, (0, 1, 0), 1e5, 1.0, 1.0, 1.0)
        return out[2] / q

    h = 1e-7
    fd = (C(W + h * v) - C(W - h * v)) / (2 * h)
    an = np.sum(g * v)
    assert abs(an - fd) <= 1e-6 * abs(fd), (an, fd)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_solve_residual(precision):
    print(f"[adjoint solve] {precision}", flush=True)
    from rhodst_amd.adjoint import BumpProblem
    prob = BumpProblem(drop=1e-10)
    prob.solve(np.array([0.02, 0.01, -0.005, 0.004]))
    lam, perf = prob.adjoint(precision=precision)
    dc = prob.dc
    g = dc.adjoint_objective_gradient(dc.spec.mesh.patch_index("bottom"), prob.direction, prob.q)
    r = dc.adjoint_apply(lam, transpose=True) - g
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(g), (np.linalg.norm(r), perf.n_iterations)
    if precision == "mixed":
        lam64, _ = prob.adjoint(precision="fp64")
        assert np.abs(lam - lam64).max() <= 1e-8 * np.abs(lam64).max()


def test_shape_gradient_through_autograd():
    """dCd/da of the Hicks-Henne amplitudes: torch.autograd backward through
    the adjoint against central differences of the re-converged Cd."""
    import torch
    from rhodst_amd.adjoint import BumpProblem, ForceCoefficient
    print("[adjoint shape gradient]", flush=True)
    a0 = np.array([0.02, 0.01, -0.005, 0.004])
    prob = BumpProblem(drop=1e-13, max_steps=6000)
    a = torch.tensor(a0, dtype=torch.float64, requires_grad=True)
    C = ForceCoefficient.apply(a, prob)
    print("Cd", float(C), flush=True)
    C.backward()
    g_adj = a.grad.numpy().copy()
    print("adjoint gradient", g_adj, flush=True)
    H = 1e-6
    fd = np.zeros(len(a0))
    for i in range(len(a0)):
        e = np.zeros(len(a0))
        e[i] = H
        cp, sp, _ = BumpProblem(drop=1e-13, max_steps=6000).solve(a0 + e)
        cm, sm, _ = BumpProblem(drop=1e-13, max_steps=6000).solve(a0 - e)
        fd[i] = (cp - cm) / (2 * H)
        print(f"mode {i}: fd {fd[i]:.10e} ({sp}, {sm} steps)", flush=True)
    print("adjoint", g_adj, "fd", fd, flush=True)
    rel = np.abs(g_adj - fd) / np.abs(fd).max()
    assert np.all(rel < 1e-5), rel


def test_state_gradient_through_autograd():
    import torch
    from rhodst_amd.adjoint import BumpProblem, StateCoefficient
    print("[adjoint state gradient]", flush=True)
    prob = BumpProblem(drop=1e-8)
    prob.solve(np.array([0.02, 0.01, -0.005, 0.004]))
    W = torch.tensor(prob.dc.get(9).reshape(-1, 5), dtype=torch.float64, requires_grad=True)
    C = StateCoefficient.apply(W, prob)
    C.backward()
    g = prob.dc.adjoint_objective_gradient(prob.dc.spec.mesh.patch_index("bottom"), prob.direction, prob.q)
    assert np.allclose(W.grad.numpy(), g, rtol=1e-12, atol=0)
