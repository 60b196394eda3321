"""GPU checks of the viscous discrete adjoint (adjoint.hip, the
second-neighbour linearisation of calcFields' viscous terms): the residual of
the linearised chain against the face loop's, the transpose identity, J v
against central differences (step 1e-7) of the device residual and of the
oracle's residual, dC/dW of the pressure + viscous force against central
differences, the adjoint solve (FP64 and FP32 transposed ILU), the frozen
eddy viscosity of a Spalart-Allmaras case, and the shape gradient of a
laminar channel bump against central differences of the re-converged Cd."""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases

pytestmark = pytest.mark.gpu


def _box(scheme="HLLC", steps=20, **kw):
    spec = cases.box_viscous_case(nx=16, ny=12, scheme=scheme, **kw)
    dc = DeviceCase(spec)
    dc.iterate(steps)
    # boundary values consistent with W (the far field reads the previous p, T)
    dc.update_fields()
    dc.update_fields()
    return spec, dc


def _fd(dc, W, v, h):
    dc.put(9, W + h * v)
    Rp = dc.adjoint_residual()
    dc.put(9, W - h * v)
    Rm = dc.adjoint_residual()
    dc.put(9, W)
    return (Rp - Rm) / (2 * h)


@pytest.mark.parametrize("scheme", ["HLLC", "AUSM+up"])
def test_chain_residual_matches_face_loop(scheme):
    print(f"[viscous adjoint residual] {scheme}", flush=True)
    spec, dc = _box(scheme)
    # the boundary conditions' masks read the stored mass flux; the face loop
    # rewrites it, so the chain is evaluated with the flux the boundary values
    # of the face loop's call were formed with
    dc.flux_residual()
    phi = dc.get(12)
    dc.update_fields()
    dc.update_fields()
    dc.flux_residual()
    R = dc.get(8).reshape(-1, 5)
    dc.put(12, phi)
    Ra = dc.adjoint_residual()
    scale = np.abs(R).max(axis=0)
    err = np.abs(R - Ra).max(axis=0) / scale
    print("relative difference per component", err, flush=True)
    assert np.all(err < 1e-11), err


@pytest.mark.parametrize("scheme", ["HLLC", "AUSM+up"])
def test_transpose_identity(scheme):
    print(f"[viscous adjoint transpose] {scheme}", flush=True)
    spec, dc = _box(scheme)
    dc.adjoint_linearise()
    rng = np.random.default_rng(4)
    v = rng.normal(size=(dc.n, 5))
    w = rng.normal(size=(dc.n, 5))
    a = np.sum(dc.adjoint_apply(v) * w)
    b = np.sum(v * dc.adjoint_apply(w, transpose=True))
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)), (a, b)


@pytest.mark.parametrize("scheme,limited", [("HLLC", False), ("AUSM+up", False), ("HLLC", True)])
def test_jacobian_against_central_differences(scheme, limited):
    """J v = (R(W + h v) - R(W - h v)) / 2h + O(h^2) with h = 1e-7, v scaled
    component-wise by W; limited: the non-orthogonal corrections limited
    (laplacian 0.5, snGrad 0.33)."""
    print(f"[viscous adjoint J fd] {scheme} limited={limited}", flush=True)
    kw = dict(laplacian=("limited", 0.5), sngrad=("limited", 0.33)) if limited else {}
    spec, dc = _box(scheme, **kw)
    dc.adjoint_linearise()
    W = dc.get(9).reshape(-1, 5)
    rng = np.random.default_rng(5)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)
    Jv = dc.adjoint_apply(v)
    fd = _fd(dc, W, v, 1e-7)
    scale = np.abs(fd).max(axis=0)
    err = np.abs(Jv - fd).max(axis=0) / scale
    print("relative error per component", err, flush=True)
    assert np.all(err < 1e-6), err


def test_jacobian_against_oracle_differences(oracle):
    """J v against central differences (h = 1e-7) of the oracle's residual:
    the oracle takes W and the stored mass flux, updates the fields twice
    (boundary values consistent with W) and forms R_ by its calcFields."""
    print("[viscous adjoint J vs oracle fd]", flush=True)
    spec, dc = _box("HLLC")
    dc.adjoint_linearise()
    W = dc.get(9).reshape(-1, 5)
    phi = dc.get(12)
    oc = oracle.OracleCase(spec)

    def R_or(Wx):
        oc.put(9, Wx)
        oc.put(12, phi)
        oc.update_fields()
        oc.update_fields()
        oc.flux_residual()
        return oc.get(8).reshape(-1, 5)

    rng = np.random.default_rng(7)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)
    Jv = dc.adjoint_apply(v)
    h = 1e-7
    fd = (R_or(W + h * v) - R_or(W - h * v)) / (2 * h)
    scale = np.abs(fd).max(axis=0)
    err = np.abs(Jv - fd).max(axis=0) / scale
    print("relative error per component", err, flush=True)
    assert np.all(err < 1e-6), err
    # and the oracle's residual at W itself against the chain's
    R0 = R_or(W)
    Ra = dc.adjoint_residual()
    assert np.all(np.abs(R0 - Ra).max(axis=0) <= 1e-11 * np.abs(R0).max(axis=0))


def test_objective_gradient_against_central_differences():
    """dC/dW of the wall's pressure + viscous force against central
    differences of rdst_forces (the boundary gradient re-formed at each W)."""
    print("[viscous adjoint dC/dW fd]", flush=True)
    spec, dc = _box("HLLC")
    patch = spec.mesh.patch_index("bottom")
    d = np.array([1.0, 0.0, 0.0])
    q = 0.5
    g = dc.adjoint_objective_gradient(patch, d, q)
    W = dc.get(9).reshape(-1, 5)
    phi = dc.get(12)
    rng = np.random.default_rng(6)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)

    def C(Wx):
        dc.put(9, Wx)
        dc.put(12, phi)
        dc.update_fields()
        dc.update_fields()
        dc.flux_residual()
        out = dc.forces(patch, d, (0, 1, 0), 1e5, 1.0, 1.0, 1.0)
        return out[2] / q

    h = 1e-7
    fd = (C(W + h * v) - C(W - h * v)) / (2 * h)
    dc.put(9, W)
    dc.put(12, phi)
    an = np.sum(g * v)
    print("analytic", an, "fd", fd, flush=True)
    assert abs(an - fd) <= 1e-6 * abs(fd), (an, fd)
    # the viscous part is not negligible here
    out = C(W)
    assert abs(dc.forces(patch, d, (0, 1, 0), 1e5, 1.0, 1.0, 1.0)[5]) > 1e-3 * abs(out * q)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_adjoint_solve_residual(precision):
    print(f"[viscous adjoint solve] {precision}", flush=True)
    spec, dc = _box("HLLC", steps=60)
    patch = spec.mesh.patch_index("bottom")
    d = np.array([1.0, 0.0, 0.0])
    lam, perf = dc.adjoint_solve(patch, d, 0.5, precision=precision, n_directions=30, max_cycles=200,
                                 tolerance=1e-13, rel_tol=0.0)
    g = dc.adjoint_objective_gradient(patch, d, 0.5)
    r = dc.adjoint_apply(lam, transpose=True) - g
    print("cycles", perf.n_iterations, "|r|/|g|", np.linalg.norm(r) / np.linalg.norm(g), flush=True)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(g)
    if precision == "mixed":
        lam64, _ = dc.adjoint_solve(patch, d, 0.5, precision="fp64", n_directions=30, max_cycles=200,
                                    tolerance=1e-13, rel_tol=0.0)
        assert np.abs(lam - lam64).max() <= 1e-8 * np.abs(lam64).max()


def test_frozen_eddy_viscosity_sa():
    """Spalart-Allmaras box: mut and alphat frozen at their stored values; J v
    against central differences of the residual with them frozen, and the
    transpose identity."""
    print("[viscous adjoint SA frozen]", flush=True)
    spec = cases.box_sa_case(model="SAneg")
    dc = DeviceCase(spec)
    dc.iterate(10)
    dc.update_fields()
    dc.update_fields()
    mut = dc.get(30)
    assert np.abs(mut).max() > 0
    dc.adjoint_linearise()
    W = dc.get(9).reshape(-1, 5)
    rng = np.random.default_rng(8)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)
    Jv = dc.adjoint_apply(v)
    fd = _fd(dc, W, v, 1e-7)
    err = np.abs(Jv - fd).max(axis=0) / np.abs(fd).max(axis=0)
    print("relative error per component", err, flush=True)
    assert np.all(err < 1e-6), err
    w = rng.normal(size=W.shape)
    a = np.sum(Jv * w)
    b = np.sum(v * dc.adjoint_apply(w, transpose=True))
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b))


def test_viscous_shape_gradient():
    """dCd/da of the Hicks-Henne amplitudes of the laminar channel bump (Re
    5000): the adjoint against central differences of the re-converged Cd."""
    from rhodst_amd.adjoint import BumpProblem
    print("[viscous adjoint shape gradient]", flush=True)
    a0 = np.array([0.02, 0.01, -0.005, 0.004])
    kw = dict(viscous=True, drop=1e-13, max_steps=8000)
    prob = BumpProblem(**kw)
    c0, steps, drop = prob.solve(a0)
    print("Cd", c0, "steps", steps, flush=True)
    g_adj = prob.gradient()
    H = 1e-6
    fd = np.zeros(len(a0))
    for i in range(len(a0)):
        e = np.zeros(len(a0))
        e[i] = H
        cp, _, _ = BumpProblem(**kw).solve(a0 + e)
        cm, _, _ = BumpProblem(**kw).solve(a0 - e)
        fd[i] = (cp - cm) / (2 * H)
        print(f"mode {i}: adjoint {g_adj[i]:.10e} fd {fd[i]:.10e}", flush=True)
    rel = np.abs(g_adj - fd) / np.abs(fd).max()
    assert np.all(rel < 1e-5), rel


def test_airfoil_node_gradient_batching():
    """Wall-node gradient of a small Spalart-Allmaras hybrid airfoil (10 k
    cells, frozen eddy viscosity): nodes 8 apart displaced together, each
    changed cell credited to the nearest displaced node, against each node
    displaced alone."""
    from rhodst_amd.adjoint import AirfoilProblem
    print("[viscous adjoint airfoil node gradient]", flush=True)
    prob = AirfoilProblem(mesh_kw=dict(n_surf=80, n_layers=16, first_cell=2e-5, far=8.0))
    cd, steps, _ = prob.solve(max_steps=200, log=None)
    lam, perf = prob.adjoint(precision="fp64", max_cycles=40, tolerance=1e-8)
    print("Cd", cd, "adjoint cycles", perf.n_iterations, flush=True)
    g = prob.node_gradient(lam)
    picks = [0, 17, 40, 63, 79]
    gs = prob.node_gradient(lam, nodes=picks)
    print("batched", g[picks], "alone", gs, flush=True)
    assert np.all(np.isfinite(g)) and np.abs(g).max() > 0
    # the two differ by the rounding of the differences (h = 1e-7)
    assert np.abs(g[picks] - gs).max() <= 1e-7 * np.abs(g).max()


def test_airfoil_full_size_jacobian_against_oracle(oracle):
    """The adjoint benchmark's workload at full size (hybrid airfoil, 205 k
    cells, HLLC + SAneg with the eddy viscosity frozen): J v against central
    differences (h = 1e-7) of the oracle's residual, the oracle holding the
    device's eddy viscosity and stored mass flux."""
    from rhodst_amd.adjoint import AirfoilProblem
    print("[viscous adjoint airfoil 205k vs oracle]", flush=True)
    prob = AirfoilProblem(precision="mixed", ilu_order="partitioned", gram_schmidt="classical")
    prob.solve(max_steps=20, chunk=20, log=None)
    dc = prob.dc
    dc.update_fields()
    dc.update_fields()
    dc.adjoint_linearise()
    W = dc.get(9).reshape(-1, 5)
    phi = dc.get(12)
    mut, alphat, bmut = dc.get(30), dc.get(31), dc.get(34)
    oc = oracle.OracleCase(dc.spec)
    print("oracle case built", flush=True)

    def R_or(Wx):
        oc.put(9, Wx)
        oc.put(12, phi)
        oc.update_fields()
        oc.update_fields()
        oc.put(30, mut)
        oc.put(31, alphat)
        oc.put(34, bmut)
        oc.flux_residual()
        return oc.get(8).reshape(-1, 5)

    rng = np.random.default_rng(9)
    v = rng.uniform(-1, 1, size=W.shape) * np.abs(W).max(axis=0)
    Jv = dc.adjoint_apply(v)
    h = 1e-7
    fd = (R_or(W + h * v) - R_or(W - h * v)) / (2 * h)
    err = np.abs(Jv - fd).max(axis=0) / np.abs(fd).max(axis=0)
    print("relative error per component", err, flush=True)
    assert np.all(err < 1e-6), err


def test_defect_correction_solve():
    """inner_operator 1: FP64 residuals against the exact viscous J^T, FP32
    inner cycles on the face-neighbour part of J^T shifted by 1/dtau; the
    solution satisfies the exact system."""
    print("[viscous adjoint defect correction]", flush=True)
    spec, dc = _box("HLLC", steps=60)
    patch = spec.mesh.patch_index("bottom")
    d = np.array([1.0, 0.0, 0.0])
    lam, perf = dc.adjoint_solve(patch, d, 0.5, precision="mixed", n_directions=30, max_cycles=2000,
                                 tolerance=0.0, rel_tol=1e-12, precond_shift=0.3, inner_operator=1)
    g = dc.adjoint_objective_gradient(patch, d, 0.5)
    r = dc.adjoint_apply(lam, transpose=True) - g
    print("cycles", perf.n_iterations, "|r|/|g|", np.linalg.norm(r) / np.linalg.norm(g), flush=True)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(g)
    lam64, _ = dc.adjoint_solve(patch, d, 0.5, precision="fp64", n_directions=30, max_cycles=200,
                                tolerance=1e-13, rel_tol=0.0)
    assert np.abs(lam - lam64).max() <= 1e-8 * np.abs(lam64).max()


@pytest.mark.parametrize("viscous", [False, True])
def test_objective_cells_sum_to_force(viscous):
    """rdst_adjoint_objective_cells: the cells' shares sum to the force
    coefficient of rdst_forces (pressure, and the viscous part when viscous)
    at a state whose boundary values and boundary gradient are consistent."""
    print(f"[objective cells] viscous={viscous}", flush=True)
    if viscous:
        spec, dc = _box("HLLC")
        patch = spec.mesh.patch_index("bottom")
    else:
        spec = cases.bump_case()
        dc = DeviceCase(spec)
        dc.iterate(50)
        dc.update_fields()
        dc.update_fields()
        patch = spec.mesh.patch_index("bottom")
    dc.flux_residual()
    phi = dc.get(12)
    dc.update_fields()
    dc.update_fields()
    dc.flux_residual()
    dc.put(12, phi)
    d = np.array([1.0, 0.0, 0.0])
    q, p_ref = 0.5, 1e5
    o = dc.adjoint_objective_cells(patch, d, q, p_ref)
    F = dc.forces(patch, d, (0, 1, 0), p_ref, 1.0, 1.0, 1.0)[2:5]
    c = float(F @ d) / q
    print("sum of shares", o.sum(), "forces", c, flush=True)
    assert abs(o.sum() - c) <= 1e-11 * max(abs(c), 1e-30)
    # only the cells on the patch hold a share
    assert 0 < np.count_nonzero(o) <= spec.mesh.patches[patch][3]
