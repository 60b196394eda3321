"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle.

Bars: face fluxes and R_ within 1e-12 relative (the plain FP64 path),
assembled block-CSR within 1e-12 of the oracle's LDU blocks through the
face->slot maps, and whole Single::iterate sequences within rounding drift.
"""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases

pytestmark = pytest.mark.gpu

SCHEMES = ["AUSM+", "AUSM+up", "HLL", "HLLC", "HLLCP", "Kurganov"]


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def make_spec(kind, scheme, **kw):
    if kind == "viscous":
        return cases.box_viscous_case(scheme=scheme, **kw)
    if kind == "step":
        return cases.forward_step_case(nx=60, ny=20, scheme=scheme, perturb=0.02, initial_co=0.5, min_co=0.1,
                                       max_co=5.0, **kw)
    return cases.box_farfield_case(nx=24, ny=16, scheme=scheme, jitter=0.2, perturb=0.05, **kw)


@pytest.mark.parametrize("kind", ["step", "box", "viscous"])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_flux_residual_parity(oracle, kind, scheme):
    print(f"[flux parity] {kind} {scheme}", flush=True)
    spec = make_spec(kind, scheme)
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    oc.flux_residual()
    dc.flux_residual()
    assert rel(dc.get(8), oc.get(8)) <= 1e-12
    for fid in (12, 13, 14, 15):
        assert rel(dc.get(fid), oc.get(fid)) <= 1e-12, fid


@pytest.mark.parametrize("kind", ["step", "box", "viscous"])
@pytest.mark.parametrize("scheme", ["AUSM+", "HLLC", "Kurganov"])
def test_assembly_parity(oracle, kind, scheme):
    print(f"[assembly parity] {kind} {scheme}", flush=True)
    spec = make_spec(kind, scheme)
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    for c in (oc, dc):
        c.flux_residual()
        c.jacobian_assemble()
    _rp, _cols, vals = dc.bcsr()
    m = dc.mesh
    D, U, L = oc.matrix(0).reshape(-1, 5, 5), oc.matrix(1).reshape(-1, 5, 5), oc.matrix(2).reshape(-1, 5, 5)
    scale = np.abs(D).max()
    assert np.abs(vals[m.get_int(6)] - D).max() <= 1e-12 * scale
    assert np.abs(vals[m.get_int(4)] - U).max() <= 1e-12 * scale
    assert np.abs(vals[m.get_int(5)] - L).max() <= 1e-12 * scale
    assert rel(dc.get(25), oc.matrix(3)) <= 1e-12


@pytest.mark.parametrize("solver", ["GMRES", "BiCGStab"])
@pytest.mark.parametrize("dependent", ["W", "deltaW"])
def test_block_solve_parity(oracle, solver, dependent):
    print(f"[solve parity] {solver} {dependent}", flush=True)
    spec = make_spec("box", "HLLC", solver=solver, dependent=dependent, max_iter=3)
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    W0 = oc.get(9)
    for c in (oc, dc):
        c.flux_residual()
        c.jacobian_assemble()
    p0, p1 = oc.block_solve(), dc.block_solve()
    assert p0.n_iterations == p1.n_iterations
    assert np.allclose(list(p1.initial_residual), list(p0.initial_residual), rtol=1e-10, atol=1e-300)
    dW0, dW1 = oc.get(9) - W0, dc.get(9) - W0
    assert rel(dW1, dW0) <= 1e-9


def test_update_fields_and_courant_parity(oracle):
    print("[update parity] box far-field", flush=True)
    spec = make_spec("box", "HLLC")
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    for c in (oc, dc):
        c.flux_residual()
        c.jacobian_assemble()
        c.block_solve()
    dc.put(9, oc.get(9))  # same W into both updates
    oc.update_fields()
    dc.update_fields()
    for fid in (0, 1, 2, 3, 4, 5, 6, 7, 16, 17, 18, 19, 20, 21, 22, 23):
        assert rel(dc.get(fid), oc.get(fid)) <= 1e-13, fid
    r0, r1 = oc.set_courant(), dc.set_courant()
    assert abs(r1 - r0) <= 1e-13 * r0
    assert rel(dc.get(10), oc.get(10)) <= 1e-13
    assert rel(dc.get(11), oc.get(11)) <= 1e-13


@pytest.mark.parametrize("kind,scheme", [("step", "AUSM+"), ("box", "HLLC"), ("box", "AUSM+up"), ("step", "Kurganov"),
                                         ("viscous", "HLLC"), ("viscous", "AUSM+up")])
def test_iterate_parity(oracle, kind, scheme):
    print(f"[iterate parity] {kind} {scheme}", flush=True)
    spec = make_spec(kind, scheme)
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    n = 20
    r0, p0 = oc.iterate(n, with_perf=True)
    r1, p1 = dc.iterate(n, with_perf=True)
    assert np.all(np.isfinite(r1))
    assert rel(r1, r0) <= 1e-9, (r1, r0)
    assert rel(dc.get(9), oc.get(9)) <= 1e-9


def test_mixed_precision_tracks_fp64(oracle):
    """FP32 Jacobian/ILU/Krylov inside FP64 defect correction: the nonlinear
    residual history follows the FP64 path (same converged state)."""
    print("[mixed] box far-field HLLC", flush=True)
    spec64 = make_spec("box", "HLLC", max_iter=2)
    spec32 = make_spec("box", "HLLC", max_iter=2, precision="mixed")
    d64, d32 = DeviceCase(spec64), DeviceCase(spec32)
    r64 = d64.iterate(40)
    r32 = d32.iterate(40)
    assert np.all(np.isfinite(r32))
    assert r32[-1] / r32[0] < 0.5 * r64[0] / r64[0]
    assert abs(np.log10(r32[-1] / r32[0]) - np.log10(r64[-1] / r64[0])) < 0.5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_partitioned_ilu_parity(oracle, precision):
    """Block-Jacobi ILU(0) over the device mesh's RCB partitions: the FP64
    path matches the oracle's ILU(0) of the partition-diagonal blocks; the
    mixed path follows it to rounding drift of the FP32 solve."""
    from rhodst_amd import _capi
    print(f"[partitioned ILU] {precision}", flush=True)
    spec = cases.box_farfield_case(nx=60, ny=50, ilu_order="partitioned", max_iter=2, precision=precision)
    dc = DeviceCase(spec)
    part = dc.mesh.get_int(_capi.M_PARTITION)
    assert part.max() + 1 == dc.mesh.info["n_parts"] > 1
    ospec = cases.box_farfield_case(nx=60, ny=50, ilu_order="partitioned", max_iter=2)
    oc = oracle.OracleCase(ospec)
    oc.set_partition(part)
    r1 = dc.iterate(10)
    r0 = oc.iterate(10)
    tol = 1e-9 if precision == "fp64" else 1e-3
    assert rel(r1, r0) <= tol, (r1, r0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_classical_gram_schmidt_parity(oracle, precision):
    """Classical Gram-Schmidt GMRES (fused multi-dot) against the oracle's
    classical Gram-Schmidt, with the partitioned ILU and the sliced-ELL SpMV."""
    from rhodst_amd import _capi
    print(f"[classical GS] {precision}", flush=True)
    kw = dict(ilu_order="partitioned", max_iter=2, gram_schmidt="classical")
    spec = cases.box_farfield_case(nx=60, ny=50, precision=precision, **kw)
    dc = DeviceCase(spec)
    oc = oracle.OracleCase(cases.box_farfield_case(nx=60, ny=50, **kw))
    oc.set_partition(dc.mesh.get_int(_capi.M_PARTITION))
    r1 = dc.iterate(10)
    r0 = oc.iterate(10)
    tol = 1e-9 if precision == "fp64" else 1e-3
    assert rel(r1, r0) <= tol, (r1, r0)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_tangential_farfield_mass_flux_is_zero(oracle, scheme):
    """On boundary faces whose state is tangential (U_b . Sf = 0 exactly) the
    reference's flux gives phi = 0 exactly, and the far-field regime masks
    (P.C:174-191) then select subsonic outflow; a rounding-level phi of the
    wrong sign would flip them to inflow."""
    print(f"[tangential phi] {scheme}", flush=True)
    spec = cases.box_viscous_case(scheme=scheme)
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    oc.flux_residual()
    dc.flux_residual()
    m = spec.mesh
    top = [p for p in m.patches if p[0] == "top"][0]
    sl = slice(top[2], top[2] + top[3])
    assert np.all(oc.get(12)[sl] == 0.0)
    assert np.all(dc.get(12)[sl] == 0.0), dc.get(12)[sl]


@pytest.mark.parametrize("kind", ["step", "box"])
@pytest.mark.parametrize("scheme", SCHEMES)
def test_flux_residual_parity_ospre(oracle, kind, scheme):
    """Second-order limited reconstruction (reconstruct(rho|T) OSPRE,
    reconstruct(U) OSPREV): face fluxes and R_ against the oracle."""
    print(f"[flux parity OSPRE] {kind} {scheme}", flush=True)
    spec = make_spec(kind, scheme, reconstruction="OSPRE")
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    oc.flux_residual()
    dc.flux_residual()
    assert rel(dc.get(8), oc.get(8)) <= 1e-12
    for fid in (12, 13, 14, 15):
        assert rel(dc.get(fid), oc.get(fid)) <= 1e-12, fid


@pytest.mark.parametrize("kind,scheme", [("step", "AUSM+"), ("box", "HLLC"), ("viscous", "AUSM+up")])
def test_iterate_parity_ospre(oracle, kind, scheme):
    print(f"[iterate parity OSPRE] {kind} {scheme}", flush=True)
    spec = make_spec(kind, scheme, reconstruction="OSPRE")
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    r0 = oc.iterate(15)
    r1 = dc.iterate(15)
    assert np.all(np.isfinite(r1))
    assert rel(r1, r0) <= 1e-9, (r1, r0)
