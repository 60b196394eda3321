"""The benchmark workloads on the device at their stated sizes, against the
oracle (plain FP64 path: natural-order ILU(0), modified Gram-Schmidt GMRES).

* Mach-3 forward step, 120 x 40 (4 032 cells), uniform inflow start, AUSM+
  first order, local time stepping, GMRES(20) + ILU(0), 500 steps.  The
  constant CFL is 3, the tutorial's initialCo (forwardStep/system/
  rhoDSTDict): at a constant CFL of 4 or more the reference's Roe operator
  sends the cells in front of the step to vacuum on the first step, in the
  oracle as on the device (the exact linear solve of that step converges;
  the operator, not the solver, is the limit).
* Transonic hybrid quad/triangle airfoil at full size (205 191 cells, far
  field at 50 chords), HLLC + SAneg, three steps.
* Converged force coefficients: the subsonic bump channel run to a 1e-12
  drop of |R|_2 with the Roe operator on both sides; Cd and Cl of the bump
  wall agree to 1e-8 (RAE2822/system/forces:19-38 is the reference's
  forceCoeffs set-up).
"""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases, meshgen

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_forward_step_500_steps(oracle):
    print("[workload] forward step 120x40, 500 steps", flush=True)
    spec = cases.forward_step_case(120, 40, initial_co=3.0, min_co=3.0, max_co=3.0)
    assert spec.mesh.n_cells == 4032 and spec.mesh.n_internal == 7904
    dc = DeviceCase(spec)
    oc = oracle.OracleCase(spec)
    r1 = dc.iterate(500)
    r0 = oc.iterate(500)
    assert np.all(np.isfinite(r1))
    assert r1[-1] < 0.2 * r1[0]
    err = np.abs(r1 - r0) / r0
    print(f"[workload] forward step: max rel diff of |R|_2 over 500 steps {err.max():.2e}", flush=True)
    assert err.max() <= 1e-9, err.max()
    assert rel(dc.get(9), oc.get(9)) <= 1e-9


def test_hybrid_airfoil_full_size(oracle):
    from rhodst_amd import _capi
    print("[workload] hybrid airfoil full size, HLLC + SAneg, 3 steps", flush=True)
    mesh = meshgen.hybrid_airfoil()
    assert 1.9e5 < mesh.n_cells < 2.2e5
    dm = _capi.DeviceMesh(mesh)
    spec = cases.airfoil_case(mesh=mesh, wall_dist=dm.wall_distance())
    dc = DeviceCase(spec, dm)
    oc = oracle.OracleCase(spec)
    oc.set_sa_partition(dc.mesh.get_int(_capi.M_SA_PARTITION))
    r1 = dc.iterate(3)
    r0 = oc.iterate(3)
    print(f"[workload] airfoil |R|_2 {r1} vs {r0}", flush=True)
    assert rel(r1, r0) <= 1e-9, (r1, r0)
    assert rel(dc.get(9), oc.get(9)) <= 1e-9
    assert rel(dc.get(29), oc.get(29)) <= 1e-7


def test_converged_force_coefficients(oracle):
    print("[workload] bump channel converged Cd/Cl", flush=True)
    spec = cases.bump_case(ilu_order="natural", gram_schmidt="modified")
    dc = DeviceCase(spec)
    oc = oracle.OracleCase(spec)
    r1 = dc.iterate(700)
    r0 = oc.iterate(700)
    assert r1[-1] < 1e-12 * r1[0] and r0[-1] < 1e-12 * r0[0], (r1[-1] / r1[0], r0[-1] / r0[0])
    wall = spec.mesh.patch_index("bottom")
    args = (wall, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 1.0e5, 1.16, 170.0, 2.0)
    fd, fo = dc.forces(*args), oc.forces(*args)
    scale = max(abs(fo[0]), abs(fo[1]))
    print(f"[workload] Cd {fd[0]:.15e} / {fo[0]:.15e}, Cl {fd[1]:.15e} / {fo[1]:.15e}", flush=True)
    assert abs(fd[0] - fo[0]) <= 1e-8 * scale and abs(fd[1] - fo[1]) <= 1e-8 * scale
