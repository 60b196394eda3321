"""The mixed-precision path's converged results against the FP64 oracle:
the channel bump (inviscid, HLLC, the adjoint's workload) run to a 1e-12
residual drop on the device in mixed precision (FP32 Jacobian, ILU and
Krylov inside the FP64 defect correction) and in the oracle in FP64; the
state and the drag coefficient are stated and bounded (Cd within 1e-6 of
FP64 is the mixed path's target; converged, the defect correction reaches
the FP64 fixed point itself)."""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases

pytestmark = pytest.mark.gpu


def _converge(run, drop=1e-12, chunk=100, max_steps=6000):
    r0, steps = None, 0
    while steps < max_steps:
        res = run(chunk)
        steps += chunk
        if r0 is None:
            r0 = res[0]
        assert np.all(np.isfinite(res))
        if res[-1] <= drop * r0:
            return steps, res[-1] / r0
    raise AssertionError(f"no {drop:g} drop in {max_steps} steps ({res[-1] / r0:.3e})")


def test_mixed_converged_against_fp64_oracle(oracle):
    print("[mixed vs FP64 oracle] bump channel", flush=True)
    dspec = cases.bump_case(precision="mixed")
    ospec = cases.bump_case(ilu_order="natural")
    dc = DeviceCase(dspec)
    oc = oracle.OracleCase(ospec)
    sd, dd = _converge(dc.iterate)
    so, do = _converge(oc.iterate)
    patch = dspec.mesh.patch_index("bottom")
    d, l = (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)
    fd = dc.forces(patch, d, l, 1e5, 1.0, 1.0, 1.0)
    fo = oc.forces(patch, d, l, 1e5, 1.0, 1.0, 1.0)
    Wd, Wo = dc.get(9).reshape(-1, 5), oc.get(9).reshape(-1, 5)
    dW = np.abs(Wd - Wo).max(axis=0) / np.abs(Wo).max(axis=0)
    dF = abs(fd[2] - fo[2]) / abs(fo[2])
    print(f"device mixed: {sd} steps (drop {dd:.2e}); oracle FP64: {so} steps (drop {do:.2e})", flush=True)
    print(f"drag force mixed {fd[2]:.15e} oracle {fo[2]:.15e} relative {dF:.2e}; state per component {dW}",
          flush=True)
    assert dF <= 1e-8
    assert np.all(dW <= 1e-9)
