"""Dual time stepping on the GPU (rdst_dual_step) against the oracle: the
forwardStepDual tutorial's controls (Kurganov, neighbouring, fixed physical
deltaT, maxInnerIter 100) on the forward step, and a closed box with the
Blazek time step and adjustTimeStep on.  Per physical step the inner
iteration counts agree, the inner residual histories and the states agree
to iterative rounding drift, and time, deltaT, deltaT0 and courant agree.
Also the error paths of the entry points.
"""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases
from rhodst_amd._capi import RdstError

from test_oracle_dual import closed_box

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def _step_spec():
    return cases.forward_step_case(nx=60, ny=20, scheme="Kurganov", time_integration="Dual", neighbouring=1,
                                   initial_co=1.0, min_co=0.1, max_co=10.0, adjust_time_step=0, max_inner_iter=100,
                                   iteration_start=10, delta_t=0.1)


@pytest.mark.parametrize("name", ["forwardStepDual", "closedBox"])
def test_dual_step_parity(oracle, name):
    print(f"[dual parity] {name}", flush=True)
    spec = _step_spec() if name == "forwardStepDual" else closed_box(neighbouring=0, adjust_time_step=1,
                                                                      min_residual=1e-8)
    oc = oracle.OracleCase(spec)
    dc = DeviceCase(spec)
    to, td = oc.time_info(), dc.time_info()
    for key in ("deltaT", "deltaT0", "courant"):
        assert td[key] == pytest.approx(to[key], rel=1e-12), key
    for step in range(3):
        ko, io, ro = oc.dual_step()
        kd, idv, rd = dc.dual_step()
        print(f"  step {step}: inner {ko} / {kd}, |R| {ro:.6e} / {rd:.6e}", flush=True)
        assert kd == ko
        assert rel(idv, io) <= 1e-6
        assert rd == pytest.approx(ro, rel=1e-7)
        assert rel(dc.get(0), oc.get(0)) <= 1e-9
        to, td = oc.time_info(), dc.time_info()
        for key in ("time", "deltaT", "deltaT0", "courant"):
            assert td[key] == pytest.approx(to[key], rel=1e-9), key


def test_dual_entry_point_errors():
    single = DeviceCase(cases.forward_step_case(nx=30, ny=10, scheme="AUSM+"))
    with pytest.raises(RdstError):
        single.dual_step()
    dual = DeviceCase(cases.forward_step_case(nx=30, ny=10, scheme="AUSM+", time_integration="Dual"))
    with pytest.raises(RdstError):
        dual.iterate(1)
    with pytest.raises(RdstError):
        DeviceCase(cases.forward_step_case(nx=30, ny=10, scheme="AUSM+", time_integration="Dual", jacobian="exact"))


def test_dual_refuses_mixed_precision():
    """The mixed form solves for the correction from zero: its normalised
    initial residual is ~1 on every inner iteration, which steadyCheck cannot
    use (the inner loop would run to maxInnerIter), so a Dual case is FP64."""
    s32 = _step_spec()
    s32.settings.precision = 1  # RDST_PREC_MIXED
    with pytest.raises(RdstError):
        DeviceCase(s32)
