"""GPU parity of grad(U) cellLimited Gauss linear (foam-extend 4.1
cellLimitedGrad; tutorials/RAE2822/system/fvSchemes:27) against the oracle:
the limited gradient itself and the residual that reads it (tauMC from the
limited gradient, the snGrad correction from the Gauss linear one) on
tetrahedra over prism layers, twelve Spalart-Allmaras steps (strain and
vorticity from the limited gradient), and the adjoint's refusal of the
limiter it does not linearise."""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases

pytestmark = pytest.mark.gpu

GRAD_U = 39


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("grad", ["Gauss linear", ("cellLimited", 1.0), ("cellLimited", 0.7)])
def test_grad_u_and_residual_parity(oracle, grad):
    print(f"[grad(U) parity] {grad}", flush=True)
    spec = cases.tet_prism_case(model=None, scheme="Kurganov", perturb=0.2, grad_u=grad)
    dc = DeviceCase(spec)
    oc = oracle.OracleCase(spec)
    dc.flux_residual()
    oc.flux_residual()
    g_d, g_o = dc.get(GRAD_U), oc.get(GRAD_U)
    assert rel(g_d, g_o) <= 1e-11, rel(g_d, g_o)
    assert rel(dc.get(8), oc.get(8)) <= 1e-11
    if grad != "Gauss linear":
        spec0 = cases.tet_prism_case(model=None, scheme="Kurganov", perturb=0.2)
        g0 = oracle.OracleCase(spec0).get(GRAD_U)
        assert rel(g_o, g0) > 1e-3  # the limiter acts on this state


def test_limited_sa_iterate_parity(oracle):
    from rhodst_amd import _capi
    print("[grad(U) cellLimited SAneg iterate]", flush=True)
    spec = cases.tet_prism_case(model="SAneg", scheme="Kurganov", turb_kw={"convection": "linearUpwind"},
                                grad_u=("cellLimited", 1.0))
    dc = DeviceCase(spec)
    oc = oracle.OracleCase(spec)
    oc.set_sa_partition(dc.mesh.get_int(_capi.M_SA_PARTITION))
    oc.set_partition(dc.mesh.get_int(_capi.M_PARTITION))
    r0 = oc.iterate(12)
    r1 = dc.iterate(12)
    assert np.all(np.isfinite(r1))
    assert rel(r1, r0) <= 1e-9, (r1, r0)
    assert rel(dc.get(9), oc.get(9)) <= 1e-8
    for f in (29, 30):
        assert rel(dc.get(f), oc.get(f)) <= 1e-7, f


def test_adjoint_refuses_cell_limited_grad():
    print("[grad(U) cellLimited adjoint refusal]", flush=True)
    dc = DeviceCase(cases.tet_prism_case(model=None, scheme="Kurganov", grad_u=("cellLimited", 1.0)))
    dc.flux_residual()
    with pytest.raises(Exception, match="cellLimited"):
        dc.adjoint_linearise()
