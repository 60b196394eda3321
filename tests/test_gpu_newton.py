"""The exact-Jacobian (Newton) implicit operator: V/dtau I + V dR/dW from the
exact linearisation of the first-order inviscid residual (with the viscous
terms' compact linearisation when viscous).  It converges the
bump channel to round-off in a few tens of steps where the reference's Roe
operator needs hundreds, in FP64 and in mixed precision, and the state it
converges to is a root of the oracle's residual (the reference's algorithm),
i.e. the same solution the Roe iteration approaches."""
import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases

pytestmark = pytest.mark.gpu

NEWTON = dict(jacobian="exact", initial_co=10.0, min_co=1.0, max_co=1e10, rel_tol=1e-3, max_iter=4)


def rnorm(R):
    return float(np.sqrt(np.sum(np.asarray(R) ** 2)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_newton_converges_to_the_oracle_root(oracle, precision):
    print(f"[newton] bump HLLC {precision}", flush=True)
    dc = DeviceCase(cases.bump_case(precision=precision, **NEWTON))
    res = dc.iterate(80)
    print("drop per 10 steps", (res[9::10] / res[0]).tolist(), flush=True)
    assert np.all(np.isfinite(res))
    k = int(np.argmax(res / res[0] <= 1e-10))
    assert res[k] / res[0] <= 1e-10, res[-1] / res[0]
    assert k < 80
    # the Roe path after the same number of steps is far from converged
    droe = DeviceCase(cases.bump_case(precision=precision))
    rroe = droe.iterate(k + 1)
    assert rroe[-1] / rroe[0] > 1e3 * res[k] / res[0]
    # the converged state is a root of the oracle's residual
    oc = oracle.OracleCase(cases.bump_case())
    oc.put(9, dc.get(9))
    oc.update_fields()
    oc.flux_residual()
    oc.update_fields()
    oc.flux_residual()
    assert rnorm(oc.get(8)) <= 1e-9 * res[0], rnorm(oc.get(8)) / res[0]


@pytest.mark.parametrize("precision,heat", [("fp64", 0), ("mixed", 0), ("fp64", 1)])
def test_viscous_newton_converges_to_the_oracle_root(oracle, precision, heat):
    """Laminar no-slip box (Sutherland air): the Newton operator with the
    compact linearisation of the viscous terms (and of the heat-conduction
    extension when on) converges where the Roe operator after the same steps
    is far behind, to a root of the oracle's viscous residual."""
    print(f"[newton] viscous box HLLC {precision} heat {heat}", flush=True)
    kw = dict(NEWTON, rel_tol=1e-4, max_iter=6, heat_conduction=heat)
    dc = DeviceCase(cases.box_viscous_case(nx=24, ny=16, precision=precision, perturb=0.02, **kw))
    res = dc.iterate(150)
    print("drop per 10 steps", (res[9::10] / res[0]).tolist(), flush=True)
    assert np.all(np.isfinite(res))
    k = int(np.argmax(res / res[0] <= 1e-10))
    assert res[k] / res[0] <= 1e-10, res[-1] / res[0]
    droe = DeviceCase(cases.box_viscous_case(nx=24, ny=16, precision=precision, perturb=0.02, heat_conduction=heat))
    rroe = droe.iterate(k + 1)
    assert rroe[-1] / rroe[0] > 1e3 * res[k] / res[0]
    oc = oracle.OracleCase(cases.box_viscous_case(nx=24, ny=16, perturb=0.02, heat_conduction=heat))
    oc.put(9, dc.get(9))
    oc.update_fields()
    oc.flux_residual()
    oc.update_fields()
    oc.flux_residual()
    assert rnorm(oc.get(8)) <= 1e-9 * res[0], rnorm(oc.get(8)) / res[0]


def test_sa_newton_outruns_roe():
    """The SA box (segregated nuTilda, muEff frozen in the operator): more
    drop in 60 Newton steps than in 60 Roe steps."""
    print("[newton] SA box", flush=True)
    dn = DeviceCase(cases.box_sa_case("SAneg", negative_fraction=0.0, **dict(NEWTON, rel_tol=1e-4, max_iter=6)))
    rn = dn.iterate(60)
    dr = DeviceCase(cases.box_sa_case("SAneg", negative_fraction=0.0))
    rr = dr.iterate(60)
    print("Newton", rn[-1] / rn[0], "Roe", rr[-1] / rr[0], flush=True)
    assert np.all(np.isfinite(rn)) and rn[-1] / rn[0] < 1e-2 * rr[-1] / rr[0]


def test_exact_jacobian_rejects_what_it_does_not_linearise():
    from rhodst_amd._capi import RdstError
    with pytest.raises(RdstError):
        DeviceCase(cases.bump_case(scheme="HLLCP", **NEWTON))
    with pytest.raises(RdstError):
        DeviceCase(cases.bump_case(reconstruction="OSPRE", **NEWTON))


def test_newton_operator_read_back_is_v_j_plus_diagonal(oracle):
    """The Newton operator is written straight into the sliced-ELL layout of
    the SpMV; read back in block-CSR it must be V dR/dW + rDeltaTDST V I:
    A v = V (J v) + rDeltaT V v for a random v, with J v from the adjoint's
    exact linearisation (rdst_adjoint_apply)."""
    print("[newton] operator read-back, bump HLLC", flush=True)
    dc = DeviceCase(cases.bump_case(**NEWTON))
    dc.iterate(3)
    dc.flux_residual()
    dc.jacobian_assemble()
    rp, cols, vals = dc.bcsr()
    n = len(rp) - 1
    rng = np.random.default_rng(5)
    v = rng.standard_normal((n, 5))
    Av = np.zeros((n, 5))
    for i in range(n):
        for s in range(rp[i], rp[i + 1]):
            Av[i] += vals[s] @ v[cols[s]]
    dc.adjoint_linearise()
    Jv = dc.adjoint_apply(v)
    V = np.asarray(oracle.geometry(dc.spec.mesh)["V"])
    rdt = dc.get(10)
    ref = V[:, None] * Jv + (rdt * V)[:, None] * v
    assert np.abs(Av - ref).max() <= 1e-10 * np.abs(ref).max()


def test_update_limit_parity(oracle):
    """The update-limit extension on the device against the oracle: ten
    steps of the perturbed forward step at CFL 5 with every cell's change of
    rho and rhoE limited to 2 %."""
    print("[newton] update limit parity, forward step", flush=True)
    kw = dict(perturb=0.02, initial_co=5.0, min_co=5.0, max_co=5.0, update_limit=0.02)
    dc = DeviceCase(cases.forward_step_case(30, 10, **kw))
    oc = oracle.OracleCase(cases.forward_step_case(30, 10, **kw))
    r1, r0 = dc.iterate(10), oc.iterate(10)
    assert np.abs(r1 - r0).max() <= 1e-9 * r0.max()
    W1, W0 = dc.get(9), oc.get(9)
    assert np.abs(W1 - W0).max() <= 1e-9 * np.abs(W0).max()
