"""Pins of the oracle's viscous calcFields terms (baseFluxDST.C:229-250) by
closed forms on a uniform grid: the viscous residual of U = (y^2, 0, 0) is
-2 mu in x (laplacian 2, div(mu dev2(grad U^T)) = 0; without the transpose
it would be -4 mu), Couette flow U = (a y, 0, 0) has zero viscous momentum
residual and dissipates -mu a^2 in the energy row."""
import numpy as np
import pytest

from rhodst_amd import cases, meshgen
from rhodst_amd._abi import make_bcs, make_settings, make_thermo

MU = 1.0e-3


def viscous_residual(oracle, U, scheme="HLLC", n=12):
    mesh = meshgen.box_2d(n, n)
    zg = {"p": "zeroGradient", "T": "zeroGradient", "U": "zeroGradient"}
    bcs = make_bcs(mesh, {k: zg for k in ("left", "right", "bottom", "top")})
    p0 = np.full(mesh.n_cells, 1.0e5)
    T0 = np.full(mesh.n_cells, 300.0)
    out = []
    for visc in (1, 0):
        spec = cases.CaseSpec("v", mesh, make_thermo(Cp=1004.5, R=287.0, mu=MU, Pr=0.72), bcs,
                              make_settings(flux_scheme=scheme, viscous=visc, p_bounds=(1, 1e7), T_bounds=(1, 1e4)),
                              p0, T0, U)
        oc = oracle.OracleCase(spec)
        oc.flux_residual()
        out.append(oc.get(8))
    C = oracle.geometry(mesh)["C"]
    h = 1.0 / n
    ij = np.floor(C[:, :2] / h).astype(int)
    interior = np.all((ij >= 3) & (ij < n - 3), axis=1)
    return out[0] - out[1], interior, C


@pytest.mark.parametrize("scheme", ["HLLC", "AUSM+", "Kurganov"])
def test_quadratic_shear_laplacian(oracle, scheme):
    mesh = meshgen.box_2d(12, 12)
    C = oracle.geometry(mesh)["C"]
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = 50.0 * C[:, 1] ** 2
    Rv, interior, _ = viscous_residual(oracle, U, scheme)
    assert np.allclose(Rv[interior, 1], -2.0 * MU * 50.0, rtol=1e-9)
    assert np.allclose(Rv[interior, 2:4], 0.0, atol=1e-12)
    assert np.allclose(Rv[interior, 0], 0.0, atol=0)


@pytest.mark.parametrize("scheme", ["HLLC", "AUSM+", "Kurganov"])
def test_couette_dissipation(oracle, scheme):
    mesh = meshgen.box_2d(12, 12)
    C = oracle.geometry(mesh)["C"]
    a = 30.0
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = a * C[:, 1]
    Rv, interior, _ = viscous_residual(oracle, U, scheme)
    assert np.allclose(Rv[interior, 1:4], 0.0, atol=1e-12)
    assert np.allclose(Rv[interior, 4], -MU * a * a, rtol=1e-9)


def test_rigid_rotation_is_stress_free(oracle):
    """U = w (-y, x, 0): antisymmetric gradient, tr = 0, constant tau and a
    zero laplacian -> no viscous momentum residual in the interior."""
    mesh = meshgen.box_2d(12, 12)
    C = oracle.geometry(mesh)["C"]
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = -20.0 * C[:, 1]
    U[:, 1] = 20.0 * C[:, 0]
    Rv, interior, _ = viscous_residual(oracle, U)
    assert np.allclose(Rv[interior, 1:4], 0.0, atol=1e-12)
