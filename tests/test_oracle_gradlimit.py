"""Pins of the oracle's cellLimited grad(U) (foam-extend 4.1 cellLimitedGrad,
selected by tutorials/RAE2822/system/fvSchemes:27 "cellLimited Gauss linear
1").

* Closed forms on a uniform 8 x 4 grid with zeroGradient boundaries and U
  varying along x only: U_x = (0, 0, 0, 1, 10, 10, 10, 10) by column, a kink
  that no admissible linear reconstruction crosses, and U_y = a x, a linear
  ramp.  The limited dU_x/dx values (k = 1: 2/h in column 3, zero elsewhere;
  k = 0.9: 2/(9h), 38/(9h), 2/h in columns 2-4) and dU_y/dx (a inside, 0 or
  2a/9 in the boundary columns) are worked out by hand from the definition,
  and the two components' limiters differ in column 3, so a limiter applied
  to the gradient's rows instead of its columns fails.
* On tetrahedra and prisms with a random U, the limited gradient is checked
  against the definition of the limiter rather than its loops: each column j
  is the Gauss linear column times one factor l_j in [0, 1]; every
  extrapolation U_j + (Cf - C) & grad to a face centre of the cell stays in
  [min, max] of U_j over the cell, its face neighbours and boundary values
  (widened by (1/k - 1)(max - min) for k < 1); and where l_j < 1 some face
  reaches that bound (the factor is the largest admissible one).
"""
import numpy as np
import pytest

from rhodst_amd import cases, meshgen
from rhodst_amd._abi import make_bcs, make_settings, make_thermo

GRAD_U = 39


def kink_grad(oracle, k, n=8, ny=4, a=3.0):
    mesh = meshgen.box_2d(n, ny)
    zg = {"p": "zeroGradient", "T": "zeroGradient", "U": "zeroGradient"}
    bcs = make_bcs(mesh, {p: zg for p in ("left", "right", "bottom", "top")})
    C = oracle.geometry(mesh)["C"]
    col = np.floor(C[:, 0] * n).astype(int)
    ux = np.array([0, 0, 0, 1, 10, 10, 10, 10], dtype=float)[col]
    U = np.zeros((mesh.n_cells, 3))
    U[:, 0] = ux
    U[:, 1] = a * C[:, 0]
    grad = ("cellLimited", k) if k is not None else "Gauss linear"
    spec = cases.CaseSpec("kink", mesh, make_thermo(Cp=1004.5, R=287.0, mu=1e-3, Pr=0.72), bcs,
                          make_settings(flux_scheme="HLLC", viscous=1, grad_u=grad, p_bounds=(1, 1e7)),
                          np.full(mesh.n_cells, 1e5), np.full(mesh.n_cells, 300.0), U)
    oc = oracle.OracleCase(spec)
    return oc.get(GRAD_U).reshape(-1, 3, 3), col, 1.0 / n, a


def by_column(g, col, i, j):
    vals = [g[col == c, i, j] for c in range(col.max() + 1)]
    for v in vals:  # every row of a column holds the same gradient
        assert np.ptp(v) <= 1e-9 * (1 + np.abs(v).max())
    return np.array([v[0] for v in vals])


def test_gauss_linear_reference_values(oracle):
    g, col, h, a = kink_grad(oracle, None)
    assert np.allclose(by_column(g, col, 0, 0), np.array([0, 0, 0.5, 5, 4.5, 0, 0, 0]) / h, rtol=1e-12, atol=1e-9)
    assert np.allclose(by_column(g, col, 0, 1), [a / 2] + [a] * 6 + [a / 2], rtol=1e-9)
    assert np.allclose(g[:, 1:, :], 0.0, atol=1e-9)


def test_cell_limited_k1_closed_form(oracle):
    g, col, h, a = kink_grad(oracle, 1.0)
    # column 3 (U_x = 1 between 0 and 10): the face towards 0 caps the
    # factor at 1 / 2.5 = 0.4 of 5/h; columns 2 and 4 sit at a local
    # min / max on one side, so their factor is 0
    assert np.allclose(by_column(g, col, 0, 0), np.array([0, 0, 0, 2, 0, 0, 0, 0]) / h, rtol=1e-12, atol=1e-9)
    # the ramp is exact inside; the boundary columns extrapolate past their
    # zeroGradient boundary value, which is the cell's own
    assert np.allclose(by_column(g, col, 0, 1), [0] + [a] * 6 + [0], rtol=1e-9, atol=1e-9)
    assert np.allclose(g[:, 1:, :], 0.0, atol=1e-9)


def test_cell_limited_k09_closed_form(oracle):
    g, col, h, a = kink_grad(oracle, 0.9)
    # the bounds widen by (1/0.9 - 1)(max - min) = (max - min)/9
    assert np.allclose(by_column(g, col, 0, 0), np.array([0, 0, 2 / 9, 38 / 9, 2, 0, 0, 0]) / h, rtol=1e-12,
                       atol=1e-9)
    assert np.allclose(by_column(g, col, 0, 1), [2 * a / 9] + [a] * 6 + [2 * a / 9], rtol=1e-9)


@pytest.mark.parametrize("k", [0.5, 0.0])
def test_cell_limited_wide_or_off_is_gauss_linear(oracle, k):
    """k = 0.5 widens every bound of this field past its extrapolations;
    k below 1e-15 returns the Gauss linear gradient unchanged."""
    g, _, _, _ = kink_grad(oracle, k)
    g0, _, _, _ = kink_grad(oracle, None)
    assert np.array_equal(g, g0)


@pytest.mark.parametrize("k", [1.0, 0.7])
def test_cell_limited_definition_on_tets_and_prisms(oracle, k):
    spec = cases.tet_prism_case("SAneg", "Kurganov", mesh=meshgen.tet_prism_box(6, 4, 3), mach=0.5)
    mesh = spec.mesh
    n = mesh.n_cells
    rng = np.random.default_rng(7)
    C = oracle.geometry(mesh)["C"]
    U = np.stack([100 + 30 * np.sin(4 * C[:, 0]) + 10 * rng.standard_normal(n),
                  20 * C[:, 1] + 5 * rng.standard_normal(n),
                  -10 + 8 * rng.standard_normal(n)], axis=1)
    out = {}
    for grad in ("Gauss linear", ("cellLimited", k)):
        spec.settings = make_settings(flux_scheme="Kurganov", viscous=1, grad_u=grad, p_bounds=(1, 1e7))
        oc = oracle.OracleCase(cases.CaseSpec(spec.name, mesh, spec.thermo, spec.bcs, spec.settings,
                                              np.full(n, 1e5), np.full(n, 300.0), U))
        out[grad if isinstance(grad, str) else "lim"] = oc.get(GRAD_U).reshape(-1, 3, 3)
        bU = oc.get(17)
    g0, g = out["Gauss linear"], out["lim"]
    geo = oracle.geometry(mesh)
    nI = mesh.n_internal
    own = np.asarray(mesh.owner)
    nb = np.asarray(mesh.neighbour)[:nI]
    bcell = own[nI:nI + len(bU)]
    assert len(geo["Cf"]) == nI + len(bU)
    # l_j: one factor per column, in [0, 1]
    den = np.where(np.abs(g0) > 1e-12 * np.abs(g0).max(), g0, np.nan)
    ratio = g / den
    lj = np.nanmax(ratio, axis=1)
    spread = np.nanmax(ratio, axis=1) - np.nanmin(ratio, axis=1)
    assert np.nanmax(spread) < 1e-12
    lj = np.nan_to_num(lj, nan=1.0)
    assert lj.min() >= 0.0 and lj.max() <= 1.0 + 1e-15
    assert (lj < 0.999).sum() > n // 20  # the field is rough enough that the limiter acts
    # bounds over the cell, its face neighbours and its boundary values
    mx, mn = U.copy(), U.copy()
    for a_, b_ in ((own[:nI], nb), (nb, own[:nI])):
        np.maximum.at(mx, a_, U[b_]); np.minimum.at(mn, a_, U[b_])
    np.maximum.at(mx, bcell, bU); np.minimum.at(mn, bcell, bU)
    dmx, dmn = mx - U, mn - U
    if k < 1.0:
        wd = (1.0 / k - 1.0) * (dmx - dmn)
        dmx, dmn = dmx + wd, dmn - wd
    cells = np.concatenate([own[:nI], nb, bcell])
    faces = np.concatenate([np.arange(nI), np.arange(nI), nI + np.arange(len(bU))])
    e = np.einsum("fi,fij->fj", geo["Cf"][faces] - C[cells], g[cells])
    scale = np.abs(U).max() * 1e-11
    assert np.all(e <= dmx[cells] + scale) and np.all(e >= dmn[cells] - scale)
    # the factor is the largest admissible: a limited column reaches a bound
    reach = np.full((n, 3), -np.inf)
    slack = np.minimum(np.abs(e - dmx[cells]), np.abs(e - dmn[cells]))
    np.maximum.at(reach, cells, -slack)
    limited = lj < 1.0 - 1e-9
    assert np.all(-reach[limited] <= scale)


def test_sngrad_correction_reads_the_unlimited_gradient(oracle):
    """The non-orthogonal correction of snGrad(U) interpolates the Gauss
    linear gradient whatever grad(U) selects, so the change of the residual
    from switching the correction on is the same with and without the
    limiter (tetrahedra: the correction vectors are far from zero)."""
    spec = cases.tet_prism_case(None, "Kurganov", mesh=meshgen.tet_prism_box(5, 4, 3), mach=0.5)
    mesh, n = spec.mesh, spec.mesh.n_cells
    rng = np.random.default_rng(3)
    U = np.stack([100 + 10 * rng.standard_normal(n), 5 * rng.standard_normal(n), 5 * rng.standard_normal(n)], axis=1)
    thermo = make_thermo(Cp=1004.5, R=287.0, mu=5.0, Pr=0.72)  # viscous terms comparable to the inviscid ones
    R = {}
    for grad in ("Gauss linear", ("cellLimited", 1.0)):
        for corr in ("corrected", "uncorrected"):
            st = make_settings(flux_scheme="Kurganov", viscous=1, grad_u=grad, laplacian=corr, sngrad=corr,
                               p_bounds=(1, 1e7))
            oc = oracle.OracleCase(cases.CaseSpec("c", mesh, thermo, spec.bcs, st, np.full(n, 1e5),
                                                  np.full(n, 300.0), U))
            oc.flux_residual()
            R[(str(grad), corr)] = oc.get(8)
    d_gl = R[("Gauss linear", "corrected")] - R[("Gauss linear", "uncorrected")]
    d_lim = R[(str(("cellLimited", 1.0)), "corrected")] - R[(str(("cellLimited", 1.0)), "uncorrected")]
    assert np.abs(d_gl).max() > 1e-3 * np.abs(R[("Gauss linear", "corrected")]).max()
    assert np.allclose(d_lim, d_gl, rtol=1e-9, atol=1e-9 * np.abs(d_gl).max())
    # and the limiter does change the residual (through tauMC)
    lim_c = R[(str(("cellLimited", 1.0)), "corrected")]
    assert np.abs(lim_c - R[("Gauss linear", "corrected")]).max() > 1e-6 * np.abs(lim_c).max()
