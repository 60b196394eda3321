"""The surface-normal gradient schemes of the viscous laplacian (fvSchemes
laplacianSchemes "Gauss linear corrected | uncorrected | limited psi", as
tutorials/oneraM6Turb/system/fvSchemes:44 selects "limited 0.33").  The
limiter follows foam-extend 4.1's limitedSnGrad (not in the reference tree):
the non-orthogonal correction c of a face is scaled by
min(psi |u| / ((1 - psi) |c| + 1e-15), 1), u = dc (U_N - U_P), |.| the vector
magnitude.

Pins: on two cells sharing one non-orthogonal face every term of the
residual but the laplacian's face correction is the same for all three
schemes, so R(limited) - R(uncorrected) = k (R(corrected) - R(uncorrected))
with one k for all three momentum components; k is checked against the
limiter evaluated from quantities known independently of the oracle (the
cell velocities, the face geometry, and the correction recovered from the
corrected-minus-uncorrected difference).  The ends: limited 0 is uncorrected
and limited 1 corrected, on the swept-wing mesh."""
import numpy as np
import pytest

from rhodst_amd import cases, meshgen
from rhodst_amd._abi import make_bcs, make_settings, make_thermo


def two_cells(U0, U1, **settings):
    """A unit cube and a parallelepiped sheared by 0.5 in y over x in [1, 2],
    sharing the face x = 1 (n = (1, 0, 0), d = (1, 0.25, 0): dc = 1,
    corr = (0, -0.25, 0)); a no-slip wall under the cube, zeroGradient
    elsewhere; constant mu = 1e-3."""
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1],
                  [2, 0.5, 0], [2, 1.5, 0], [2, 0.5, 1], [2, 1.5, 1]], dtype=np.float64)
    H = np.array([[0, 1, 2, 3, 4, 5, 6, 7], [1, 8, 9, 2, 5, 10, 11, 6]])

    def classify(cells, lf, cen):
        return np.where((np.abs(cen[:, 1]) < 1e-12) & (cen[:, 0] < 1.0), 0, 1)

    m = meshgen.hex_polymesh(P, H, classify, [("wall", meshgen.PATCH_WALL), ("sides", meshgen.PATCH_GENERIC)])
    zg = {"p": "zeroGradient", "T": "zeroGradient", "U": "zeroGradient"}
    wall = {"p": "zeroGradient", "T": "zeroGradient", "U": ("fixedValue", (0.0, 0.0, 0.0))}
    kw = dict(flux_scheme="AUSM+", viscous=1, p_bounds=(1.0, 1e7), T_bounds=(1.0, 1e4), rho_bounds=(1e-4, 1e3))
    kw.update(settings)
    return cases.CaseSpec("twoCells", m, make_thermo(Cp=1004.5, R=287.0, mu=1e-3, Pr=0.72),
                          make_bcs(m, {"wall": wall, "sides": zg}), make_settings(**kw),
                          np.full(2, 1e5), np.full(2, 300.0), np.ascontiguousarray([U0, U1], dtype=np.float64))


def residual(oracle, spec):
    oc = oracle.OracleCase(spec)
    oc.update_fields()
    oc.flux_residual()
    return oc.get(8)


@pytest.mark.parametrize("psi,U1", [(0.33, (12.0, -3.0, 0.5)), (0.33, (10.05, 1.98, 1.0)), (0.8, (10.3, 2.2, 0.9)),
                                    (0.5, (10.0, 2.0, 1.0))])
def test_limited_correction_on_one_face(oracle, psi, U1):
    U0 = (10.0, 2.0, 1.0)
    Rc = residual(oracle, two_cells(U0, U1))
    Ru = residual(oracle, two_cells(U0, U1, laplacian="uncorrected"))
    Rl = residual(oracle, two_cells(U0, U1, laplacian=("limited", psi)))
    # the face: |Sf| = 1, V0 = 1, mu = 1e-3, dc = 1, owner = the cube
    corr_term = Rc[0, 1:4] - Ru[0, 1:4]            # = -mu |Sf| c / V0
    c = -corr_term / 1e-3
    assert np.linalg.norm(c) > 1e-3                 # the face is non-orthogonal and sees a gradient
    u = np.array(U1) - np.array(U0)
    k_expected = min(psi * np.linalg.norm(u) / ((1.0 - psi) * np.linalg.norm(c) + 1e-15), 1.0)
    k = (Rl[0, 1:4] - Ru[0, 1:4])
    assert np.allclose(k, k_expected * corr_term, rtol=1e-9, atol=1e-15 * np.abs(Rc).max())
    # the neighbour sees the same face term with the opposite sign and its own volume (1)
    assert np.allclose(Rl[1, 1:4] - Ru[1, 1:4], -k, rtol=1e-9, atol=1e-15 * np.abs(Rc).max())
    # continuity and energy do not see the laplacian's scheme (snGradSchemes stays corrected)
    assert np.array_equal(Rl[:, [0, 4]], Rc[:, [0, 4]])


def test_limited_ends_on_the_wing(oracle):
    """limited 0 = uncorrected and limited 1 = corrected (every face with a
    velocity difference), viscous terms with both laplacian and snGrad
    schemes switched, on the small swept wing with a seeded state."""
    mesh = meshgen.swept_wing(n_wake=6, n_surf=12, n_normal=12, n_span_wing=6, n_span_out=3, n_layers=6)
    rng = np.random.default_rng(3)

    def run(**kw):
        spec = cases.swept_wing_case(mesh=mesh, viscous=True, **kw)
        spec.U0 = np.ascontiguousarray(spec.U0 * (1 + 0.05 * rng.uniform(-1, 1, spec.U0.shape)))
        rng.bit_generator.state = state
        return residual(oracle, spec)

    state = rng.bit_generator.state
    Rc = run()
    Ru = run(laplacian="uncorrected", sngrad="uncorrected")
    assert not np.array_equal(Rc, Ru)
    assert np.array_equal(run(laplacian=("limited", 0.0), sngrad=("limited", 0.0)), Ru)
    R1 = run(laplacian=("limited", 1.0), sngrad=("limited", 1.0))
    assert np.abs(R1 - Rc).max() <= 1e-13 * np.abs(Rc).max()


def test_heat_conduction_term_on_one_face(oracle):
    """The heat-conduction extension on the two cells: with T zeroGradient on
    every patch the only conductive flux is the shared face's,
    alpha |Sf| dc (h_1 - h_0) with alpha = mu / Pr (constant transport),
    h = Cp T, |Sf| = 1, dc = 1, V = 1: R4 of the cube drops by it, R4 of the
    other cell rises by it, nothing else changes."""
    U0, U1 = (10.0, 2.0, 1.0), (12.0, -3.0, 0.5)
    spec0 = two_cells(U0, U1)
    spec1 = two_cells(U0, U1, heat_conduction=1)
    for s in (spec0, spec1):
        s.T0 = np.array([300.0, 340.0])
    R0 = residual(oracle, spec0)
    R1 = residual(oracle, spec1)
    q = (1e-3 / 0.72) * 1.0 * 1.0 * (1004.5 * 340.0 - 1004.5 * 300.0)
    assert np.isclose(R1[0, 4] - R0[0, 4], -q, rtol=1e-12)
    assert np.isclose(R1[1, 4] - R0[1, 4], q, rtol=1e-12)
    assert np.array_equal(R1[:, :4], R0[:, :4])
