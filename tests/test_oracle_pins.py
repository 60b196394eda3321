"""Pins of oracle functions that consistency and conservation leave open,
at hand-chosen states where the published method gives plain numbers:

* the AUSM+ / AUSM+up split Mach and pressure polynomials (Liou, J. Comput.
  Phys. 129 (1996) 364-382: beta = 1/8, alpha = 3/16; AUSM+up without
  lowMach keeps alpha = 0, AUSMPLUPDST.C:118-131), whose sums M4+ + M4- and
  P5+ + P5- do not depend on alpha and beta;
* the characteristic far-field conditions in all four regimes
  (farField{Pressure,Temperature,Velocity}DST, P.C:144-215, T.C:144-221,
  U.C:144-217) against Blazek's far-field relations (Computational Fluid
  Dynamics: Principles and Applications, 2001, sect. 8.4), with the two
  places the reference departs from them stated;
* the local time step in both modes (timeDST.C:218-268), including the
  wall-patch term of the neighbouring mode;
* the CFL ramp of setCourantDST on a scripted residual and rho/e history
  (timeDST.C:316-414);
* updateFields.H:40-94: primitives from W and back, and the field bounds.
"""
import numpy as np
import pytest

from rhodst_amd import cases, meshgen
from rhodst_amd._abi import make_bcs, make_settings, make_thermo

F_RHO, F_RHOU, F_RHOE, F_P, F_T, F_U, F_H = 0, 1, 2, 3, 4, 5, 6
F_R, F_W, F_RDELTAT, F_COURANT = 8, 9, 10, 11
F_BU, F_BP, F_BT = 17, 18, 19
F_RHO_OLD, F_E_OLD = 27, 28


# ------------------------------------------------------------------ AUSM polynomials
def face_side(rho, U, e, rPsi, a_star):
    """Reconstructed face state [rho, rhoU, e, rPsi, a*, c] (c unused by AUSM)."""
    U = np.asarray(U, float)
    return np.array([rho, *(rho * U), e, rPsi, a_star, 1.0])


SF = np.array([2.0, 0.0, 0.0])  # |S| = 2, n = x


def _ausm(oracle, scheme, L, R):
    return oracle.face_flux(scheme, L, R, SF, gamma=1.4)


def test_ausm_plus_split_polynomials(oracle):
    """AUSM+ with a* = 1 on both sides (a1/2 = 1): M_L = 0, M_R = 1/2 gives
    Ma1/2 = M4+(0) + M4-(1/2) = 3/8 - 0.1328125 and p1/2 = P5+(0) p_L +
    P5-(1/2) p_R = 0.5 p_L + 0.103515625 p_R; M_L = 1/2, M_R = 0 gives
    M4+(1/2) + M4-(0) = 0.6328125 - 3/8 and P5+(1/2) = 0.896484375."""
    pL, pR = 1.2 * 0.8, 0.9 * 1.1
    L = face_side(1.2, (0.0, 0.3, 0.0), 2.0, 0.8, 1.0)
    R = face_side(0.9, (0.5, -0.2, 0.1), 1.5, 1.1, 1.0)
    out = _ausm(oracle, 0, L, R)
    ma = 0.375 - 0.1328125
    phi = ma * 1.0 * 1.2 * 2.0                      # upwind: Ma1/2 > 0 takes the left rho, U, H
    P = 0.5 * pL + 0.103515625 * pR
    HL = 2.0 + 0.5 * 0.09 + 0.8
    assert out[0] == pytest.approx(phi, rel=1e-14)
    assert out[1:4] == pytest.approx(phi * np.array([0.0, 0.3, 0.0]) + P * SF, rel=1e-14, abs=1e-15)
    assert out[4] == pytest.approx(phi * HL, rel=1e-14)
    assert out[5:8] == pytest.approx([0.0, 0.3, 0.0], abs=1e-15)
    # mirrored Mach numbers
    L2 = face_side(1.2, (0.5, 0.3, 0.0), 2.0, 0.8, 1.0)
    R2 = face_side(0.9, (0.0, -0.2, 0.1), 1.5, 1.1, 1.0)
    out2 = _ausm(oracle, 0, L2, R2)
    ma2 = 0.6328125 - 0.375
    P2 = 0.896484375 * pL + 0.5 * pR
    assert out2[0] == pytest.approx(ma2 * 1.2 * 2.0, rel=1e-14)
    assert out2[1] == pytest.approx(ma2 * 1.2 * 2.0 * 0.5 + P2 * 2.0, rel=1e-14)
    # both sides at M = -1/2: upwinding to the right state, P5+(-1/2) = 0.103515625
    L3 = face_side(1.2, (-0.5, 0.0, 0.0), 2.0, 0.8, 1.0)
    R3 = face_side(0.9, (-0.5, 0.0, 0.0), 1.5, 1.1, 1.0)
    out3 = _ausm(oracle, 0, L3, R3)
    assert out3[0] == pytest.approx((0.1328125 - 0.6328125) * 0.9 * 2.0, rel=1e-14)
    assert out3[1] == pytest.approx(-0.5 * 0.9 * 2.0 * -0.5 + (0.103515625 * pL + 0.896484375 * pR) * 2.0,
                                    rel=1e-14)


def test_ausm_plus_up_split_polynomials(oracle):
    """AUSM+up without lowMach: alpha = 0 (P5+(1/2) = 0.84375,
    P5-(1/2) = 0.15625, P5-(1/4) = 0.31640625) and the interface speed
    a1/2 = min(a*^2/max(a*, U_L), a*^2/max(a*, -U_R)) (AUSMPLUPDST.C:249-252)."""
    pL, pR = 1.2 * 0.8, 0.9 * 1.1
    L = face_side(1.2, (0.0, 0.3, 0.0), 2.0, 0.8, 1.0)
    R = face_side(0.9, (0.5, -0.2, 0.1), 1.5, 1.1, 1.0)
    out = _ausm(oracle, 1, L, R)
    ma = 0.375 - 0.1328125
    assert out[0] == pytest.approx(ma * 1.2 * 2.0, rel=1e-14)
    assert out[1] == pytest.approx((0.5 * pL + 0.15625 * pR) * 2.0, rel=1e-14)
    # U_L = 1.25 a*: a_L = 1/1.25 = 0.8 < a_R = 1, so a1/2 = 0.8, M_L = 1.5625
    # (supersonic: M4+ = M, P5+ = 1), M_R = 0.25 (M4-(1/4) = -0.25048828125)
    L2 = face_side(1.2, (1.25, 0.0, 0.0), 2.0, 0.8, 1.0)
    R2 = face_side(0.9, (0.2, 0.0, 0.0), 1.5, 1.1, 1.0)
    out2 = _ausm(oracle, 1, L2, R2)
    ma2 = 1.5625 - 0.25048828125
    phi2 = ma2 * 0.8 * 1.2 * 2.0
    assert out2[0] == pytest.approx(phi2, rel=1e-14)
    assert out2[1] == pytest.approx(phi2 * 1.25 + (1.0 * pL + 0.31640625 * pR) * 2.0, rel=1e-14)


# ------------------------------------------------------------------ far field
GAMMA = 1.4
R_N = 1.0 / 1.4       # normalised gas: c = 1 at T = 1
CP = 2.5


def _farfield_cell(U_inf, p_I, T_I, U_I):
    """One cell, far field on all four sides (p_inf = T_inf = 1, c_inf = 1);
    the state is set to the interior state I, so the boundary values start
    at the free stream and the face mach numbers are U_inf . n."""
    mesh = meshgen.box_2d(1, 1)
    ff = {"p": ("farField", 1.0), "T": ("farField", 1.0), "U": ("farField", tuple(U_inf)),
          "inf": {"p": 1.0, "T": 1.0, "U": tuple(U_inf)}}
    spec = cases.CaseSpec("ff", mesh, make_thermo(Cp=CP, R=R_N, mu=0.0, Pr=1.0),
                          make_bcs(mesh, {k: ff for k in ("left", "right", "bottom", "top")}),
                          make_settings(flux_scheme="AUSM+", p_bounds=(1e-6, 1e6), T_bounds=(1e-6, 1e6),
                                        rho_bounds=(1e-6, 1e6)),
                          np.array([p_I]), np.array([T_I]), np.array([U_I], float))
    return spec, mesh


def test_farfield_four_regimes(oracle):
    """U_inf = (1.5, 0.5): left face supersonic inflow, right supersonic
    outflow, bottom subsonic inflow, top subsonic outflow.  Interior state
    p 1.2, T 1.1, U (1.3, 0.2, 0)."""
    U_inf = np.array([1.5, 0.5, 0.0])
    p_I, T_I, U_I = 1.2, 1.1, np.array([1.3, 0.2, 0.0])
    spec, mesh = _farfield_cell(U_inf, p_I, T_I, U_I)
    oc = oracle.OracleCase(spec)
    oc.update_fields()          # W is the interior state: only the boundary values change
    bp, bT, bU = oc.get(F_BP), oc.get(F_BT), oc.get(F_BU)
    rho_I = p_I / (R_N * T_I)
    c_I = np.sqrt(GAMMA * R_N * T_I)
    rho_inf = 1.0 / R_N
    n = {"left": np.array([-1.0, 0, 0]), "right": np.array([1.0, 0, 0]),
         "bottom": np.array([0, -1.0, 0]), "top": np.array([0, 1.0, 0])}
    b = {name: k for k, name in enumerate(("left", "right", "bottom", "top"))}  # boundary order = patch order
    # supersonic inflow: the free stream
    assert bp[b["left"]] == pytest.approx(1.0, rel=1e-14)
    assert bT[b["left"]] == pytest.approx(1.0, rel=1e-14)
    assert bU[b["left"]] == pytest.approx(U_inf, rel=1e-14)
    # supersonic outflow: the interior
    assert bp[b["right"]] == pytest.approx(p_I, rel=1e-14)
    assert bT[b["right"]] == pytest.approx(T_I, rel=1e-14)
    assert bU[b["right"]] == pytest.approx(U_I, rel=1e-14)
    # subsonic inflow (Blazek): p_b = (p_a + p_d - rho0 c0 n.(U_a - U_d))/2,
    # rho_b = rho_a + (p_b - p_a)/c0^2, U_b = U_a - n (p_a - p_b)/(rho0 c0)
    nb = n["bottom"]
    pb = 0.5 * (1.0 + p_I - rho_I * c_I * nb @ (U_inf - U_I))
    rhob = rho_inf + (pb - 1.0) / c_I ** 2
    assert bp[b["bottom"]] == pytest.approx(pb, rel=1e-13)
    assert bT[b["bottom"]] == pytest.approx(pb / (R_N * rhob), rel=1e-13)
    assert bU[b["bottom"]] == pytest.approx(U_inf - nb * (1.0 - pb) / (rho_I * c_I), rel=1e-13)
    # subsonic outflow: p_b = p_a (Blazek).  The reference departs from
    # Blazek twice here: U is zero-gradient (valueFraction 1 - supOut - subOut
    # = 0, U.C:193-214), and T takes rho_b = rho_d + (pb - p_d)/c0^2 with the
    # averaged pb of the inflow formula instead of p_a (T.C:194-218)
    nt = n["top"]
    pb_t = 0.5 * (1.0 + p_I - rho_I * c_I * nt @ (U_inf - U_I))
    assert bp[b["top"]] == pytest.approx(1.0, rel=1e-14)
    assert bU[b["top"]] == pytest.approx(U_I, rel=1e-14)
    assert bT[b["top"]] == pytest.approx(pb_t / (R_N * (rho_I + (pb_t - p_I) / c_I ** 2)), rel=1e-13)


# ------------------------------------------------------------------ local time step
def _uniform_box(nx, ny, lx, ly, U, T=1.0, wall_bottom=False, **settings):
    mesh = meshgen.box_2d(nx, ny, lx=lx, ly=ly)
    if wall_bottom:
        mesh.patches = [(nm, meshgen.PATCH_WALL if nm == "bottom" else t, s, sz) for nm, t, s, sz in mesh.patches]
    zg = {"p": "zeroGradient", "T": "zeroGradient", "U": "zeroGradient"}
    spec = {k: zg for k in ("left", "right", "top")}
    spec["bottom"] = {"p": "zeroGradient", "T": "zeroGradient", "U": "slip"} if wall_bottom else zg
    n = mesh.n_cells
    kw = dict(flux_scheme="AUSM+", initial_co=5.0, min_co=0.5, max_co=20.0)
    kw.update(settings)
    return cases.CaseSpec("lts", mesh, make_thermo(Cp=CP, R=R_N, mu=0.0, Pr=1.0), make_bcs(mesh, spec),
                          make_settings(**kw), np.ones(n), np.full(n, T), np.tile(np.asarray(U, float), (n, 1))), mesh


def test_lts_blazek_closed_form(oracle):
    """neighbouring = false: rDeltaTDST = ((|u|+c)/dx + (|v|+c)/dy)/CFL on a
    uniform one-cell-thick grid (dS = half the summed face-area components,
    empty faces excluded), with the scalar courant = initialCo."""
    U = (-2.0, 0.5, 0.0)
    spec, mesh = _uniform_box(6, 4, 3.0, 0.5, U, neighbouring=0)   # dx = 0.5, dy = 0.125
    oc = oracle.OracleCase(spec)
    r = oc.get(F_RDELTAT)
    c = 1.0
    expect = ((2.0 + c) / 0.5 + (0.5 + c) / 0.125) / 5.0
    assert r == pytest.approx(np.full(mesh.n_cells, expect), rel=1e-13)


def test_lts_neighbouring_closed_form_and_wall_term(oracle):
    """neighbouring = true: rDeltaTDST_c = max over the cell's internal faces
    of (|U_f| + c_f) deltaCoeffs, and on wall faces
    0.5 deltaCoeffs_b c_c + |U_c| (the |U| term is not scaled by deltaCoeffs,
    timeDST.C:253-262), over courant.  With h = 4, |U| = 3.5, c = 1 the wall
    term (0.5*0.5*1 + 3.5 = 3.75) beats the face term ((3.5+1)/4 = 1.125)."""
    U = (3.5, 0.0, 0.0)
    spec, mesh = _uniform_box(3, 3, 12.0, 12.0, U, wall_bottom=True, neighbouring=1)
    oc = oracle.OracleCase(spec)
    r = oc.get(F_RDELTAT)
    C = oracle.geometry(mesh)["C"]
    bottom = C[:, 1] < 4.0
    assert r[~bottom] == pytest.approx(np.full((~bottom).sum(), (3.5 + 1.0) / 4.0 / 5.0), rel=1e-13)
    assert r[bottom] == pytest.approx(np.full(bottom.sum(), (0.5 * 0.5 * 1.0 + 3.5) / 5.0), rel=1e-13)
    # the same without the wall patch: the face term everywhere
    spec2, _ = _uniform_box(3, 3, 12.0, 12.0, U, wall_bottom=False, neighbouring=1)
    r2 = oracle.OracleCase(spec2).get(F_RDELTAT)
    assert r2 == pytest.approx(np.full(9, 1.125 / 5.0), rel=1e-13)


def test_cfl_ramp_scripted_history(oracle):
    """setCourantDST on five cells in a row (initialCo 5, minCo 0.5,
    maxCo 20): residualRatio = clamp(|R|prev/|R|, 0.1, 2); a cell whose rho
    or e fell below 0.95 of the previous value gets factor 0.5 and its face
    neighbours 0.75; courant = clamp(courant ratio factor, minCo, maxCo)."""
    spec, mesh = _uniform_box(5, 1, 5.0, 1.0, (0.5, 0.0, 0.0), neighbouring=1)
    oc = oracle.OracleCase(spec)
    n = mesh.n_cells
    rho = oc.get(F_RHO)
    rhoU = oc.get(F_RHOU)
    e = oc.get(F_RHOE) / rho - 0.5 * np.sum((rhoU / rho[:, None]) ** 2, axis=1)
    order = np.argsort(oracle.geometry(mesh)["C"][:, 0])   # cells left to right

    def step(res, rho_old=None, e_old=None):
        R = np.zeros((n, 5))
        R[0, 0] = res
        oc.put(F_R, R)
        oc.put(F_RHO_OLD, rho if rho_old is None else rho_old)
        oc.put(F_E_OLD, e if e_old is None else e_old)
        assert oc.set_courant() == pytest.approx(res, rel=1e-15)
        return oc.get(F_COURANT)[order]

    assert step(1.0) == pytest.approx([5.0] * 5)           # first call: ratio 1
    assert step(0.1) == pytest.approx([10.0] * 5)          # ratio 10 -> 2
    assert step(0.05) == pytest.approx([20.0] * 5)         # ratio 2
    assert step(0.025) == pytest.approx([20.0] * 5)        # 40 -> maxCo
    assert step(1.0) == pytest.approx([2.0] * 5)           # ratio 0.025 -> 0.1
    ro = rho.copy()
    ro[order[2]] = rho[order[2]] / 0.94                     # rho fell by 6 % in the middle cell
    assert step(1.0, rho_old=ro) == pytest.approx([2.0, 1.5, 1.0, 1.5, 2.0])
    eo = e.copy()
    eo[order[0]] = e[order[0]] / 0.9                        # e fell by 10 % in the first cell
    ro2 = rho.copy()
    ro2[order[4]] = rho[order[4]] / 0.96                    # a 4 % drop is no drop
    assert step(1.0, rho_old=ro2, e_old=eo) == pytest.approx([1.0, 1.125, 1.0, 1.5, 2.0])
    # a 4 % drop in rho or in e is no drop, in any cell (owner or neighbour side)
    assert step(1.0, rho_old=rho / 0.96, e_old=e / 0.96) == pytest.approx([1.0, 1.125, 1.0, 1.5, 2.0])
    assert step(20.0) == pytest.approx([0.5] * 5)          # ratio 0.05 -> 0.1, then minCo
    # the time step follows the ramped courant (neighbouring: (|U|+c)/h / courant)
    assert oc.get(F_RDELTAT)[order] == pytest.approx(np.full(5, 1.5 / 0.5), rel=1e-13)


# ------------------------------------------------------------------ updateFields
def test_update_fields_round_trip_and_bounds(oracle):
    """W -> rho, U = rhoU/rho, h = Cp/Cv (rhoE/rho - |U|^2/2), T = h/Cp,
    p = rho R T recovers the primitives W was built from; out-of-bounds
    rho and h are clamped (rho first, then U from the clamped rho), p after,
    and W itself is not re-synced (updateFields.H:40-94)."""
    spec, mesh = _uniform_box(3, 2, 3.0, 2.0, (0.0, 0.0, 0.0), rho_bounds=(0.2, 5.0), T_bounds=(0.5, 3.0),
                              p_bounds=(0.1, 4.0))
    oc = oracle.OracleCase(spec)
    rng = np.random.default_rng(7)
    n = mesh.n_cells
    p = rng.uniform(0.5, 2.0, n)
    T = rng.uniform(0.8, 2.0, n)
    U = rng.uniform(-1.0, 1.0, (n, 3))
    rho = p / (R_N * T)
    Cv = CP - R_N
    W = np.column_stack([rho, rho[:, None] * U, rho * (Cv * T + 0.5 * np.sum(U * U, axis=1))])
    oc.put(F_W, W)
    oc.update_fields()
    assert oc.get(F_RHO) == pytest.approx(rho, rel=1e-15)
    assert oc.get(F_U) == pytest.approx(U, rel=1e-14, abs=1e-15)
    assert oc.get(F_T) == pytest.approx(T, rel=1e-14)
    assert oc.get(F_P) == pytest.approx(p, rel=1e-14)
    assert oc.get(F_H) == pytest.approx(CP * T, rel=1e-14)
    # bounds: cell 0 rho below rhoMin, cell 1 T above TMax, cell 2 p above pMax
    Wb = W.copy()
    Wb[0] = [0.1, 0.05, 0.0, 0.0, 0.1 * (Cv * 1.0) + 0.5 * 0.05 ** 2 / 0.1]
    Wb[1] = [1.0, 0.0, 0.0, 0.0, 1.0 * Cv * 5.0]
    Wb[2] = [4.0, 0.0, 0.0, 0.0, 4.0 * Cv * 2.0]
    oc.put(F_W, Wb)
    oc.update_fields()
    rho_b, U_b, T_b, p_b = oc.get(F_RHO), oc.get(F_U), oc.get(F_T), oc.get(F_P)
    assert rho_b[0] == 0.2
    assert U_b[0] == pytest.approx([0.25, 0.0, 0.0], rel=1e-15)
    # h = Cp/Cv (rhoE/rho - |U|^2/2) with the clamped rho: (0.1 Cv + 0.0125)/0.2 - 0.03125
    h0 = CP / Cv * ((0.1 * Cv + 0.0125) / 0.2 - 0.5 * 0.25 ** 2)
    assert T_b[0] == pytest.approx(h0 / CP, rel=1e-14)
    assert T_b[1] == pytest.approx(3.0, rel=1e-15)          # h clamped to Cp TMax
    assert p_b[1] == pytest.approx(1.0 * R_N * 3.0, rel=1e-14)
    assert p_b[2] == 4.0                                    # rho R T = 4 R_N 2 = 5.71 -> pMax
    assert oc.get(F_W) == pytest.approx(Wb, rel=1e-15)      # W is not re-synced


def test_update_limit_scales_large_changes_only(oracle):
    """The update-limit extension: after one step from the forward step's
    uniform start, every cell's change of rho and rhoE is within the limit of
    its old value, the cells whose unlimited change stayed inside it are
    unchanged, and the others moved along the unlimited direction with the
    binding component at the limit."""
    from rhodst_amd import cases
    lim = 0.02
    a = oracle.OracleCase(cases.forward_step_case(30, 10, perturb=0.02, initial_co=5.0, min_co=5.0, max_co=5.0))
    b = oracle.OracleCase(cases.forward_step_case(30, 10, perturb=0.02, initial_co=5.0, min_co=5.0, max_co=5.0,
                                                 update_limit=lim))
    W0 = a.get(9).copy()
    a.iterate(1)
    b.iterate(1)
    Wa, Wb = a.get(9), b.get(9)
    da, db = Wa - W0, Wb - W0
    ra = np.maximum(np.abs(da[:, 0]) / W0[:, 0], np.abs(da[:, 4]) / W0[:, 4])
    rb = np.maximum(np.abs(db[:, 0]) / W0[:, 0], np.abs(db[:, 4]) / W0[:, 4])
    assert (ra > lim).any() and (ra <= lim).any()
    assert (rb <= lim * (1 + 1e-12)).all()
    inside = ra <= lim
    assert np.array_equal(Wb[inside], Wa[inside])
    out = ~inside
    assert np.allclose(rb[out], lim, rtol=1e-12)
    scale = (rb[out] / ra[out])[:, None]
    assert np.allclose(db[out], scale * da[out], rtol=1e-9, atol=1e-12 * np.abs(W0).max())
