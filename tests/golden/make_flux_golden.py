"""Golden face-flux vectors for the six rhoDST schemes, from a second,
independent restatement in numpy of the reference's formulas (written from
the reference sources, not from the oracle):

  AUSM+     AUSMPLDST.C:127-217        AUSM+up   AUSMPLUPDST.C:209-320 (+ :84-87 constants)
  HLL       HLLDST.C:91-204            HLLC      HLLCDST.C:135-272
  HLLCP     HLLCPDST.C:166-317         Kurganov  KurganovDST.C:93-167

Inputs per face are the reconstructed side quantities the schemes read
(rho, rhoU, e = h - p/rho, rPsi = 1/psi, a* from the cell H, c = sqrt(gamma
rPsi)), the face area vector and, for HLLCP, the interpolated sensor f.  The
states are seeded and cover subsonic and supersonic flow in both directions,
|M| = 1 exactly on one side (where foam's sign(0) = +1 lets both AUSM
branches add), and the lowMach form of AUSM+up.  Run from the repository
root:  python tests/golden/make_flux_golden.py  -> tests/golden/flux_faces.npz
"""
from __future__ import annotations

import os

import numpy as np

GAMMA = 1.4
R_GAS = 287.0
CV = R_GAS / (GAMMA - 1.0)
SMALL, VSMALL = 1.0e-15, 1.0e-300


def foam_sign(x):
    return 1.0 if x >= 0.0 else -1.0


def pos(x):   # DST::pos: x > 0
    return 1.0 if x > 0.0 else 0.0


def pos0(x):  # DST::pos0: x >= 0
    return 1.0 if x >= 0.0 else 0.0


def neg(x):   # DST::neg: x < 0
    return 1.0 if x < 0.0 else 0.0


def neg0(x):  # DST::neg0: x <= 0
    return 1.0 if x <= 0.0 else 0.0


def side_from_primitive(rho, u, p):
    """[rho, rhoU(3), e, rPsi, a*, c] of an ideal-gas cell (e = Cv T)."""
    u = np.asarray(u, dtype=np.float64)
    T = p / (rho * R_GAS)
    e = CV * T
    rhoE = rho * (e + 0.5 * float(u @ u))
    H = max(rhoE / rho, SMALL) + max(p / rho, SMALL)
    a_star = np.sqrt(2.0 * (GAMMA - 1.0) / (GAMMA + 1.0) * H)
    c = np.sqrt(GAMMA * R_GAS * T)
    return np.array([rho, *(rho * u), e, R_GAS * T, a_star, c])


def _unpack(s):
    rho = s[0]
    rhoU = np.array(s[1:4])
    U = rhoU / rho
    e, rPsi, a, c = s[4], s[5], s[6], s[7]
    p = rho * rPsi
    rhoE = rho * (e + 0.5 * float(U @ U))
    return rho, rhoU, U, e, rPsi, a, c, p, rhoE


def ausm(L, R, Sf, up=False, low_mach=False, Mainf=0.5):
    beta = 0.125
    rho_o, rhoU_o, U_o, e_o, rPsi_o, a_o, c_o, p_o, rhoE_o = _unpack(L)
    rho_n, rhoU_n, U_n, e_n, rPsi_n, a_n, c_n, p_n, rhoE_n = _unpack(R)
    magSf = float(np.linalg.norm(Sf))
    n = Sf / magSf
    Uv_o, Uv_n = float(U_o @ n), float(U_n @ n)
    alpha = 3.0 / 16.0
    if not up:
        a12 = 0.5 * (a_o + a_n)
    else:
        ao = a_o ** 2 / max(a_o, Uv_o)
        an = a_n ** 2 / max(a_n, -Uv_n)
        a12 = min(ao, an)
        alpha = 0.0
    M_o, M_n = Uv_o / a12, Uv_n / a12
    fa = 0.0
    Mmean2 = 0.0
    if up and low_mach:
        Mmean2 = 0.5 * (Uv_o ** 2 + Uv_n ** 2) / a12 ** 2
        M0 = np.sqrt(min(1.0, max(Mmean2, Mainf ** 2)))
        fa = M0 * (2.0 - M0)
        alpha = 3.0 / 16.0 * (-4.0 + 5.0 * fa ** 2)

    def sub(M):   # 0.5 (sign(1 - |M|) + |sign(1 - |M|)|): 1 where |M| <= 1
        s = foam_sign(1.0 - abs(M))
        return 0.5 * (s + abs(s))

    def sup(M):   # 1 where |M| >= 1
        s = foam_sign(-1.0 + abs(M))
        return 0.5 * (s + abs(s))

    Ma4_o = sub(M_o) * (0.25 * (M_o + 1.0) ** 2) * (1.0 + 4.0 * beta * (M_o - 1.0) ** 2) + sup(M_o) * max(M_o, 0.0)
    P5_o = (sub(M_o) * 0.25 * (M_o + 1.0) ** 2 * (2.0 - M_o + 4.0 * alpha * M_o * (M_o - 1.0) ** 2)
            + sup(M_o) * pos(M_o))
    Ma4_n = sub(M_n) * (-0.25) * (M_n - 1.0) ** 2 * (1.0 + 4.0 * beta * (M_n + 1.0) ** 2) + sup(M_n) * min(M_n, 0.0)
    P5_n = (sub(M_n) * (-0.25) * (M_n - 1.0) ** 2 * (-2.0 - M_n + 4.0 * alpha * M_n * (M_n + 1.0) ** 2)
            + sup(M_n) * neg(M_n))
    Ma12 = Ma4_o + Ma4_n
    P12 = P5_o * p_o + P5_n * p_n
    if up and low_mach:
        Kp, Ku, sigma = 0.25, 0.75, 1.0
        Mp = -Kp / fa * max(1.0 - sigma * Mmean2, 0.0) * (p_n - p_o) / (0.5 * (rho_o + rho_n)) / a12 ** 2
        pu = -Ku * P5_o * P5_n * (rho_o + rho_n) * a12 * fa * (Uv_o - Uv_n)
        Ma12 += Mp
        P12 += pu
    Up = pos(Ma12) * U_o + neg0(Ma12) * U_n
    phi = Ma12 * a12 * (pos(Ma12) * rho_o + neg0(Ma12) * rho_n) * magSf
    phiU = phi * Up + P12 * Sf
    H_o, H_n = (rhoE_o + p_o) / rho_o, (rhoE_n + p_n) / rho_n
    phiE = phi * (pos(Ma12) * H_o + neg0(Ma12) * H_n)
    return np.array([phi, *phiU, phiE, *Up])


def _roe(L, R, n):
    rho_o, rhoU_o, U_o, e_o, rPsi_o, a_o, c_o, p_o, rhoE_o = _unpack(L)
    rho_n, rhoU_n, U_n, e_n, rPsi_n, a_n, c_n, p_n, rhoE_n = _unpack(R)
    E_o, E_n = e_o + 0.5 * float(U_o @ U_o), e_n + 0.5 * float(U_n @ U_n)
    H_o, H_n = E_o + p_o / rho_o, E_n + p_n / rho_n
    k = np.sqrt(max(VSMALL, rho_n) / max(VSMALL, rho_o))
    Ht = (k * H_n + H_o) / (k + 1.0)
    Ut = (k * U_n + U_o) / (k + 1.0)
    at = np.sqrt((GAMMA - 1.0) * (Ht - 0.5 * float(Ut @ Ut)))
    return k, Ht, Ut, at, float(Ut @ n), E_o, E_n, H_o, H_n


def hll(L, R, Sf):
    rho_o, rhoU_o, U_o, e_o, rPsi_o, a_o, c_o, p_o, rhoE_o = _unpack(L)
    rho_n, rhoU_n, U_n, e_n, rPsi_n, a_n, c_n, p_n, rhoE_n = _unpack(R)
    magSf = float(np.linalg.norm(Sf))
    n = Sf / magSf
    _, _, _, at, Uvt, E_o, E_n, H_o, H_n = _roe(L, R, n)
    Uv_o, Uv_n = float(U_o @ n), float(U_n @ n)
    S_o = min(Uv_o - a_o, Uvt - at)
    S_n = max(Uv_n + a_n, Uvt + at)
    c1, cm, c2 = pos0(S_o), neg(S_o) * pos0(S_n), neg(S_n) * neg(S_o)
    F_o = rho_o * U_o * Uv_o + p_o * n
    F_n = rho_n * U_n * Uv_n + p_n * n
    d = S_n - S_o
    phi = (c1 * rho_o * Uv_o + cm * (S_n * rho_o * Uv_o - S_o * rho_n * Uv_n + S_o * S_n * (rho_n - rho_o)) / d
           + c2 * rho_n * Uv_n) * magSf
    phiU = (c1 * F_o + cm * (S_n * F_o - S_o * F_n + S_o * S_n * (rho_n * U_n - rho_o * U_o)) / d + c2 * F_n) * magSf
    phiE = (c1 * rho_o * Uv_o * H_o
            + cm * (S_n * rho_o * Uv_o * H_o - S_o * rho_n * Uv_n * H_n + S_o * S_n * (rho_n * E_n - rho_o * E_o)) / d
            + c2 * rho_n * Uv_n * H_n) * magSf
    Up = c1 * U_o + cm * (S_n * U_o - S_o * U_n + Uv_o * U_o - Uv_n * U_n) / d + c2 * U_n
    return np.array([phi, *phiU, phiE, *Up])


def hllc(L, R, Sf, f=None):
    """HLLC; with f (the HLLCP sensor value) the HLLCP variant."""
    rho_o, rhoU_o, U_o, e_o, rPsi_o, a_o, c_o, p_o, rhoE_o = _unpack(L)
    rho_n, rhoU_n, U_n, e_n, rPsi_n, a_n, c_n, p_n, rhoE_n = _unpack(R)
    magSf = float(np.linalg.norm(Sf))
    n = Sf / magSf
    k, _, Ut, at, Uvt, E_o, E_n, H_o, H_n = _roe(L, R, n)
    Uv_o, Uv_n = float(U_o @ n), float(U_n @ n)
    S_o = min(Uv_o - a_o, Uvt - at)
    S_n = max(Uv_n + a_n, Uvt + at)
    Ss = ((p_n - p_o + rho_o * Uv_o * (S_o - Uv_o) - rho_n * Uv_n * (S_n - Uv_n))
          / (rho_o * (S_o - Uv_o) - rho_n * (S_n - Uv_n)))
    ps_o = p_o + rho_o * (S_o - Uv_o) * (Ss - Uv_o)
    ps_n = p_n + rho_n * (S_n - Uv_n) * (Ss - Uv_n)
    phip = 0.0
    pm_o = pm_n = None
    pE_o, pE_n = ps_o, ps_n
    if f is not None:
        M_o, M_n = Uv_o / a_o, Uv_n / a_n
        Mt = (k * M_n + M_o) / (k + 1.0)
        cO, cN = rho_o * (S_o - Uv_o), rho_n * (S_n - Uv_n)
        pst = (cN * p_o - cO * p_n - cO * cN * (Uv_o - Uv_n)) / (cN - cO)
        theta = min(max(abs(M_o), abs(M_n)), 1.0)
        pss = pst * theta + (1.0 - theta) * 0.5 * (p_o + p_n)
        psss = f * pss + (1.0 - f) * pst
        phip = (f - 1.0) * S_o * S_n / (S_n - S_o) / (1.0 + abs(Mt)) * (p_n - p_o) / at ** 2
        pm_o = pm_n = psss
        pE_o = pE_n = pst
    else:
        pm_o, pm_n = ps_o, ps_n
    c1 = pos(S_o)
    cm1 = neg0(S_o) * pos(Ss)
    cm2 = neg0(S_o) * neg0(Ss) * pos0(S_n)
    c2 = neg0(S_o) * neg0(Ss) * neg(S_n)
    rE_o, rE_n = rho_o * E_o, rho_n * E_n
    F_o = rhoU_o * Uv_o + p_o * n
    F_n = rhoU_n * Uv_n + p_n * n
    G_o, G_n = (rE_o + p_o) * Uv_o, (rE_n + p_n) * Uv_n
    d1, d2 = S_o - Ss, S_n - Ss
    phi = (c1 * rho_o * Uv_o + cm1 * (Ss * rho_o * (S_o - Uv_o) / d1 + phip)
           + cm2 * (Ss * rho_n * (S_n - Uv_n) / d2 + phip) + c2 * rho_n * Uv_n) * magSf
    phiU = (c1 * F_o + cm1 * ((Ss * (S_o * rhoU_o - F_o) + S_o * pm_o * n) / d1 + phip * Ut)
            + cm2 * ((Ss * (S_n * rhoU_n - F_n) + S_n * pm_n * n) / d2 + phip * Ut) + c2 * F_n) * magSf
    UT2 = float(Ut @ Ut)
    phiE = (c1 * G_o + cm1 * (Ss * (S_o * rE_o - G_o + S_o * pE_o) / d1 + 0.5 * phip * UT2)
            + cm2 * (Ss * (S_n * rE_n - G_n + S_n * pE_n) / d2 + 0.5 * phip * UT2) + c2 * G_n) * magSf
    Up = c1 * U_o + cm1 * (S_o - Uv_o) * U_o / d1 + cm2 * (S_n - Uv_n) * U_n / d2 + c2 * U_n
    return np.array([phi, *phiU, phiE, *Up])


def kurganov(L, R, Sf):
    rho_o, rhoU_o, U_o, e_o, rPsi_o, a_o, c_o, p_o, rhoE_o = _unpack(L)
    rho_n, rhoU_n, U_n, e_n, rPsi_n, a_n, c_n, p_n, rhoE_n = _unpack(R)
    magSf = float(np.linalg.norm(Sf))
    pv_o, pv_n = float(U_o @ Sf), float(U_n @ Sf)
    cS_o, cS_n = c_o * magSf, c_n * magSf
    ap = max(max(pv_o + cS_o, pv_n + cS_n), 0.0)
    am = min(min(pv_o - cS_o, pv_n - cS_n), 0.0)
    w_o = ap / (ap - am)
    aSf = am * w_o
    w_n = 1.0 - w_o
    ao = pv_o * w_o - aSf
    an = pv_n * w_n + aSf
    phi = ao * rho_o + an * rho_n
    phiU = (ao * rhoU_o + an * rhoU_n) + (w_o * p_o + w_n * p_n) * Sf
    phiE = ao * (rhoE_o + p_o) + an * (rhoE_n + p_n) + aSf * p_o - aSf * p_n
    Up = w_o * U_o + w_n * U_n
    return np.array([phi, *phiU, phiE, *Up])


SCHEMES = ["AUSM+", "AUSM+up", "HLL", "HLLC", "HLLCP", "Kurganov"]


def evaluate(scheme, L, R, Sf, f=1.0, low_mach=0, Mainf=0.5):
    if scheme == "AUSM+":
        return ausm(L, R, Sf)
    if scheme == "AUSM+up":
        return ausm(L, R, Sf, up=True, low_mach=bool(low_mach), Mainf=Mainf)
    if scheme == "HLL":
        return hll(L, R, Sf)
    if scheme == "HLLC":
        return hllc(L, R, Sf)
    if scheme == "HLLCP":
        return hllc(L, R, Sf, f=f)
    return kurganov(L, R, Sf)


def cases(rng):
    """(L, R, Sf, f, low_mach) tuples: random subsonic / supersonic pairs in
    both directions, plus |M| = 1 exactly against the face normal."""
    out = []
    for k in range(48):
        Sf = rng.normal(size=3) * 1e-2
        n = Sf / np.linalg.norm(Sf)
        mach = [0.3, 0.8, 1.6, 2.5][k % 4]
        sides = []
        for _ in range(2):
            rho = rng.uniform(0.5, 2.0)
            p = rng.uniform(0.5e5, 2.0e5)
            c = np.sqrt(GAMMA * p / rho)
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            sgn = 1.0 if k % 8 < 4 else -1.0
            d = 0.5 * d + sgn * n
            d /= np.linalg.norm(d)
            sides.append(side_from_primitive(rho, mach * rng.uniform(0.8, 1.2) * c * d, p))
        f = rng.uniform(0.2, 1.0)
        out.append((sides[0], sides[1], Sf, f, k % 3 == 0))
    # |M| = 1 exactly for AUSM+ (a12 = a* on both sides when the two sides match)
    for sgn in (1.0, -1.0):
        Sf = np.array([1e-2, 0.0, 0.0])
        s = side_from_primitive(1.2, [0.0, 0.0, 0.0], 1.0e5)
        a = s[6]
        u = np.array([sgn * a, 0.0, 0.0])
        L = side_from_primitive(1.2, u, 1.0e5)
        # keep a* of the moving state equal to the one M is measured with
        L[6] = a
        out.append((L, L.copy(), Sf, 1.0, False))
    return out


def main():
    rng = np.random.default_rng(3)
    cs = cases(rng)
    data = {"L": np.array([c[0] for c in cs]), "R": np.array([c[1] for c in cs]),
            "Sf": np.array([c[2] for c in cs]), "f": np.array([c[3] for c in cs]),
            "low_mach": np.array([int(c[4]) for c in cs]), "gamma": np.array(GAMMA)}
    for s in SCHEMES:
        data[s] = np.array([evaluate(s, L, R, Sf, f, lm) for L, R, Sf, f, lm in cs])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "flux_faces.npz")
    np.savez(path, **data)
    print(f"wrote {path}: {len(cs)} faces x {len(SCHEMES)} schemes")


if __name__ == "__main__":
    main()
