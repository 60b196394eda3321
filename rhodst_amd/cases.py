"""Case descriptions (mesh + thermo + BCs + settings + initial state) for
the workloads the block-coupled implicit step is run on."""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from . import meshgen
from ._abi import make_bcs, make_settings, make_thermo


@dataclass
class CaseSpec:
    name: str
    mesh: meshgen.PolyMesh
    thermo: object
    bcs: object
    settings: object
    p0: np.ndarray
    T0: np.ndarray
    U0: np.ndarray
    # Spalart-Allmaras: (rdst_turbulence, rdst_scalar_bc[n_patches], wall
    # distance [n_cells], initial nuTilda [n_cells]) or None (laminar)
    turbulence: tuple | None = None


def forward_step_case(nx: int = 120, ny: int = 40, scheme: str = "AUSM+", perturb: float = 0.0,
                      seed: int = 2, **settings) -> CaseSpec:
    """Mach-3 forward step (tutorials/forwardStep/0/{p,T,U}, thermo
    constant/thermophysicalProperties 'normalisedGas', Cp 2.5, mu 0), with R
    set to 1/1.4 so that c = 1 at T = 1: inflow p = 1, T = 1, U = (3,0,0).
    inlet fixedValue; outlet p zeroGradient, T/U inletOutlet; bottom/top
    symmetryPlane; obstacle p/T zeroGradient, U slip.  Uniform initial state,
    optionally multiplied by a seeded +-perturb factor (seed 2)."""
    mesh = meshgen.forward_step(nx, ny)
    spec = {
        "inlet": {"p": ("fixedValue", 1.0), "T": ("fixedValue", 1.0), "U": ("fixedValue", (3.0, 0.0, 0.0))},
        "outlet": {"p": "zeroGradient", "T": ("inletOutlet", 1.0), "U": ("inletOutlet", (3.0, 0.0, 0.0))},
        "bottom": {"p": "symmetryPlane", "T": "symmetryPlane", "U": "symmetryPlane"},
        "top": {"p": "symmetryPlane", "T": "symmetryPlane", "U": "symmetryPlane"},
        "obstacle": {"p": "zeroGradient", "T": "zeroGradient", "U": "slip"},
    }
    n = mesh.n_cells
    p0 = np.ones(n)
    T0 = np.ones(n)
    U0 = np.zeros((n, 3))
    U0[:, 0] = 3.0
    if perturb > 0:
        rng = np.random.default_rng(seed)
        p0 = p0 * (1 + perturb * rng.uniform(-1, 1, n))
        T0 = T0 * (1 + perturb * rng.uniform(-1, 1, n))
        U0[:, 0] *= 1 + perturb * rng.uniform(-1, 1, n)
        U0[:, 1] = 3.0 * perturb * rng.uniform(-1, 1, n)
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=5.0, min_co=5.0, max_co=5.0,
              solver="GMRES", precond="ILUC0", n_directions=20, min_iter=1, max_iter=1,
              tolerance=1e-15, rel_tol=1e-4)
    kw.update(settings)
    return CaseSpec("forwardStep", mesh, make_thermo(Cp=2.5, R=1.0 / 1.4, mu=0.0, Pr=1.0),
                    make_bcs(mesh, spec), make_settings(**kw), p0, T0, np.ascontiguousarray(U0))


def box_farfield_case(nx: int = 24, ny: int = 16, scheme: str = "HLLC", mach: float = 0.5, aoa_deg: float = 5.0,
                      jitter: float = 0.2, seed: int = 1, perturb: float = 0.05, **settings) -> CaseSpec:
    """A jittered 2-D box with the characteristic far-field conditions
    (farField{Pressure,Temperature,Velocity}DST) on all four sides and a
    seeded perturbed subsonic initial state; exercises every far-field regime
    the flow produces.  Air-like thermo, inviscid."""
    mesh = meshgen.box_2d(nx, ny, jitter=jitter, seed=seed)
    gamma, R = 1.4, 287.0
    Cp = gamma * R / (gamma - 1)
    Tinf, pinf = 300.0, 1.0e5
    c = np.sqrt(gamma * R * Tinf)
    a = np.deg2rad(aoa_deg)
    Uinf = (mach * c * np.cos(a), mach * c * np.sin(a), 0.0)
    ff = {"p": ("farField", pinf), "T": ("farField", Tinf), "U": ("farField", Uinf),
          "inf": {"p": pinf, "T": Tinf, "U": Uinf}}
    spec = {k: ff for k in ("left", "right", "bottom", "top")}
    n = mesh.n_cells
    rng = np.random.default_rng(2)
    p0 = pinf * (1 + perturb * rng.uniform(-1, 1, n))
    T0 = Tinf * (1 + perturb * rng.uniform(-1, 1, n))
    U0 = np.tile(np.array(Uinf), (n, 1)) * (1 + perturb * rng.uniform(-1, 1, (n, 1)))
    U0[:, 2] = 0.0
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=2.0, min_co=0.5, max_co=20.0,
              solver="GMRES", precond="ILUC0", n_directions=10, min_iter=1, max_iter=2,
              tolerance=1e-15, rel_tol=1e-3, p_bounds=(1.0, 1e7), T_bounds=(1.0, 1e4),
              rho_bounds=(1e-4, 1e3))
    kw.update(settings)
    return CaseSpec("boxFarField", mesh, make_thermo(Cp=Cp, R=R, mu=0.0, Pr=0.72),
                    make_bcs(mesh, spec), make_settings(**kw), p0, T0, np.ascontiguousarray(U0))


def swept_wing_case(scheme: str = "AUSM+up", mach: float = 0.84, aoa_deg: float = 3.06, viscous: bool = False,
                    mesh=None, mesh_kw: dict | None = None, turbulence: str | None = None, reynolds: float = 1.2e7,
                    wall_dist=None, turb_kw: dict | None = None, **settings) -> CaseSpec:
    """Transonic swept wing: semi-span 1, 30 deg leading-edge sweep, taper
    0.56, 10 % analytic section, C-H hex mesh (meshgen.swept_wing, 2.03 M
    cells by default), Mach 0.84 at 3.06 deg, characteristic far field,
    symmetry plane at the root.  Air (R 287, gamma 1.4, T 288.15 K,
    p 101325 Pa), uniform free-stream start.  Inviscid: slip wing.
    turbulence="SARC2"/"SAneg": no-slip wing, viscous terms on, a constant
    mu giving Re = reynolds on the root chord, nuTilda fixedValue 0 on the
    wing, inletOutlet 3 nu on the far field, the exact wall distance
    (meshgen.wall_distance, or wall_dist when given) and nuTilda = 3 nu."""
    mesh = mesh if mesh is not None else meshgen.swept_wing(**(mesh_kw or {}))
    gamma, R = 1.4, 287.0
    Cp = gamma * R / (gamma - 1)
    Tinf, pinf = 288.15, 101325.0
    c = np.sqrt(gamma * R * Tinf)
    a = np.deg2rad(aoa_deg)
    Uinf = (mach * c * np.cos(a), mach * c * np.sin(a), 0.0)
    ff = {"p": ("farField", pinf), "T": ("farField", Tinf), "U": ("farField", Uinf),
          "inf": {"p": pinf, "T": Tinf, "U": Uinf}}
    if turbulence:
        viscous = True
    wall = {"p": "zeroGradient", "T": "zeroGradient", "U": ("fixedValue", (0.0, 0.0, 0.0)) if viscous else "slip"}
    spec = {"wing": wall, "symmetry": {"p": "symmetryPlane", "T": "symmetryPlane", "U": "symmetryPlane"},
            "farfield": ff}
    n = mesh.n_cells
    p0 = np.full(n, pinf)
    T0 = np.full(n, Tinf)
    U0 = np.tile(np.array(Uinf), (n, 1))
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=1.0, min_co=0.5, max_co=20.0,
              solver="GMRES", precond="ILUC0", n_directions=30, min_iter=1, max_iter=1,
              tolerance=1e-15, rel_tol=1e-2, p_bounds=(100.0, 1e7), T_bounds=(50.0, 2000.0),
              rho_bounds=(1e-3, 100.0))
    kw.update(settings)
    rho_inf = pinf / (R * Tinf)
    mu = rho_inf * mach * c / reynolds if viscous else 0.0
    kw.setdefault("viscous", 1 if viscous else 0)
    spec_out = CaseSpec("sweptWing", mesh, make_thermo(Cp=Cp, R=R, mu=mu, Pr=0.72),
                        make_bcs(mesh, spec), make_settings(**kw), p0, T0, np.ascontiguousarray(U0))
    if turbulence:
        from ._abi import make_scalar_bcs, make_turbulence
        nu = mu / rho_inf
        sbc = make_scalar_bcs(mesh, {"wing": ("fixedValue", 0.0), "symmetry": "symmetryPlane",
                                     "farfield": ("inletOutlet", 3.0 * nu)})
        d = wall_dist if wall_dist is not None else meshgen.wall_distance(mesh, ["wing"])
        spec_out.turbulence = (make_turbulence(turbulence, **(turb_kw or {})), sbc, d, np.full(n, 3.0 * nu))
        spec_out.name = f"sweptWing_{turbulence}"
    return spec_out


def airfoil_case(scheme: str = "HLLC", mach: float = 0.73, aoa_deg: float = 2.3, reynolds: float = 6.5e6,
                 turbulence: str | None = "SAneg", mesh=None, mesh_kw: dict | None = None, wall_dist=None,
                 **settings) -> CaseSpec:
    """Transonic viscous flow over the 12 % section on the hybrid quad/triangle
    mesh (meshgen.hybrid_airfoil, about 2e5 cells), Mach 0.73 at 2.3 deg,
    Re = reynolds on the unit chord: air (R 287, gamma 1.4, T 288.15 K,
    p 101325 Pa), a constant mu, no-slip adiabatic wall, characteristic far
    field at 50 chords, frontAndBack empty; uniform free-stream start.
    turbulence "SAneg"/"SARC2": nuTilda fixedValue 0 on the wall, inletOutlet
    3 nu on the far field, the exact wall distance (or wall_dist) and
    nuTilda = 3 nu; None: laminar."""
    mesh = mesh if mesh is not None else meshgen.hybrid_airfoil(**(mesh_kw or {}))
    gamma, R = 1.4, 287.0
    Cp = gamma * R / (gamma - 1)
    Tinf, pinf = 288.15, 101325.0
    c = np.sqrt(gamma * R * Tinf)
    a = np.deg2rad(aoa_deg)
    Uinf = (mach * c * np.cos(a), mach * c * np.sin(a), 0.0)
    ff = {"p": ("farField", pinf), "T": ("farField", Tinf), "U": ("farField", Uinf),
          "inf": {"p": pinf, "T": Tinf, "U": Uinf}}
    wall = {"p": "zeroGradient", "T": "zeroGradient", "U": ("fixedValue", (0.0, 0.0, 0.0))}
    n = mesh.n_cells
    rho_inf = pinf / (R * Tinf)
    mu = rho_inf * mach * c / reynolds
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=1.0, min_co=0.5, max_co=20.0, viscous=1,
              solver="GMRES", precond="ILUC0", n_directions=30, min_iter=1, max_iter=1,
              tolerance=1e-15, rel_tol=1e-2, p_bounds=(100.0, 1e7), T_bounds=(50.0, 2000.0),
              rho_bounds=(1e-3, 100.0))
    kw.update(settings)
    spec_out = CaseSpec("airfoil", mesh, make_thermo(Cp=Cp, R=R, mu=mu, Pr=0.72),
                        make_bcs(mesh, {"airfoil": wall, "farfield": ff}), make_settings(**kw),
                        np.full(n, pinf), np.full(n, Tinf), np.ascontiguousarray(np.tile(np.array(Uinf), (n, 1))))
    if turbulence:
        from ._abi import make_scalar_bcs, make_turbulence
        nu = mu / rho_inf
        sbc = make_scalar_bcs(mesh, {"airfoil": ("fixedValue", 0.0), "farfield": ("inletOutlet", 3.0 * nu)})
        d = wall_dist if wall_dist is not None else meshgen.wall_distance(mesh, ["airfoil"])
        spec_out.turbulence = (make_turbulence(turbulence), sbc, d, np.full(n, 3.0 * nu))
        spec_out.name = f"airfoil_{turbulence}"
    return spec_out


RAE2822_MESH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden",
                            "rae2822_polymesh.npz")


def rae2822_case(scheme: str = "Kurganov", turbulence: str | None = "SpalartAllmaras", mesh=None, wall_dist=None,
                 **settings) -> CaseSpec:
    """The reference's RAE2822 tutorial (tutorials/RAE2822): its polyMesh
    (23 552 cells; patches frontAndBackPlanes empty, freestreamBC, airfoil
    wall), free stream 0/include/freestreamConditions (U (233.44 9.417 0),
    p 108987.773, T 255.556) with the farField*DST conditions, no-slip
    adiabatic airfoil, hPsiThermo with Sutherland transport (W 28.966,
    Cp 1005, As 1.4584e-6, Ts 110.33; R = RR/W with RR = 8314.47), rhoDSTDict
    (Kurganov, neighbouring, CFL 1 -> 500), fvSchemes (OSPRE/OSPREV
    reconstruction), fvSolution (BiCGStab + ILUC0, maxIter 100, tolerance
    and relTol 1e-10; fieldBounds rho 0.01 10, p 100 1e6, T 10 800) and the
    SpalartAllmaras model with 0/nuTilda (4.4e-5 inside and at the far field,
    0 on the wall), its div(phi,nuTilda) Gauss linearUpwind grad(nuTilda)
    and grad(U) cellLimited Gauss linear 1 (fvSchemes:27).  The mesh is read
    from the committed array copy of the tutorial's polyMesh."""
    mesh = mesh if mesh is not None else meshgen.polymesh_from_npz(RAE2822_MESH)
    Uinf, pinf, Tinf = (233.44, 9.417, 0.0), 108987.773, 255.556
    R = 8314.47 / 28.966
    ff = {"p": ("farField", pinf), "T": ("farField", Tinf), "U": ("farField", Uinf),
          "inf": {"p": pinf, "T": Tinf, "U": Uinf}}
    wall = {"p": "zeroGradient", "T": "zeroGradient", "U": ("fixedValue", (0.0, 0.0, 0.0))}
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=1.0, min_co=1.0, max_co=500.0, viscous=1,
              solver="BiCGStab", precond="ILUC0", min_iter=1, max_iter=100, tolerance=1e-10, rel_tol=1e-10,
              reconstruction="OSPRE", p_bounds=(100.0, 1e6), T_bounds=(10.0, 800.0), rho_bounds=(0.01, 10.0),
              grad_u=("cellLimited", 1.0))
    kw.update(settings)
    n = mesh.n_cells
    spec = CaseSpec("RAE2822", mesh, make_thermo(Cp=1005.0, R=R, Pr=0.72, sutherland=(1.4584e-6, 110.33)),
                    make_bcs(mesh, {"freestreamBC": ff, "airfoil": wall}), make_settings(**kw),
                    np.full(n, pinf), np.full(n, Tinf), np.ascontiguousarray(np.tile(np.array(Uinf), (n, 1))))
    if turbulence:
        from ._abi import make_scalar_bcs, make_turbulence
        sbc = make_scalar_bcs(mesh, {"airfoil": ("fixedValue", 0.0), "freestreamBC": ("inletOutlet", 4.4e-5)})
        d = wall_dist if wall_dist is not None else meshgen.wall_distance(mesh, ["airfoil"])
        spec.turbulence = (make_turbulence(turbulence, convection="linearUpwind"), sbc, d, np.full(n, 4.4e-5))
        spec.name = f"RAE2822_{turbulence}"
    return spec


def box_viscous_case(nx: int = 24, ny: int = 16, scheme: str = "HLLC", mach: float = 0.3, jitter: float = 0.15,
                     seed: int = 1, perturb: float = 0.05, wall_stretch: bool = True, **settings) -> CaseSpec:
    """Laminar flow over a no-slip wall: a jittered 2-D box (0.01 x 0.01 m)
    with a wall at the bottom (U fixedValue 0, p and T zeroGradient) and the
    characteristic far field on the other three sides; Sutherland air, seeded
    perturbed state.  Exercises the viscous calcFields terms, the wall
    boundary gradient and the viscous spectral radius."""
    mesh = meshgen.box_2d(nx, ny, lx=0.01, ly=0.01, jitter=jitter, seed=seed)
    gamma, R = 1.4, 287.0
    Cp = gamma * R / (gamma - 1)
    Tinf, pinf = 300.0, 1.0e5
    c = np.sqrt(gamma * R * Tinf)
    Uinf = (mach * c, 0.0, 0.0)
    ff = {"p": ("farField", pinf), "T": ("farField", Tinf), "U": ("farField", Uinf),
          "inf": {"p": pinf, "T": Tinf, "U": Uinf}}
    wall = {"p": "zeroGradient", "T": "zeroGradient", "U": ("fixedValue", (0.0, 0.0, 0.0))}
    spec = {"left": ff, "right": ff, "top": ff, "bottom": wall}
    n = mesh.n_cells
    rng = np.random.default_rng(2)
    p0 = pinf * (1 + perturb * rng.uniform(-1, 1, n))
    T0 = Tinf * (1 + perturb * rng.uniform(-1, 1, n))
    U0 = np.tile(np.array(Uinf), (n, 1)) * (1 + perturb * rng.uniform(-1, 1, (n, 1)))
    U0[:, 1] = 0.1 * Uinf[0] * rng.uniform(-1, 1, n)
    U0[:, 2] = 0.0
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=2.0, min_co=0.5, max_co=20.0, viscous=1,
              solver="GMRES", precond="ILUC0", n_directions=10, min_iter=1, max_iter=2,
              tolerance=1e-15, rel_tol=1e-3, p_bounds=(1.0, 1e7), T_bounds=(1.0, 1e4), rho_bounds=(1e-4, 1e3))
    kw.update(settings)
    return CaseSpec("boxViscous", mesh, make_thermo(Cp=Cp, R=R, Pr=0.72, sutherland=(1.458e-6, 110.4)),
                    make_bcs(mesh, spec), make_settings(**kw), p0, T0, np.ascontiguousarray(U0))


def box_sa_case(model: str = "SAneg", scheme: str = "HLLC", nx: int = 24, ny: int = 16, negative_fraction: float = 0.1,
                seed: int = 2, turb_kw: dict | None = None, **settings) -> CaseSpec:
    """The no-slip-wall box of box_viscous_case with the Spalart-Allmaras model:
    nuTilda fixedValue 0 on the wall, inletOutlet (inletValue 3 nu) on the far
    field, wall distance to the bottom wall, and a seeded initial nuTilda of
    about 3 nu with a fraction of the cells negative (the SAneg branch)."""
    from ._abi import make_scalar_bcs, make_turbulence
    spec = box_viscous_case(scheme=scheme, nx=nx, ny=ny, **settings)
    mesh = spec.mesh
    rng = np.random.default_rng(seed)
    mu = 1.458e-6 * np.sqrt(300.0) / (1.0 + 110.4 / 300.0)
    nu = mu / (1.0e5 / (287.0 * 300.0))
    nut0 = 3.0 * nu * (1.0 + 0.5 * rng.uniform(-1, 1, mesh.n_cells))
    if negative_fraction > 0:
        k = rng.choice(mesh.n_cells, int(negative_fraction * mesh.n_cells), replace=False)
        nut0[k] = -0.3 * nut0[k]
    ff = ("inletOutlet", 3.0 * nu)
    bcs = make_scalar_bcs(mesh, {"left": ff, "right": ff, "top": ff, "bottom": ("fixedValue", 0.0)})
    d = meshgen.wall_distance(mesh, ["bottom"])
    turb = make_turbulence(model, **(turb_kw or {}))
    spec.name = f"boxSA_{model}"
    spec.turbulence = (turb, bcs, d, nut0)
    return spec


def tet_prism_case(model: str | None = "SAneg", scheme: str = "Kurganov", mesh=None, mach: float = 0.3,
                   perturb: float = 0.05, seed: int = 2, wall_dist=None, turb_kw: dict | None = None,
                   **settings) -> CaseSpec:
    """Three-dimensional flow over a no-slip wall on tetrahedra above prism
    layers (meshgen.tet_prism_box, 600 cells by default): the wall at z = 0 (U
    fixedValue 0, p and T zeroGradient), the characteristic far field on the
    other five sides with the free stream along x, Sutherland air, a seeded
    perturbed state, viscous terms on; model "SAneg"/"SARC2"/... adds the
    Spalart-Allmaras model (nuTilda fixedValue 0 on the wall, inletOutlet 3 nu
    elsewhere, the exact wall distance, a seeded nuTilda of about 3 nu),
    None: laminar."""
    mesh = mesh if mesh is not None else meshgen.tet_prism_box(jitter=0.15)
    gamma, R = 1.4, 287.0
    Cp = gamma * R / (gamma - 1)
    Tinf, pinf = 300.0, 1.0e5
    c = np.sqrt(gamma * R * Tinf)
    Uinf = (mach * c, 0.0, 0.0)
    ff = {"p": ("farField", pinf), "T": ("farField", Tinf), "U": ("farField", Uinf),
          "inf": {"p": pinf, "T": Tinf, "U": Uinf}}
    wall = {"p": "zeroGradient", "T": "zeroGradient", "U": ("fixedValue", (0.0, 0.0, 0.0))}
    spec = {"bottom": wall, **{k: ff for k in ("top", "left", "right", "front", "back")}}
    n = mesh.n_cells
    rng = np.random.default_rng(seed)
    p0 = pinf * (1 + perturb * rng.uniform(-1, 1, n))
    T0 = Tinf * (1 + perturb * rng.uniform(-1, 1, n))
    U0 = np.tile(np.array(Uinf), (n, 1)) * (1 + perturb * rng.uniform(-1, 1, (n, 1)))
    U0[:, 1] = 0.1 * Uinf[0] * rng.uniform(-1, 1, n)
    U0[:, 2] = 0.05 * Uinf[0] * rng.uniform(-1, 1, n)
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=2.0, min_co=0.5, max_co=20.0, viscous=1,
              solver="GMRES", precond="ILUC0", n_directions=10, min_iter=1, max_iter=2,
              tolerance=1e-15, rel_tol=1e-3, p_bounds=(1.0, 1e7), T_bounds=(1.0, 1e4), rho_bounds=(1e-4, 1e3))
    kw.update(settings)
    out = CaseSpec("tetPrism", mesh, make_thermo(Cp=Cp, R=R, Pr=0.72, sutherland=(1.458e-6, 110.4)),
                   make_bcs(mesh, spec), make_settings(**kw), p0, T0, np.ascontiguousarray(U0))
    if model:
        from ._abi import make_scalar_bcs, make_turbulence
        mu = 1.458e-6 * np.sqrt(Tinf) / (1.0 + 110.4 / Tinf)
        nu = mu / (pinf / (R * Tinf))
        nut0 = 3.0 * nu * (1.0 + 0.5 * rng.uniform(-1, 1, n))
        sff = ("inletOutlet", 3.0 * nu)
        sbc = make_scalar_bcs(mesh, {"bottom": ("fixedValue", 0.0),
                                     **{k: sff for k in ("top", "left", "right", "front", "back")}})
        d = wall_dist if wall_dist is not None else meshgen.wall_distance(mesh, ["bottom"])
        out.turbulence = (make_turbulence(model, **(turb_kw or {})), sbc, d, nut0)
        out.name = f"tetPrism_{model}"
    return out


def bump_case(amplitudes=(0.02, 0.01, -0.005, 0.004), nx: int = 48, ny: int = 24, mach: float = 0.5,
              scheme: str = "HLLC", node_disp=None, viscous: bool = False, reynolds: float = 5000.0,
              **settings) -> CaseSpec:
    """Subsonic flow through a channel over a Hicks-Henne bump on the bottom
    wall (meshgen.bump_channel): characteristic far-field inlet and outlet
    (subsonic inflow/outflow everywhere), symmetry top.  Air at 300 K / 1e5 Pa.
    Inviscid: slip bottom wall.  viscous: laminar, a no-slip adiabatic bottom
    wall and a constant mu giving Re = reynolds on the channel length.  The
    workload of the discrete-adjoint checks (force coefficients of the bottom
    wall against wall shape)."""
    mesh = meshgen.bump_channel(nx, ny, amplitudes=amplitudes, node_disp=node_disp)
    gamma, R = 1.4, 287.0
    Cp = gamma * R / (gamma - 1)
    Tinf, pinf = 300.0, 1.0e5
    U = (mach * np.sqrt(gamma * R * Tinf), 0.0, 0.0)
    ff = {"p": ("farField", pinf), "T": ("farField", Tinf), "U": ("farField", U),
          "inf": {"p": pinf, "T": Tinf, "U": U}}
    bottom = {"p": "zeroGradient", "T": "zeroGradient",
              "U": ("fixedValue", (0.0, 0.0, 0.0)) if viscous else "slip"}
    spec = {"inlet": ff, "outlet": ff, "bottom": bottom,
            "top": {"p": "symmetryPlane", "T": "symmetryPlane", "U": "symmetryPlane"}}
    mu = (pinf / (R * Tinf)) * U[0] * 2.0 / reynolds if viscous else 0.0
    n = mesh.n_cells
    kw = dict(flux_scheme=scheme, neighbouring=1, initial_co=5.0, min_co=1.0, max_co=1.0e5,
              solver="GMRES", precond="ILUC0", n_directions=30, min_iter=1, max_iter=2,
              tolerance=1e-14, rel_tol=1e-4, p_bounds=(100.0, 1e7), T_bounds=(50.0, 2000.0),
              rho_bounds=(1e-3, 100.0), ilu_order="partitioned", gram_schmidt="classical")
    if viscous:
        kw["viscous"] = 1
    kw.update(settings)
    return CaseSpec("bumpViscous" if viscous else "bump", mesh, make_thermo(Cp=Cp, R=R, mu=mu, Pr=0.72),
                    make_bcs(mesh, spec), make_settings(**kw), np.full(n, pinf), np.full(n, Tinf),
                    np.tile(np.array(U), (n, 1)))
