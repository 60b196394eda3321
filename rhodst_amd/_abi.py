"""ctypes mirror of include/rhodst.h (structs and enums) and helpers that
turn a PolyMesh + case description into the C descriptors."""
from __future__ import annotations

import ctypes as C

import numpy as np

# rdst_flux_scheme, in the order of the baseFluxDST RTS table (newbaseFluxDST.C:64-84)
FLUX_SCHEMES = {"AUSM+": 0, "AUSM+up": 1, "HLL": 2, "HLLC": 3, "HLLCP": 4, "Kurganov": 5}
BC = {"fixedValue": 0, "zeroGradient": 1, "inletOutlet": 2, "symmetryPlane": 3, "slip": 3,
      "symmetry": 3, "farField": 4, "empty": 5}
SOLVERS = {"GMRES": 0, "BiCGStab": 1}
PRECONDS = {"ILUC0": 0, "ILU0": 0, "none": 1, "diagonal": 2, "Diagonal": 2}
DEPENDENT = {"W": 0, "deltaW": 1}
PRECISION = {"fp64": 0, "mixed": 1}

F_RHO, F_RHOU, F_RHOE, F_P, F_T, F_U, F_H, F_PSI, F_R, F_W, F_RDELTAT, F_COURANT = range(12)
F_PHI, F_PHIUP, F_PHIEP, F_UP = 12, 13, 14, 15
F_BRHO, F_BU, F_BP, F_BT, F_BRHOU, F_BRHOE, F_BH, F_BPSI = 16, 17, 18, 19, 20, 21, 22, 23
F_BCSR_VALUES, F_SOURCE, F_MU, F_RHO_OLD, F_E_OLD = 24, 25, 26, 27, 28


class rdst_patch(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("type", C.c_int32), ("start", C.c_int32),
                ("size", C.c_int32), ("pad", C.c_int32)]


class rdst_mesh_desc(C.Structure):
    _fields_ = [("n_points", C.c_int32), ("n_faces", C.c_int32), ("n_internal_faces", C.c_int32),
                ("n_cells", C.c_int32), ("n_patches", C.c_int32), ("pad", C.c_int32),
                ("points", C.POINTER(C.c_double)), ("face_offsets", C.POINTER(C.c_int32)),
                ("face_points", C.POINTER(C.c_int32)), ("owner", C.POINTER(C.c_int32)),
                ("neighbour", C.POINTER(C.c_int32)), ("patches", C.POINTER(rdst_patch))]


class rdst_bc(C.Structure):
    _fields_ = [("p_type", C.c_int32), ("T_type", C.c_int32), ("U_type", C.c_int32), ("pad", C.c_int32),
                ("p_value", C.c_double), ("T_value", C.c_double), ("U_value", C.c_double * 3),
                ("p_inf", C.c_double), ("T_inf", C.c_double), ("U_inf", C.c_double * 3)]


class rdst_thermo(C.Structure):
    _fields_ = [("Cp", C.c_double), ("R", C.c_double), ("transport", C.c_int32), ("pad", C.c_int32),
                ("mu", C.c_double), ("Pr", C.c_double), ("As", C.c_double), ("Ts", C.c_double)]


class rdst_settings(C.Structure):
    _fields_ = [("flux_scheme", C.c_int32), ("low_mach", C.c_int32), ("Mainf", C.c_double),
                ("dependent", C.c_int32), ("neighbouring", C.c_int32), ("beta", C.c_double),
                ("initial_co", C.c_double), ("min_co", C.c_double), ("max_co", C.c_double),
                ("solver", C.c_int32), ("precond", C.c_int32), ("n_directions", C.c_int32),
                ("min_iter", C.c_int32), ("max_iter", C.c_int32), ("precision", C.c_int32),
                ("tolerance", C.c_double), ("rel_tol", C.c_double),
                ("rho_min", C.c_double), ("rho_max", C.c_double), ("p_min", C.c_double),
                ("p_max", C.c_double), ("T_min", C.c_double), ("T_max", C.c_double),
                ("ilu_order", C.c_int32), ("viscous", C.c_int32), ("gram_schmidt", C.c_int32),
                ("reconstruction", C.c_int32), ("jacobian", C.c_int32),
                ("time_integration", C.c_int32), ("adjust_time_step", C.c_int32), ("max_inner_iter", C.c_int32),
                ("iteration_start", C.c_int32), ("mag_converged_slope", C.c_double), ("min_residual", C.c_double),
                ("delta_t", C.c_double), ("laplacian_sngrad", C.c_int32), ("sngrad", C.c_int32),
                ("laplacian_limit", C.c_double), ("sngrad_limit", C.c_double), ("heat_conduction", C.c_int32),
                ("update_limit", C.c_double), ("grad_u", C.c_int32), ("grad_u_pad", C.c_int32),
                ("grad_u_limit", C.c_double)]


class rdst_solver_perf(C.Structure):
    _fields_ = [("initial_residual", C.c_double * 5), ("final_residual", C.c_double * 5),
                ("n_iterations", C.c_int32), ("converged", C.c_int32)]


class rdst_turbulence(C.Structure):
    _fields_ = [("model", C.c_int32), ("negative_nu_tilda", C.c_int32), ("stilda_modification", C.c_int32),
                ("rcm_correction", C.c_int32),
                ("sigma_nut", C.c_double), ("kappa", C.c_double), ("Prt", C.c_double),
                ("Cb1", C.c_double), ("Cb2", C.c_double), ("Cw2", C.c_double), ("Cw3", C.c_double),
                ("Cv1", C.c_double), ("Cv2", C.c_double),
                ("Cn1", C.c_double), ("Cn2", C.c_double), ("Cn3", C.c_double), ("Ct3", C.c_double),
                ("Ct4", C.c_double),
                ("Cr1", C.c_double), ("Cr2", C.c_double), ("Cr3", C.c_double),
                ("relax", C.c_double), ("tolerance", C.c_double), ("rel_tol", C.c_double),
                ("min_iter", C.c_int32), ("max_iter", C.c_int32), ("convection", C.c_int32), ("pad", C.c_int32)]


class rdst_scalar_bc(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad", C.c_int32), ("value", C.c_double)]


TURB_MODELS = {"laminar": 0, "SAneg": 1, "SARC2": 2, "SARC1": 3, "SpalartAllmaras": 4}


def make_turbulence(model="SAneg", **kw) -> rdst_turbulence:
    """RASProperties <model>Coeffs with the reference's defaults
    (SAneg.C:168-307, SARC2.C:172-285, SARC1.C:196-330: Cr2 = 12) plus the nuTilda solver settings of
    tutorials/RAE2822/system/fvSolution (symGaussSeidel, tolerance 1e-6,
    relTol 0.1, minIter 1, relaxation 0.5); convection "upwind" or "linearUpwind"
    (divSchemes div(phi,nuTilda), the tutorials' Gauss linearUpwind grad(nuTilda))."""
    t = rdst_turbulence()
    t.model = TURB_MODELS[model] if isinstance(model, str) else int(model)
    t.negative_nu_tilda = int(kw.get("negative_nu_tilda", 1))
    t.stilda_modification = int(kw.get("stilda_modification", 1))
    t.rcm_correction = int(kw.get("rcm_correction", 1))
    defaults = dict(sigma_nut=0.66666, kappa=0.41, Prt=1.0, Cb1=0.1355, Cb2=0.622, Cw2=0.3, Cw3=2.0, Cv1=7.1,
                    Cv2=5.0, Cn1=16.0, Cn2=0.7, Cn3=0.9, Ct3=1.2, Ct4=0.5, Cr1=1.0, Cr2=2.0, Cr3=1.0,
                    relax=0.5, tolerance=1e-6, rel_tol=0.1)
    if t.model == 3:
        defaults["Cr2"] = 12.0
    for k, v in defaults.items():
        setattr(t, k, kw.get(k, v))
    t.min_iter = int(kw.get("min_iter", 1))
    t.max_iter = int(kw.get("max_iter", 1000))
    conv = kw.get("convection", "upwind")
    t.convection = {"upwind": 0, "linearUpwind": 1}[conv] if isinstance(conv, str) else int(conv)
    return t


def make_scalar_bcs(mesh, spec: dict):
    """spec: patch name -> (type, value) or type for 0/nuTilda; empty patches
    default to empty."""
    arr = (rdst_scalar_bc * len(mesh.patches))()
    for i, (name, t, _s, _n) in enumerate(mesh.patches):
        if t == 3:
            arr[i].type = BC["empty"]
            continue
        d = spec[name]
        typ, val = d if isinstance(d, tuple) else (d, 0.0)
        arr[i].type = BC[typ]
        arr[i].value = val
    return arr


class rdst_adjoint_settings(C.Structure):
    _fields_ = [("precision", C.c_int32), ("n_directions", C.c_int32), ("max_cycles", C.c_int32),
                ("inner_operator", C.c_int32),
                ("tolerance", C.c_double), ("rel_tol", C.c_double), ("precond_shift", C.c_double)]


def make_adjoint_settings(precision="fp64", n_directions=30, max_cycles=200, tolerance=1e-12, rel_tol=0.0,
                          precond_shift=0.0, inner_operator=0):
    s = rdst_adjoint_settings()
    s.inner_operator = int(inner_operator)
    s.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
    s.n_directions, s.max_cycles = int(n_directions), int(max_cycles)
    s.tolerance, s.rel_tol = float(tolerance), float(rel_tol)
    s.precond_shift = float(precond_shift)
    return s


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class rdst_mesh_geometry(C.Structure):
    _fields_ = [("Sf", C.POINTER(C.c_double)), ("mag_sf", C.POINTER(C.c_double)), ("Cf", C.POINTER(C.c_double)),
                ("C", C.POINTER(C.c_double)), ("V", C.POINTER(C.c_double)), ("weights", C.POINTER(C.c_double)),
                ("delta_coeffs", C.POINTER(C.c_double))]


class MeshGeometry:
    """Owns the arrays behind an rdst_mesh_geometry: the host fvMesh's Sf,
    magSf (or None), Cf, C, V, weights and deltaCoeffs in polyMesh face
    order (internal faces, then every patch's faces)."""

    def __init__(self, Sf, Cf, C_, V, weights, delta_coeffs, mag_sf=None):
        self.arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in (Sf, Cf, C_, V, weights, delta_coeffs)]
        self.mag = None if mag_sf is None else np.ascontiguousarray(mag_sf, dtype=np.float64)
        Sf_, Cf_, C2, V_, w_, dc_ = self.arrays
        self.geom = rdst_mesh_geometry(Sf=dptr(Sf_), mag_sf=dptr(self.mag) if self.mag is not None else None,
                                       Cf=dptr(Cf_), C=dptr(C2), V=dptr(V_), weights=dptr(w_), delta_coeffs=dptr(dc_))


class MeshDesc:
    """Owns the contiguous arrays behind an rdst_mesh_desc."""

    def __init__(self, mesh):
        self.points = np.ascontiguousarray(mesh.points, dtype=np.float64)
        self.face_offsets = np.ascontiguousarray(mesh.face_offsets, dtype=np.int32)
        self.face_points = np.ascontiguousarray(mesh.face_points, dtype=np.int32)
        self.owner = np.ascontiguousarray(mesh.owner, dtype=np.int32)
        self.neighbour = np.ascontiguousarray(mesh.neighbour, dtype=np.int32)
        self.patches = (rdst_patch * len(mesh.patches))()
        for i, (name, t, start, size) in enumerate(mesh.patches):
            self.patches[i].name = name.encode()[:63]
            self.patches[i].type = t
            self.patches[i].start = start
            self.patches[i].size = size
        self.desc = rdst_mesh_desc(
            n_points=len(self.points), n_faces=len(self.owner), n_internal_faces=len(self.neighbour),
            n_cells=mesh.n_cells, n_patches=len(mesh.patches), pad=0,
            points=dptr(self.points), face_offsets=iptr(self.face_offsets),
            face_points=iptr(self.face_points), owner=iptr(self.owner),
            neighbour=iptr(self.neighbour), patches=C.cast(self.patches, C.POINTER(rdst_patch)))


def make_bcs(mesh, spec: dict):
    """spec: patch name -> dict(p=(type, value), T=(type, value), U=(type, (x,y,z)),
    inf=dict(p=, T=, U=)).  Empty patches default to empty BCs."""
    arr = (rdst_bc * len(mesh.patches))()
    for i, (name, t, _s, _n) in enumerate(mesh.patches):
        b = arr[i]
        if t == 3:
            b.p_type = b.T_type = b.U_type = BC["empty"]
            continue
        d = spec[name]
        pt, pv = d["p"] if isinstance(d["p"], tuple) else (d["p"], 0.0)
        tt, tv = d["T"] if isinstance(d["T"], tuple) else (d["T"], 0.0)
        ut, uv = d["U"] if isinstance(d["U"], tuple) else (d["U"], (0.0, 0.0, 0.0))
        b.p_type, b.T_type, b.U_type = BC[pt], BC[tt], BC[ut]
        b.p_value, b.T_value = pv, tv
        for k in range(3):
            b.U_value[k] = uv[k]
        inf = d.get("inf")
        if inf:
            b.p_inf, b.T_inf = inf["p"], inf["T"]
            for k in range(3):
                b.U_inf[k] = inf["U"][k]
    return arr


def make_settings(**kw) -> rdst_settings:
    s = rdst_settings()
    s.flux_scheme = FLUX_SCHEMES[kw.get("flux_scheme", "AUSM+")] if isinstance(kw.get("flux_scheme", "AUSM+"), str) else kw["flux_scheme"]
    s.low_mach = int(kw.get("low_mach", 0))
    s.Mainf = kw.get("Mainf", 0.5)
    s.dependent = DEPENDENT[kw.get("dependent", "W")]
    s.neighbouring = int(kw.get("neighbouring", 1))
    s.beta = kw.get("beta", 1.0)
    s.initial_co = kw.get("initial_co", 5.0)
    s.min_co = kw.get("min_co", 5.0)
    s.max_co = kw.get("max_co", 5.0)
    s.solver = SOLVERS[kw.get("solver", "GMRES")]
    s.precond = PRECONDS[kw.get("precond", "ILUC0")]
    s.n_directions = kw.get("n_directions", 20)
    s.min_iter = kw.get("min_iter", 1)
    s.max_iter = kw.get("max_iter", 1)
    s.precision = PRECISION[kw.get("precision", "fp64")]
    s.tolerance = kw.get("tolerance", 1e-15)
    s.rel_tol = kw.get("rel_tol", 1e-4)
    s.rho_min, s.rho_max = kw.get("rho_bounds", (1e-3, 1e4))
    s.p_min, s.p_max = kw.get("p_bounds", (1e-3, 1e4))
    s.T_min, s.T_max = kw.get("T_bounds", (1e-3, 1e4))
    order = kw.get("ilu_order", "natural")
    s.ilu_order = {"natural": 0, "partitioned": 1}[order] if isinstance(order, str) else int(order)
    s.viscous = int(kw.get("viscous", 0))
    gs = kw.get("gram_schmidt", "modified")
    s.gram_schmidt = {"modified": 0, "classical": 1}[gs] if isinstance(gs, str) else int(gs)
    rc = kw.get("reconstruction", "upwind")
    s.reconstruction = {"upwind": 0, "OSPRE": 1}[rc] if isinstance(rc, str) else int(rc)
    jk = kw.get("jacobian", "Roe")
    s.jacobian = {"Roe": 0, "exact": 1}[jk] if isinstance(jk, str) else int(jk)
    ti = kw.get("time_integration", "Single")
    s.time_integration = {"Single": 0, "Dual": 1}[ti] if isinstance(ti, str) else int(ti)
    s.adjust_time_step = int(kw.get("adjust_time_step", 1))
    s.max_inner_iter = int(kw.get("max_inner_iter", 20))
    s.iteration_start = int(kw.get("iteration_start", 10 if s.time_integration == 1 else 500))
    s.mag_converged_slope = float(kw.get("mag_converged_slope", 0.1))
    s.min_residual = float(kw.get("min_residual", 1e-5))
    s.delta_t = float(kw.get("delta_t", 0.0))
    # fvSchemes laplacianSchemes / snGradSchemes: "corrected", "uncorrected"
    # or ("limited", psi)
    s.laplacian_sngrad, s.laplacian_limit = _sngrad(kw.get("laplacian", "corrected"))
    s.sngrad, s.sngrad_limit = _sngrad(kw.get("sngrad", "corrected"))
    s.heat_conduction = int(kw.get("heat_conduction", 0))
    s.update_limit = float(kw.get("update_limit", 0.0))
    # gradSchemes grad(U): "Gauss linear" or ("cellLimited", k)
    s.grad_u, s.grad_u_limit = _grad_scheme(kw.get("grad_u", "Gauss linear"))
    return s


def _grad_scheme(spec):
    if isinstance(spec, str):
        if spec != "Gauss linear":
            raise ValueError("grad(U) scheme: 'Gauss linear' or ('cellLimited', k)")
        return 0, 0.0
    kind, k = spec
    if kind != "cellLimited":
        raise ValueError("grad(U) scheme: 'Gauss linear' or ('cellLimited', k)")
    return 1, float(k)


SNGRAD = {"corrected": 0, "uncorrected": 1, "limited": 2}


def _sngrad(spec):
    if isinstance(spec, str):
        return SNGRAD[spec], 1.0
    kind, psi = spec
    return SNGRAD[kind], float(psi)


def make_thermo(Cp=2.5, R=1.0 / 1.4, mu=0.0, Pr=1.0, sutherland=None) -> rdst_thermo:
    t = rdst_thermo()
    t.Cp, t.R, t.mu, t.Pr = Cp, R, mu, Pr
    if sutherland:
        t.transport = 1
        t.As, t.Ts = sutherland
    return t
