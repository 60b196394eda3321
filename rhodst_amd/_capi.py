"""ctypes binding of librhodst_amd.so (the C-ABI of include/rhodst.h).

This is the only way the package reaches the step: there is no CPU
fallback.  Loading fails with a RuntimeError when the shared library has not
been built, and rdst_init fails with RDST_ERR_NO_DEVICE when no GPU is
visible.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import (MeshDesc, MeshGeometry, dptr, iptr, rdst_mesh_geometry, make_adjoint_settings, rdst_adjoint_settings, rdst_bc, rdst_mesh_desc,
                   rdst_scalar_bc, rdst_settings, rdst_solver_perf, rdst_thermo, rdst_turbulence)

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "librhodst_amd.so")

RDST_OK, RDST_ERR_ARG, RDST_ERR_HIP, RDST_ERR_UNKNOWN_NAME, RDST_ERR_STATE, RDST_ERR_NO_DEVICE = range(6)

# every entry point declared in include/rhodst.h
EXPORTS = ["rdst_last_error", "rdst_version", "rdst_init", "rdst_mesh_create", "rdst_mesh_create_geometry",
           "rdst_mesh_destroy",
           "rdst_mesh_info", "rdst_mesh_get_int", "rdst_mesh_get_geom", "rdst_mesh_wall_distance",
           "rdst_case_create",
           "rdst_case_destroy", "rdst_case_set_state", "rdst_case_get", "rdst_case_put",
           "rdst_flux_residual", "rdst_jacobian_assemble", "rdst_block_solve", "rdst_ldu_solve", "rdst_update_fields",
           "rdst_set_courant", "rdst_iterate", "rdst_dual_step", "rdst_time_info", "rdst_case_set_turbulence", "rdst_turbulence_correct",
           "rdst_adjoint_linearise", "rdst_adjoint_apply", "rdst_adjoint_residual", "rdst_adjoint_objective_cells",
           "rdst_adjoint_objective_gradient", "rdst_adjoint_solve",
           "rdst_forces", "rdst_forces_noalloc", "rdst_last_flux_ms", "rdst_time_flux"]

_lib = None


class RdstError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rhodst error {code}: {msg}")
        self.code = code


def lib():
    """Load the HIP extension; raise if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP extension not built: {LIB_PATH} missing (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.rdst_last_error.restype = C.c_char_p
    L.rdst_init.argtypes = [C.c_int]
    L.rdst_mesh_create.argtypes = [P(rdst_mesh_desc), P(C.c_void_p)]
    L.rdst_mesh_create_geometry.argtypes = [P(rdst_mesh_desc), P(rdst_mesh_geometry), P(C.c_void_p)]
    L.rdst_mesh_destroy.argtypes = [C.c_void_p]
    L.rdst_mesh_info.argtypes = [C.c_void_p, P(C.c_int64)]
    L.rdst_mesh_get_int.argtypes = [C.c_void_p, C.c_int, P(C.c_int32)]
    L.rdst_mesh_get_geom.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
    L.rdst_case_create.argtypes = [C.c_void_p, P(rdst_thermo), P(rdst_bc), P(rdst_settings), P(C.c_void_p)]
    L.rdst_case_destroy.argtypes = [C.c_void_p]
    L.rdst_case_set_state.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double)]
    L.rdst_case_get.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
    L.rdst_case_put.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
    for name in ("rdst_flux_residual", "rdst_jacobian_assemble", "rdst_update_fields"):
        getattr(L, name).argtypes = [C.c_void_p]
    L.rdst_block_solve.argtypes = [C.c_void_p, P(rdst_solver_perf)]
    L.rdst_ldu_solve.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double),
                                 P(C.c_double), C.c_int, P(rdst_solver_perf)]
    L.rdst_set_courant.argtypes = [C.c_void_p, P(C.c_double)]
    L.rdst_iterate.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(rdst_solver_perf)]
    L.rdst_dual_step.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_double), P(C.c_double)]
    L.rdst_time_info.argtypes = [C.c_void_p, P(C.c_double)]
    L.rdst_forces.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double), C.c_double, C.c_double,
                              C.c_double, C.c_double, P(C.c_double)]
    L.rdst_forces_noalloc.argtypes = L.rdst_forces.argtypes
    L.rdst_case_set_turbulence.argtypes = [C.c_void_p, P(rdst_turbulence), P(rdst_scalar_bc), P(C.c_double),
                                           P(C.c_double)]
    L.rdst_turbulence_correct.argtypes = [C.c_void_p, P(rdst_solver_perf)]
    L.rdst_adjoint_linearise.argtypes = [C.c_void_p]
    L.rdst_adjoint_apply.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double)]
    L.rdst_adjoint_residual.argtypes = [C.c_void_p, P(C.c_double)]
    L.rdst_adjoint_objective_cells.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_double, C.c_double,
                                               P(C.c_double)]
    L.rdst_adjoint_objective_gradient.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_double, P(C.c_double)]
    L.rdst_adjoint_solve.argtypes = [C.c_void_p, C.c_int, P(C.c_double), C.c_double, P(rdst_adjoint_settings),
                                     P(C.c_double), P(rdst_solver_perf)]
    L.rdst_last_flux_ms.argtypes = [C.c_void_p, P(C.c_double)]
    L.rdst_time_flux.argtypes = [C.c_void_p, C.c_int, P(C.c_double)]
    _lib = L
    return L


def check(rc):
    if rc != RDST_OK:
        raise RdstError(rc, lib().rdst_last_error().decode())


_initialised = False


def init(device: int = 0):
    global _initialised
    if not _initialised:
        check(lib().rdst_init(device))
        _initialised = True


INFO_KEYS = ["n_cells", "n_internal", "n_bfaces", "nnzb", "n_colours", "n_levels_l", "n_levels_u", "n_cell_colours",
             "n_parts", "max_part_levels"]
M_OWNER, M_NEIGHBOUR, M_ROWPTR, M_COLS, M_SLOT_UPPER, M_SLOT_LOWER, M_SLOT_DIAG, M_FACE_COLOUR, M_ILU_LEVEL, \
    M_BFACE_PATCH, M_CELL_COLOUR, M_PARTITION, M_SA_PARTITION = range(13)
G_SF, G_MAGSF, G_CF, G_C, G_V, G_WEIGHTS, G_DELTACOEFFS = range(7)


class DeviceMesh:
    """A mesh uploaded to the GPU (rdst_mesh_create)."""

    def __init__(self, polymesh, geometry: MeshGeometry | None = None):
        """geometry: the host fvMesh's own geometry (rdst_mesh_create_geometry),
        kept bit for bit; None: computed from the points (rdst_mesh_create)."""
        self.polymesh = polymesh
        self._md = MeshDesc(polymesh)
        h = C.c_void_p()
        if geometry is None:
            check(lib().rdst_mesh_create(C.byref(self._md.desc), C.byref(h)))
        else:
            self._geometry = geometry
            check(lib().rdst_mesh_create_geometry(C.byref(self._md.desc), C.byref(geometry.geom), C.byref(h)))
        self.h = h
        info = (C.c_int64 * 10)()
        check(lib().rdst_mesh_info(self.h, info))
        self.info = dict(zip(INFO_KEYS, [int(v) for v in info]))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.rdst_mesh_destroy(self.h)
            self.h = None

    def get_int(self, which):
        i = self.info
        sizes = {M_OWNER: i["n_internal"] + i["n_bfaces"], M_NEIGHBOUR: i["n_internal"], M_ROWPTR: i["n_cells"] + 1,
                 M_COLS: i["nnzb"], M_SLOT_UPPER: i["n_internal"], M_SLOT_LOWER: i["n_internal"],
                 M_SLOT_DIAG: i["n_cells"], M_FACE_COLOUR: i["n_internal"], M_ILU_LEVEL: i["n_cells"],
                 M_BFACE_PATCH: i["n_bfaces"], M_CELL_COLOUR: i["n_cells"], M_PARTITION: i["n_cells"], M_SA_PARTITION: i["n_cells"]}
        out = np.zeros(sizes[which], dtype=np.int32)
        check(lib().rdst_mesh_get_int(self.h, which, iptr(out)))
        return out

    def get_geom(self, which):
        i = self.info
        nf = i["n_internal"] + i["n_bfaces"]
        shapes = {G_SF: (nf, 3), G_MAGSF: (nf,), G_CF: (nf, 3), G_C: (i["n_cells"], 3), G_V: (i["n_cells"],),
                  G_WEIGHTS: (nf,), G_DELTACOEFFS: (nf,)}
        out = np.zeros(shapes[which])
        check(lib().rdst_mesh_get_geom(self.h, which, dptr(out)))
        return out

    def wall_distance(self):
        """Exact distance of every cell centre to the wall patches (device)."""
        init()
        out = np.zeros(self.info["n_cells"])
        check(lib().rdst_mesh_wall_distance(self.h, dptr(out)))
        return out


_FIELD_SHAPES = {1: 3, 5: 3, 8: 5, 9: 5, 13: 3, 15: 3, 17: 3, 20: 3, 24: 25, 25: 5, 39: 9}


class DeviceCase:
    """The device-resident Single::iterate state of one case (rdst_case_*)."""

    def __init__(self, spec, mesh: DeviceMesh | None = None):
        init()
        self.spec = spec
        self.mesh = mesh if mesh is not None else DeviceMesh(spec.mesh)
        h = C.c_void_p()
        check(lib().rdst_case_create(self.mesh.h, C.byref(spec.thermo), spec.bcs, C.byref(spec.settings),
                                     C.byref(h)))
        self.h = h
        i = self.mesh.info
        self.n, self.nI, self.nB, self.nnzb = i["n_cells"], i["n_internal"], i["n_bfaces"], i["nnzb"]
        self.set_state(spec.p0, spec.T0, spec.U0)
        if getattr(spec, "turbulence", None) is not None:
            self.set_turbulence(*spec.turbulence)

    def set_turbulence(self, turb, bcs, wall_dist, nu_tilda):
        """wall_dist None: the device computes the exact wall distance."""
        d = None if wall_dist is None else np.ascontiguousarray(wall_dist, dtype=np.float64)
        nt = np.ascontiguousarray(nu_tilda, dtype=np.float64)
        check(lib().rdst_case_set_turbulence(self.h, C.byref(turb), bcs, None if d is None else dptr(d), dptr(nt)))

    def turbulence_correct(self):
        perf = rdst_solver_perf()
        check(lib().rdst_turbulence_correct(self.h, C.byref(perf)))
        return perf

    # ---- discrete adjoint (first-order inviscid residual)
    def adjoint_linearise(self):
        check(lib().rdst_adjoint_linearise(self.h))

    def adjoint_apply(self, x, transpose=False):
        xa = np.ascontiguousarray(x, dtype=np.float64).reshape(self.n, 5)
        y = np.zeros((self.n, 5))
        check(lib().rdst_adjoint_apply(self.h, 1 if transpose else 0, dptr(xa), dptr(y)))
        return y

    def adjoint_residual(self):
        R = np.zeros((self.n, 5))
        check(lib().rdst_adjoint_residual(self.h, dptr(R)))
        return R

    def adjoint_objective_cells(self, patch, direction, q, p_ref):
        d = np.ascontiguousarray(direction, dtype=np.float64)
        o = np.zeros(self.n)
        check(lib().rdst_adjoint_objective_cells(self.h, patch, dptr(d), float(q), float(p_ref), dptr(o)))
        return o

    def adjoint_objective_gradient(self, patch, direction, q):
        d = np.ascontiguousarray(direction, dtype=np.float64)
        g = np.zeros((self.n, 5))
        check(lib().rdst_adjoint_objective_gradient(self.h, patch, dptr(d), float(q), dptr(g)))
        return g

    def adjoint_solve(self, patch, direction, q, lambda0=None, **settings):
        d = np.ascontiguousarray(direction, dtype=np.float64)
        lam = np.zeros((self.n, 5)) if lambda0 is None else np.ascontiguousarray(lambda0, dtype=np.float64).copy()
        s = make_adjoint_settings(**settings)
        perf = rdst_solver_perf()
        check(lib().rdst_adjoint_solve(self.h, patch, dptr(d), float(q), C.byref(s), dptr(lam), C.byref(perf)))
        return lam, perf

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.rdst_case_destroy(self.h)
            self.h = None

    def set_state(self, p, T, U):
        p = np.ascontiguousarray(p, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        U = np.ascontiguousarray(U, dtype=np.float64)
        check(lib().rdst_case_set_state(self.h, dptr(p), dptr(T), dptr(U)))

    def _size(self, field):
        if field in (12, 13, 14, 15):
            return self.nI + self.nB
        if 16 <= field <= 23 or field in (33, 34):
            return self.nB
        if field in (36, 37):
            return self.nI
        if field == 24:
            return self.nnzb
        return self.n

    def get(self, field):
        k = _FIELD_SHAPES.get(field, 1)
        out = np.zeros((self._size(field), k))
        check(lib().rdst_case_get(self.h, field, dptr(out)))
        return out if k > 1 else out[:, 0]

    def put(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        check(lib().rdst_case_put(self.h, field, dptr(a)))

    def ldu_solve(self, diag, upper, lower, b, x, mode=0):
        """rdst_ldu_solve on a caller-assembled LDU system: diag [n,5,5],
        upper/lower [nI,5,5] (polyMesh internal-face order), b [n,5], x [n,5]
        initial guess; mode 0 PLAIN (A x = b), 1 W (A x' = A x + b), 2 deltaW
        (A d = b, x += d).  Returns (x, perf)."""
        d = np.ascontiguousarray(diag, dtype=np.float64)
        u = np.ascontiguousarray(upper, dtype=np.float64)
        l_ = np.ascontiguousarray(lower, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        xx = np.ascontiguousarray(x, dtype=np.float64).copy()
        perf = rdst_solver_perf()
        check(lib().rdst_ldu_solve(self.h, dptr(d), dptr(u), dptr(l_), dptr(bb), dptr(xx), mode, C.byref(perf)))
        return xx, perf

    def flux_residual(self):
        check(lib().rdst_flux_residual(self.h))

    def jacobian_assemble(self):
        check(lib().rdst_jacobian_assemble(self.h))

    def block_solve(self):
        perf = rdst_solver_perf()
        check(lib().rdst_block_solve(self.h, C.byref(perf)))
        return perf

    def update_fields(self):
        check(lib().rdst_update_fields(self.h))

    def set_courant(self):
        r = C.c_double()
        check(lib().rdst_set_courant(self.h, C.byref(r)))
        return r.value

    def iterate(self, n_steps=1, with_perf=False):
        res = np.zeros(n_steps)
        perfs = (rdst_solver_perf * n_steps)()
        check(lib().rdst_iterate(self.h, n_steps, dptr(res), perfs))
        return (res, perfs) if with_perf else res

    def dual_step(self):
        """One physical time step of a Dual case (rdst_dual_step): (inner
        iterations, their mean initial residuals, |R|_2 of the closing
        setCourantDST)."""
        k = C.c_int32()
        inner = np.zeros(self.spec.settings.max_inner_iter + 1)
        r = C.c_double()
        check(lib().rdst_dual_step(self.h, C.byref(k), dptr(inner), C.byref(r)))
        return k.value, inner[:k.value].copy(), r.value

    def time_info(self):
        out = np.zeros(4)
        check(lib().rdst_time_info(self.h, dptr(out)))
        return {"time": out[0], "deltaT": out[1], "deltaT0": out[2], "courant": out[3]}

    def forces(self, patch, drag_dir, lift_dir, p_ref, rho_inf, mag_u_inf, a_ref=1.0):
        d = np.ascontiguousarray(drag_dir, dtype=np.float64)
        l_ = np.ascontiguousarray(lift_dir, dtype=np.float64)
        out = np.zeros(8)
        check(lib().rdst_forces(self.h, patch, dptr(d), dptr(l_), p_ref, rho_inf, mag_u_inf, a_ref, dptr(out)))
        return out

    def last_flux_ms(self):
        v = C.c_double()
        check(lib().rdst_last_flux_ms(self.h, C.byref(v)))
        return v.value

    def time_flux(self, n=20):
        v = C.c_double()
        check(lib().rdst_time_flux(self.h, n, C.byref(v)))
        return v.value

    def bcsr(self):
        """(row_ptr, cols, values [nnzb,5,5]) of the assembled FP64 matrix."""
        rp = self.mesh.get_int(M_ROWPTR)
        cols = self.mesh.get_int(M_COLS)
        vals = self.get(24).reshape(-1, 5, 5)
        return rp, cols, vals
