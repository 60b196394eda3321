"""The ctypes descriptors in rhodst_amd/_abi.py against include/rhodst.h: a
C program generated from the ctypes field lists is compiled against the
header (a field the header does not have fails to compile) and prints
sizeof and offsetof of every field, which must equal ctypes' own.  Also the
caller-geometry mesh entry point (host only, no GPU)."""
import ctypes as C
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from rhodst_amd import _abi, _capi, meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = ["rdst_patch", "rdst_mesh_desc", "rdst_mesh_geometry", "rdst_bc", "rdst_thermo", "rdst_settings",
           "rdst_turbulence", "rdst_scalar_bc", "rdst_solver_perf", "rdst_adjoint_settings"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_ctypes_layouts_match_header(tmp_path):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "rhodst.h"', "int main(void) {",
             '    printf("{");']
    first = True
    for s in STRUCTS:
        cls = getattr(_abi, s)
        sep = "" if first else ","
        first = False
        lines.append(f'    printf("{sep}\\"{s}\\": {{\\"__sizeof\\": %zu", sizeof({s}));')
        for name, _t in cls._fields_:
            lines.append(f'    printf(", \\"{name}\\": %zu", offsetof({s}, {name}));')
        lines.append('    printf("}");')
    lines += ['    printf("}\\n");', "    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    for s in STRUCTS:
        cls = getattr(_abi, s)
        assert got[s]["__sizeof"] == C.sizeof(cls), s
        for name, _t in cls._fields_:
            assert got[s][name] == getattr(cls, name).offset, (s, name)


def _poly_order(mesh, arr_internal_bnd, dm):
    """Spread a surface field [nI + nB(non-empty)] back to polyMesh face order
    (faces of empty patches get garbage values: the entry point ignores them)."""
    nf = len(mesh.owner)
    k = arr_internal_bnd.shape[1:] if arr_internal_bnd.ndim > 1 else ()
    out = np.full((nf,) + k, 1.0e300)
    nI = mesh.n_internal
    out[:nI] = arr_internal_bnd[:nI]
    b = nI
    for name, t, start, size in mesh.patches:
        if t == meshgen.PATCH_EMPTY:
            continue
        out[start:start + size] = arr_internal_bnd[b:b + size]
        b += size
    return out


def _geometry_of(dm, mesh, scale_V=1.0):
    g = {w: dm.get_geom(w) for w in (_capi.G_SF, _capi.G_MAGSF, _capi.G_CF, _capi.G_C, _capi.G_V, _capi.G_WEIGHTS,
                                      _capi.G_DELTACOEFFS)}
    return _abi.MeshGeometry(_poly_order(mesh, g[_capi.G_SF], dm), _poly_order(mesh, g[_capi.G_CF], dm),
                             g[_capi.G_C], g[_capi.G_V] * scale_V, _poly_order(mesh, g[_capi.G_WEIGHTS], dm),
                             _poly_order(mesh, g[_capi.G_DELTACOEFFS], dm),
                             mag_sf=_poly_order(mesh, g[_capi.G_MAGSF], dm))


def test_mesh_create_geometry_keeps_caller_geometry():
    mesh = meshgen.box_2d(9, 7, jitter=0.2, seed=3)
    dm = _capi.DeviceMesh(mesh)
    geo = _geometry_of(dm, mesh)
    dm2 = _capi.DeviceMesh(mesh, geometry=geo)
    for w in (_capi.G_SF, _capi.G_MAGSF, _capi.G_CF, _capi.G_C, _capi.G_V, _capi.G_WEIGHTS, _capi.G_DELTACOEFFS):
        assert np.array_equal(dm2.get_geom(w), dm.get_geom(w)), w
    for w in (_capi.M_ROWPTR, _capi.M_COLS, _capi.M_SLOT_UPPER, _capi.M_PARTITION, _capi.M_SA_PARTITION):
        assert np.array_equal(dm2.get_int(w), dm.get_int(w)), w
    # a different geometry is kept as given, not recomputed
    dm3 = _capi.DeviceMesh(mesh, geometry=_geometry_of(dm, mesh, scale_V=2.0))
    assert np.array_equal(dm3.get_geom(_capi.G_V), 2.0 * dm.get_geom(_capi.G_V))
    assert np.array_equal(dm3.get_geom(_capi.G_SF), dm.get_geom(_capi.G_SF))


def test_mesh_create_geometry_rejects_bad_input():
    mesh = meshgen.box_2d(3, 2)
    dm = _capi.DeviceMesh(mesh)
    geo = _geometry_of(dm, mesh, scale_V=-1.0)
    with pytest.raises(_capi.RdstError, match="non-positive cell volume"):
        _capi.DeviceMesh(mesh, geometry=geo)
    md = _abi.MeshDesc(mesh)
    h = C.c_void_p()
    rc = _capi.lib().rdst_mesh_create_geometry(C.byref(md.desc), None, C.byref(h))
    assert rc == _capi.RDST_ERR_ARG
