"""rdst_forces_noalloc (the force integral with its block partials in the
library's static device scratch, no device allocation) against rdst_forces on
the same case: identical kernels and reduction order, so the eight outputs
agree bit for bit; both agree with the oracle's force integral to 1e-9.  Two
cases in a row, the second with fewer wall faces, so the second call reuses
the scratch the first one filled."""
import ctypes as C

import numpy as np
import pytest

from rhodst_amd import DeviceCase, cases
from rhodst_amd._abi import dptr
from rhodst_amd._capi import check, lib

pytestmark = pytest.mark.gpu


def _both(dc, patch, drag, lift, p_ref, rho, u, a_ref):
    d = np.ascontiguousarray(drag, dtype=np.float64)
    l_ = np.ascontiguousarray(lift, dtype=np.float64)
    ref = np.zeros(8)
    out = np.zeros(8)
    check(lib().rdst_forces(dc.h, patch, dptr(d), dptr(l_), p_ref, rho, u, a_ref, dptr(ref)))
    check(lib().rdst_forces_noalloc(dc.h, patch, dptr(d), dptr(l_), p_ref, rho, u, a_ref, dptr(out)))
    return ref, out


def test_forces_noalloc_matches_forces(oracle):
    a = np.radians(2.3)
    gamma, R, T, p = 1.4, 287.0, 288.15, 101325.0
    U = 0.73 * np.sqrt(gamma * R * T)
    drag, lift = (np.cos(a), np.sin(a), 0.0), (-np.sin(a), np.cos(a), 0.0)
    # first call: viscous SA airfoil (pressure and viscous parts), 160 wall faces
    print("[no_alloc] airfoil SAneg, 3 steps", flush=True)
    big = cases.airfoil_case(turbulence="SAneg", mesh_kw=dict(n_surf=80, n_layers=16, first_cell=2e-5, far=8.0))
    dc = DeviceCase(big)
    oc = oracle.OracleCase(big)
    from rhodst_amd import _capi
    oc.set_sa_partition(dc.mesh.get_int(_capi.M_SA_PARTITION))
    dc.iterate(3)
    oc.iterate(3)
    wall = big.mesh.patch_index("airfoil")
    ref, out = _both(dc, wall, drag, lift, p, p / (R * T), U, 0.1)
    assert np.array_equal(ref, out), (ref, out)
    o = oc.forces(wall, drag, lift, p, p / (R * T), U, 0.1)
    assert np.abs(out[:2] - o[:2]).max() <= 1e-9 * np.abs(o[:2]).max(), (out, o)
    assert abs(out[5]) > 0.0  # the viscous part was integrated
    # second call, smaller input: the laminar no-slip box, 12 wall faces
    print("[no_alloc] viscous box 12x8, 2 steps", flush=True)
    small = cases.box_viscous_case(nx=12, ny=8)
    dc2 = DeviceCase(small)
    oc2 = oracle.OracleCase(small)
    dc2.iterate(2)
    oc2.iterate(2)
    wall2 = small.mesh.patch_index("bottom")
    ref2, out2 = _both(dc2, wall2, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 1.0e5, 1.16, 100.0, 1e-4)
    assert np.array_equal(ref2, out2), (ref2, out2)
    o2 = oc2.forces(wall2, (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 1.0e5, 1.16, 100.0, 1e-4)
    assert np.abs(out2[:2] - o2[:2]).max() <= 1e-9 * np.abs(o2[:2]).max(), (out2, o2)
    # and the first case again after the smaller one
    ref3, out3 = _both(dc, wall, drag, lift, p, p / (R * T), U, 0.1)
    assert np.array_equal(ref3, out3) and np.array_equal(out3, out)


def test_forces_noalloc_rejects_bad_patch():
    spec = cases.box_viscous_case(nx=6, ny=4)
    dc = DeviceCase(spec)
    out = np.zeros(8)
    d = np.array([1.0, 0.0, 0.0])
    rc = lib().rdst_forces_noalloc(dc.h, 99, dptr(d), dptr(d), 0.0, 1.0, 1.0, 1.0, dptr(out))
    assert rc == 1
    assert b"patch index out of range" in lib().rdst_last_error()
