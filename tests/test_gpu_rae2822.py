"""The reference's RAE2822 tutorial through the device path: its polyMesh,
boundary conditions, thermo and solver settings (Kurganov, OSPRE/OSPREV
reconstruction, BiCGStab + ILUC0, foam-extend's stock SpalartAllmaras),
against the oracle over ten Single::iterate steps; the device state is then
written as a foam time directory with a forceCoeffs-style log."""
import os

import numpy as np
import pytest

from rhodst_amd import DeviceCase, _capi, cases, foamio, meshgen

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def coefficients(case, spec):
    # RAE2822/system/forces:19-38
    return case.forces(spec.mesh.patch_index("airfoil"), (0.9991873755767, 0.0403061842411, 0.0),
                       (-0.0403061842411, 0.9991873755767, 0.0), 108987.773, 1.48571291822, 233.6216, 1.0)


def test_rae2822_tutorial_parity(oracle, tmp_path):
    print("[RAE2822] SpalartAllmaras, Kurganov, OSPRE, BiCGStab: 10 steps", flush=True)
    spec = cases.rae2822_case()
    dm = _capi.DeviceMesh(spec.mesh)
    spec = cases.rae2822_case(mesh=spec.mesh, wall_dist=dm.wall_distance())
    dc = DeviceCase(spec, dm)
    oc = oracle.OracleCase(spec)
    oc.set_sa_partition(dm.get_int(_capi.M_SA_PARTITION))
    r1 = dc.iterate(10)
    r0 = oc.iterate(10)
    print(f"[RAE2822] |R|_2 {r1[0]:.6e} -> {r1[-1]:.6e}", flush=True)
    assert np.all(np.isfinite(r1)) and r1[-1] < 0.2 * r1[0]
    assert rel(r1, r0) <= 1e-9, (r1, r0)
    assert rel(dc.get(9), oc.get(9)) <= 1e-9
    assert rel(dc.get(29), oc.get(29)) <= 1e-7
    cd, co = coefficients(dc, spec), coefficients(oc, spec)
    assert rel(cd[:2], co[:2]) <= 1e-8, (cd, co)
    # foam time directory and force log of the device state
    case_dir = str(tmp_path)
    meshgen.write_polymesh(spec.mesh, os.path.join(case_dir, "constant", "polyMesh"))
    foamio.write_time(case_dir, "10", dc)
    log = foamio.ForceLog(os.path.join(case_dir, "postProcessing", "forces", "0", "forceCoeffs.dat"))
    log.append(10, cd[0], cd[1])
    back = foamio.ForceLog.read(log.path)
    assert back.shape == (1, 3) and np.isclose(back[0, 1], cd[0], rtol=1e-10)
    cells, patches = foamio.read_field(os.path.join(case_dir, "10", "p"))
    assert np.array_equal(cells, dc.get(3))
    assert patches["frontAndBackPlanes"] is None
