"""The reference's RAE2822 tutorial case from the committed copy of its
polyMesh, and foam-extend's stock SpalartAllmaras model (the RASModel of
every turbulent tutorial, RAE2822/constant/RASProperties:19), which is
SARC2's equation with the rotation-curvature multiplier fr1 = 1
(SARC2.C:121-140 returns 1 when RCMCorrection is off)."""
import os

import numpy as np
import pytest

from rhodst_amd import cases, meshgen

RAE = "/root/reference/tutorials/RAE2822"


def test_rae2822_fixture_counts():
    m = meshgen.polymesh_from_npz(cases.RAE2822_MESH)
    assert m.n_cells == 23552 and len(m.owner) == 94608 and len(m.neighbour) == 46704
    assert [p[0] for p in m.patches] == ["frontAndBackPlanes", "freestreamBC", "airfoil"]
    assert [p[3] for p in m.patches] == [47104, 496, 304]


@pytest.mark.skipif(not os.path.isdir(RAE), reason="reference tutorials not mounted")
def test_rae2822_fixture_is_the_tutorial_mesh():
    a = meshgen.polymesh_from_npz(cases.RAE2822_MESH)
    b = meshgen.read_polymesh(RAE)
    for k in ("points", "face_offsets", "face_points", "owner", "neighbour"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert a.patches == b.patches


def test_stock_spalart_allmaras_is_sarc2_without_rotation_correction(oracle):
    spec_sa = cases.rae2822_case()
    spec_rc = cases.rae2822_case(turbulence="SARC2")
    spec_rc.turbulence[0].rcm_correction = 0
    o1, o2 = oracle.OracleCase(spec_sa), oracle.OracleCase(spec_rc)
    r1, r2 = o1.iterate(3), o2.iterate(3)
    assert np.array_equal(r1, r2)
    for f in (9, 29, 30):
        assert np.array_equal(o1.get(f), o2.get(f)), f
    # and the rotation-curvature multiplier does change the answer
    o3 = oracle.OracleCase(cases.rae2822_case(turbulence="SARC2"))
    o3.iterate(3)
    assert not np.array_equal(o3.get(29), o1.get(29))


def test_rae2822_runs_in_the_oracle(oracle):
    oc = oracle.OracleCase(cases.rae2822_case())
    r = oc.iterate(8)
    assert np.all(np.isfinite(r)) and r[-1] < 0.2 * r[0]
