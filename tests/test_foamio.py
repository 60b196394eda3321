"""foam time-directory writer and force log: a state written as ASCII
volFields reads back exactly (internal and per-patch values, empty patches
marked empty), next to the polyMesh written by meshgen.write_polymesh."""
import os

import numpy as np

from rhodst_amd import cases, foamio, meshgen


def test_time_directory_round_trip(oracle, tmp_path):
    spec = cases.box_sa_case("SAneg")
    oc = oracle.OracleCase(spec)
    oc.iterate(3)
    case_dir = str(tmp_path)
    meshgen.write_polymesh(spec.mesh, os.path.join(case_dir, "constant", "polyMesh"))
    paths = foamio.write_time(case_dir, "3", oc)
    assert sorted(os.path.basename(p) for p in paths) == ["T", "U", "mut", "nuTilda", "p", "rho"]
    back = meshgen.read_polymesh(case_dir)
    assert back.n_cells == spec.mesh.n_cells
    nonempty = [(n, s) for n, t, _st, s in spec.mesh.patches if t != meshgen.PATCH_EMPTY]
    for name, cid, bid in (("p", 3, 18), ("T", 4, 19), ("U", 5, 17), ("rho", 0, 16), ("nuTilda", 29, 33)):
        cells, patches = foamio.read_field(os.path.join(case_dir, "3", name))
        ref = oc.get(cid)
        assert np.array_equal(cells, ref.reshape(cells.shape)), name
        bref = oc.get(bid)
        off = 0
        for pname, size in nonempty:
            assert np.array_equal(patches[pname], bref[off:off + size].reshape(patches[pname].shape)), (name, pname)
            off += size
        for pname, t, _s, _n in spec.mesh.patches:
            if t == meshgen.PATCH_EMPTY:
                assert patches[pname] is None
    with open(os.path.join(case_dir, "3", "U")) as fh:
        head = fh.read(400)
    assert "class       volVectorField;" in head and "object      U;" in head
    assert "dimensions      [0 1 -1 0 0 0 0];" in open(os.path.join(case_dir, "3", "U")).read()


def test_force_log(tmp_path):
    path = os.path.join(str(tmp_path), "postProcessing", "forces", "0", "forceCoeffs.dat")
    log = foamio.ForceLog(path)
    log.append(1, 0.0123, 0.45)
    log.append(2, 0.0121, 0.46)
    data = foamio.ForceLog.read(path)
    assert data.shape == (2, 3)
    assert np.allclose(data[:, 1], [0.0123, 0.0121], rtol=1e-12) and np.allclose(data[:, 2], [0.45, 0.46])
