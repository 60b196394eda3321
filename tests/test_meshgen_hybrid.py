"""The hybrid airfoil mesh generator: closed, positive cells, hexes in the
wall layers and prisms outside, bounded non-orthogonality, wall distance."""
import numpy as np

from rhodst_amd import meshgen


def test_hybrid_airfoil_quality(oracle):
    m = meshgen.hybrid_airfoil(n_surf=80, n_layers=16, first_cell=2e-5, far=8.0)
    g = oracle.geometry(m)
    assert (g["V"] > 0).all()
    # closed cells: the in-plane parts of the outward face-area vectors of
    # every cell sum to zero (the empty front/back faces, left out of the
    # geometry, are normal to z)
    nf_all = len(g["magSf"])
    acc = np.zeros((m.n_cells, 3))
    np.add.at(acc, m.owner[:nf_all], g["Sf"])
    np.add.at(acc, m.neighbour, -g["Sf"][:m.n_internal])
    assert np.abs(acc[:, :2]).max() <= 1e-12 * g["magSf"].max()
    nf = np.bincount(m.owner, minlength=m.n_cells) + np.bincount(m.neighbour, minlength=m.n_cells)
    counts = np.bincount(nf, minlength=7)
    assert counts[5] > 0 and counts[6] > 0 and counts[5] + counts[6] == m.n_cells
    nI = m.n_internal
    d = g["C"][m.neighbour] - g["C"][m.owner[:nI]]
    cosang = (d * g["Sf"][:nI]).sum(1) / (np.linalg.norm(d, axis=1) * g["magSf"][:nI])
    assert np.degrees(np.arccos(cosang.min())) < 65.0
    names = [p[0] for p in m.patches]
    assert names == ["airfoil", "farfield", "frontAndBack"]
    # far-field faces lie on the 8-chord circle about (0.5, 0)
    _n, _t, start, size = m.patches[1]
    r = np.linalg.norm(g["Cf"][start:start + size, :2] - np.array([0.5, 0.0]), axis=1)
    assert np.abs(r - 8.0).max() < 0.05
    # the first wall-layer cells sit half the first height off the wall
    # (closer at the corners of the blunt trailing edge)
    wd = meshgen.wall_distance(m, ["airfoil"])
    _n, _t, start, size = m.patches[0]
    wall_d = wd[m.owner[start:start + size]]
    assert abs(np.median(wall_d) - 1e-5) < 1e-7
    assert wall_d.min() > 0.75e-5 and wall_d.max() < 1.01e-5
