"""The tetrahedron/prism box generator: cell counts, closed positive cells,
the box volume, and the conformity of Kuhn's subdivision with the prism
layer (every interior face shared by exactly two cells)."""
import numpy as np

from rhodst_amd import meshgen


def test_tet_prism_box(oracle):
    nx, ny, nz, pl = 6, 5, 4, 1
    m = meshgen.tet_prism_box(nx, ny, nz, prism_layers=pl, jitter=0.15)
    assert m.n_cells == 2 * nx * ny * pl + 6 * nx * ny * (nz - pl)
    g = oracle.geometry(m)
    assert (g["V"] > 0).all()
    assert abs(g["V"].sum() - 0.01 * 0.008 * 0.006) <= 1e-12 * 0.01 * 0.008 * 0.006
    acc = np.zeros((m.n_cells, 3))
    np.add.at(acc, m.owner, g["Sf"])
    np.add.at(acc, m.neighbour, -g["Sf"][:m.n_internal])
    assert np.abs(acc).max() <= 1e-12 * g["magSf"].max()
    # faces: triangles and quads only; boundary faces tile each side of the box
    sizes = np.diff(m.face_offsets)
    assert set(np.unique(sizes).tolist()) <= {3, 4}
    areas = {p[0]: g["magSf"][p[2]:p[2] + p[3]].sum() for p in m.patches}
    assert abs(areas["bottom"] - 0.01 * 0.008) < 1e-15 and abs(areas["left"] - 0.008 * 0.006) < 1e-15
    # every cell has 4 (tetrahedron) or 5 (prism) faces
    nf = np.bincount(m.owner, minlength=m.n_cells) + np.bincount(m.neighbour, minlength=m.n_cells)
    assert np.count_nonzero(nf == 5) == 2 * nx * ny * pl and np.count_nonzero(nf == 4) == 6 * nx * ny * (nz - pl)
