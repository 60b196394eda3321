"""Quality of the swept-wing C-H mesh at the benchmark size (2 027 520 cells):
closed positive cells, every wall cell centre on the flow side of its wall
face, bounded non-orthogonality and neighbour volume ratio, and the wake's
first spacing growing behind the trailing edge."""
import numpy as np
import pytest

from rhodst_amd import meshgen


@pytest.fixture(scope="module")
def wing(oracle):
    m = meshgen.swept_wing()
    return m, oracle.geometry(m)


def test_swept_wing_cells_closed_and_positive(wing):
    m, g = wing
    assert m.n_cells == 352 * 96 * 60
    assert (g["V"] > 0).all()
    acc = np.zeros((m.n_cells, 3))
    nf = len(g["magSf"])
    np.add.at(acc, m.owner[:nf], g["Sf"])
    np.add.at(acc, m.neighbour, -g["Sf"][:m.n_internal])
    scale = np.zeros(m.n_cells)
    np.maximum.at(scale, m.owner[:nf], g["magSf"])
    assert (np.linalg.norm(acc, axis=1) <= 1e-9 * scale).all()


def test_swept_wing_wall_cells(wing):
    """Each wall face's cell centre lies on the flow side at about half the
    first-cell height (2e-6 of the normal line's length / far, 2e-6 to 2.6e-6
    here), so the wall deltaCoeffs stay positive."""
    m, g = wing
    _n, _t, start, size = m.patches[m.patch_index("wing")]
    f = np.arange(start, start + size)
    n = g["Sf"][f] / g["magSf"][f][:, None]
    dn = np.einsum("ij,ij->i", n, g["Cf"][f] - g["C"][m.owner[f]])
    assert dn.min() > 0.25e-6 and dn.max() < 1.35e-6
    assert 0.95e-6 < np.median(dn) < 1.3e-6


def test_swept_wing_nonorthogonality_and_volume_ratio(wing):
    m, g = wing
    nI = m.n_internal
    d = g["C"][m.neighbour] - g["C"][m.owner[:nI]]
    cosang = (d * g["Sf"][:nI]).sum(1) / (np.linalg.norm(d, axis=1) * g["magSf"][:nI])
    ang = np.degrees(np.arccos(np.clip(cosang, -1.0, 1.0)))
    # the worst faces are at the collapsed tip's leading edge; the swept
    # leading edge shears the spanwise faces of its boundary-layer cells
    assert ang.max() < 87.0
    assert np.percentile(ang, 99.9) < 66.0
    assert (ang > 70.0).sum() <= 400
    V = g["V"]
    vr = np.maximum(V[m.owner[:nI]] / V[m.neighbour], V[m.neighbour] / V[m.owner[:nI]])
    assert vr.max() < 30.0


def test_swept_wing_wake_spacing_grows(wing):
    """Wake cells next to the cut at the root: height ~ first_cell + 0.01 s,
    s the distance behind the local trailing edge (x_te = 1 + (tan 30 deg -
    0.44) z for the taper 0.56 planform); measured as the centre of the
    second row minus that of the first."""
    m, g = wing
    ni = 352
    i = np.arange(10, 36)               # lower wake, from the far field towards the TE
    C = g["C"]
    c0, c1 = C[i], C[i + ni]            # j = 0 and j = 1 cells at k = 0
    s = c0[:, 0] - (1.0 + (np.tan(np.deg2rad(30.0)) - 0.44) * c0[:, 2])
    h = np.abs(c1[:, 1] - c0[:, 1])
    f = 2e-6 + 0.01 * s
    assert (s > 0.005).all()
    assert np.all(h > 0.7 * f) and np.all(h < 1.6 * f)


def test_tile_order_is_a_renumbering(oracle):
    """order="tiles" renumbers the cells (4 x 4 x 4 tiles in Morton order)
    without changing the mesh: new cell n is old cell perm[n] (same centre and
    volume), the same numbers of cells and faces, and the two cells of nine
    faces in ten within 1000 of each other."""
    kw = dict(n_wake=6, n_surf=16, n_normal=16, n_span_wing=8, n_span_out=4, n_layers=8)
    a, b = meshgen.swept_wing(**kw), meshgen.swept_wing(order="tiles", **kw)
    ga, gb = oracle.geometry(a), oracle.geometry(b)
    assert a.n_cells == b.n_cells and a.n_internal == b.n_internal and len(a.owner) == len(b.owner)
    assert not np.array_equal(ga["C"], gb["C"])
    # the permutation as tile_order gives it
    ni, nj, nk = 44, 16, 12
    ii, jj, kk = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    ii, jj, kk = (x.transpose(2, 1, 0).ravel() for x in (ii, jj, kk))
    perm = meshgen.tile_order(ii, jj, kk)
    assert np.array_equal(np.sort(perm), np.arange(a.n_cells))
    # the same cells (geometry to round-off: the faces are summed in another order)
    assert np.allclose(gb["C"], ga["C"][perm], rtol=0, atol=1e-12)
    assert np.allclose(gb["V"], ga["V"][perm], rtol=1e-10, atol=0)
    d = np.abs(np.asarray(b.neighbour) - np.asarray(b.owner[:b.n_internal]))
    assert np.percentile(d, 90) < 1000
