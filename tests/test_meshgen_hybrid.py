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


def test_airfoil_wall_displacement(oracle):
    """displace_airfoil_wall: each wall point moves by its displacement along
    its wall-layer line, the line's points by the same times a weight falling
    to 0 at ring n_layers; linear in the displacement; seeded bumps of 1e-4
    chord on the full first-cell height keep every cell positive."""
    kw = dict(n_surf=80, n_layers=16, first_cell=2e-6, far=8.0)
    m0 = meshgen.hybrid_airfoil(**kw)
    loop, kind = meshgen.airfoil_loop(80)
    ns = len(loop)
    rng = np.random.default_rng(1)
    d1, d2 = 1e-4 * rng.uniform(-1, 1, ns), 1e-4 * rng.uniform(-1, 1, ns)
    m1 = meshgen.displace_airfoil_wall(m0, d1, 16)
    m2 = meshgen.displace_airfoil_wall(m0, d2, 16)
    m12 = meshgen.displace_airfoil_wall(m0, d1 + d2, 16)
    assert np.abs((m12.points - m0.points) - (m1.points - m0.points) - (m2.points - m0.points)).max() < 1e-15
    nP = len(m0.points) // 2
    P0 = m0.points[:nP, :2].reshape(-1, ns, 2)
    P1 = m1.points[:nP, :2].reshape(-1, ns, 2)
    line = (P0[1] - P0[0]) / np.linalg.norm(P0[1] - P0[0], axis=1)[:, None]
    assert np.abs((P1[0] - P0[0]) - d1[:, None] * line).max() < 1e-15
    assert np.array_equal(P1[16:], P0[16:])
    assert np.array_equal(m1.points[:, 2], m0.points[:, 2])
    assert np.array_equal(m1.face_points, m0.face_points) and np.array_equal(m1.owner, m0.owner)
    # the same as generating with wall_disp
    assert np.array_equal(meshgen.hybrid_airfoil(**kw, wall_disp=d1).points, m1.points)
    g = oracle.geometry(m1)
    assert (g["V"] > 0).all()


def test_airfoil_problem_design_nodes():
    """The wall-node design set of the adjoint's airfoil problem: the loop's
    surface points without the lower trailing-edge corner (n_surf of them),
    and seeded Hicks-Henne bumps of at most ~1e-4 chord that vanish at the
    leading and trailing edges."""
    from rhodst_amd.adjoint import AirfoilProblem
    p = AirfoilProblem(mesh_kw=dict(n_surf=80))
    loop, kind = p.loop, p.kind
    assert len(p.design) == 80
    assert np.all(kind[p.design] < 2)
    lower_te = np.nonzero(kind == 0)[0][np.argmax(loop[kind == 0, 0])]
    assert lower_te not in set(p.design.tolist())
    b = p.base_disp
    assert np.all(b[kind == 2] == 0.0)
    assert 0.2e-4 < np.abs(b).max() < 2e-4
    assert abs(b[np.argmin(loop[:, 0])]) < 1e-12   # the leading edge
