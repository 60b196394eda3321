"""The device wall distance (rdst_mesh_wall_distance) against brute force on
the host: the exact point-triangle distance (meshgen._point_triangle_distance)
to every wall triangle, minimum over all of them."""
import numpy as np
import pytest

from rhodst_amd import cases, meshgen
from rhodst_amd._capi import DeviceMesh, G_C

pytestmark = pytest.mark.gpu


def brute_force(mesh, patches, C):
    faces = []
    for name in patches:
        _n, _t, start, size = mesh.patches[mesh.patch_index(name)]
        faces.extend(range(start, start + size))
    A, B, Cc, _ = meshgen._face_triangles(mesh, np.array(faces))
    d = np.full(len(C), np.inf)
    for t in range(len(A)):
        n = len(C)
        d = np.minimum(d, meshgen._point_triangle_distance(C, np.broadcast_to(A[t], (n, 3)),
                                                           np.broadcast_to(B[t], (n, 3)),
                                                           np.broadcast_to(Cc[t], (n, 3))))
    return d


@pytest.mark.parametrize("kind", ["box", "airfoil", "wing"])
def test_wall_distance_brute_force(kind):
    print(f"[wall distance] {kind}", flush=True)
    if kind == "box":
        mesh, walls = cases.box_sa_case().mesh, ["bottom"]
        # the box's bottom is a generic patch carrying wall BCs; type it as a wall
        i = mesh.patch_index("bottom")
        name, _t, start, size = mesh.patches[i]
        mesh.patches[i] = (name, meshgen.PATCH_WALL, start, size)
    elif kind == "airfoil":
        mesh, walls = meshgen.hybrid_airfoil(n_surf=80, n_layers=16, first_cell=2e-5, far=8.0), ["airfoil"]
    else:
        mesh, walls = meshgen.swept_wing(n_wake=8, n_surf=16, n_normal=16, n_span_wing=6, n_span_out=2, n_layers=8), ["wing"]
    dm = DeviceMesh(mesh)
    d = dm.wall_distance()
    ref = brute_force(mesh, walls, dm.get_geom(G_C))
    assert np.all(d > 0)
    assert np.abs(d - ref).max() <= 1e-12 * ref.max(), np.abs(d - ref).max()
    # the k-nearest host method used by the cases agrees as well
    assert np.abs(meshgen.wall_distance(mesh, walls, centres=dm.get_geom(G_C)) - ref).max() <= 1e-12 * ref.max()
