"""Mesh generation and polyMesh I/O for the block-coupled implicit step.

The reference reads its mesh through foam-extend's fvMesh (blockMesh for the
forward step, tutorials/forwardStep/constant/polyMesh/blockMeshDict:18-98).
foam-extend is not available here, so this module builds the same polyMesh
primitives (points, faces, owner, neighbour, boundary) from hexahedral cell
shapes the way polyMesh's cell-shape constructor does:

* cells keep the order they are given in (block by block, i fastest, then j, k);
* internal faces are in upper-triangular order: sorted by owner, then by
  neighbour, owner < neighbour, with the normal pointing owner -> neighbour;
* boundary faces follow, grouped by patch in patch order, each patch sorted by
  (face cell, local hex face), with outward normals.

Face point lists use the hex cell model of OpenFOAM (faces (0 4 7 3),
(1 2 6 5), (0 1 5 4), (3 7 6 2), (0 3 2 1), (4 5 6 7), all outward).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field, replace

import numpy as np

PATCH_GENERIC, PATCH_WALL, PATCH_SYMMETRY, PATCH_EMPTY = 0, 1, 2, 3
_PATCH_TYPE_NAMES = {"patch": PATCH_GENERIC, "wall": PATCH_WALL,
                     "symmetryPlane": PATCH_SYMMETRY, "symmetry": PATCH_SYMMETRY,
                     "empty": PATCH_EMPTY}

# OpenFOAM hex cell model faces (outward)
HEX_FACES = np.array([[0, 4, 7, 3], [1, 2, 6, 5], [0, 1, 5, 4],
                      [3, 7, 6, 2], [0, 3, 2, 1], [4, 5, 6, 7]], dtype=np.int64)


@dataclass
class PolyMesh:
    points: np.ndarray          # [nP,3] float64
    face_offsets: np.ndarray    # [nF+1] int32
    face_points: np.ndarray     # [..] int32
    owner: np.ndarray           # [nF] int32
    neighbour: np.ndarray       # [nI] int32
    n_cells: int
    patches: list = field(default_factory=list)  # (name, type, start, size)

    @property
    def n_faces(self) -> int:
        return len(self.owner)

    @property
    def n_internal(self) -> int:
        return len(self.neighbour)

    def patch_index(self, name: str) -> int:
        for i, p in enumerate(self.patches):
            if p[0] == name:
                return i
        raise KeyError(name)

    def n_boundary_nonempty(self) -> int:
        return sum(p[3] for p in self.patches if p[1] != PATCH_EMPTY)


def hex_polymesh(points: np.ndarray, hexes: np.ndarray, classify, patch_defs) -> PolyMesh:
    """Build a polyMesh from hex cells.

    points: [nP,3]; hexes: [nC,8] point labels in hex-model order.
    classify(cell_ids, local_face_ids, face_centres) -> patch index per
    boundary face.  patch_defs: list of (name, type) in patch order.
    """
    nC = hexes.shape[0]
    hexes = np.asarray(hexes, dtype=np.int64)
    faces = hexes[:, HEX_FACES]                      # [nC,6,4]
    flat = faces.reshape(-1, 4)
    key = np.sort(flat, axis=1)
    # lexicographic unique over the sorted point quadruples
    order = np.lexsort((key[:, 3], key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    same_next = np.all(ks[1:] == ks[:-1], axis=1)
    cell_of = np.repeat(np.arange(nC, dtype=np.int64), 6)
    lf_of = np.tile(np.arange(6, dtype=np.int64), nC)
    # pairs (order[i], order[i+1]) with equal keys are internal faces
    pair_idx = np.nonzero(same_next)[0]
    a = order[pair_idx]
    b = order[pair_idx + 1]
    ca, cb = cell_of[a], cell_of[b]
    own_slot = np.where(ca < cb, a, b)
    nei_cell = np.where(ca < cb, cb, ca)
    own_cell = np.minimum(ca, cb)
    io = np.lexsort((nei_cell, own_cell))
    own_slot, own_cell, nei_cell = own_slot[io], own_cell[io], nei_cell[io]
    internal_faces = flat[own_slot]
    # two cells sharing more than one face (a folded C-mesh row): merge the
    # faces of the pair into one polygon (they share an edge, both are
    # oriented outward from the owner)
    dup = np.nonzero((own_cell[1:] == own_cell[:-1]) & (nei_cell[1:] == nei_cell[:-1]))[0]
    merged_faces = {}
    if len(dup):
        for d0 in dup:
            merged_faces[int(d0)] = _merge_polygons(list(map(int, internal_faces[d0])),
                                                    list(map(int, internal_faces[d0 + 1])))
        keep = np.ones(len(own_cell), dtype=bool)
        keep[dup + 1] = False
        new_index = np.cumsum(keep) - 1
        merged_faces = {int(new_index[k]): v for k, v in merged_faces.items()}
        internal_faces = internal_faces[keep]
        own_cell, nei_cell = own_cell[keep], nei_cell[keep]
    # boundary faces: slots that appear once
    is_paired = np.zeros(flat.shape[0], dtype=bool)
    is_paired[a] = True
    is_paired[b] = True
    bslots = np.nonzero(~is_paired)[0]
    bcell = cell_of[bslots]
    blf = lf_of[bslots]
    bfaces = flat[bslots]
    bcentres = points[bfaces].mean(axis=1)
    bpatch = np.asarray(classify(bcell, blf, bcentres), dtype=np.int64)
    if np.any(bpatch < 0):
        raise ValueError("unclassified boundary faces")
    ob = np.lexsort((blf, bcell, bpatch))
    bfaces, bcell, bpatch = bfaces[ob], bcell[ob], bpatch[ob]
    owner = np.concatenate([own_cell, bcell]).astype(np.int32)
    patches = []
    n_int = len(own_cell)
    start = n_int
    for pi, (name, ptype) in enumerate(patch_defs):
        size = int(np.count_nonzero(bpatch == pi))
        patches.append((name, ptype, start, size))
        start += size
    all_faces = np.concatenate([internal_faces, bfaces], axis=0)
    nF = all_faces.shape[0]
    sizes = np.full(nF, 4, dtype=np.int64)
    for k, v in merged_faces.items():
        sizes[k] = len(v)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    fpts = np.empty(offsets[-1], dtype=np.int64)
    regular = np.ones(nF, dtype=bool)
    regular[list(merged_faces.keys())] = False
    ridx = np.nonzero(regular)[0]
    fpts[offsets[ridx][:, None] + np.arange(4)] = all_faces[ridx]
    for k, v in merged_faces.items():
        fpts[offsets[k]:offsets[k + 1]] = v
    offsets = offsets.astype(np.int32)
    fpts = fpts.astype(np.int32)
    return PolyMesh(points=np.ascontiguousarray(points, dtype=np.float64),
                    face_offsets=offsets, face_points=fpts,
                    owner=owner, neighbour=nei_cell.astype(np.int32),
                    n_cells=nC, patches=patches)


def _merge_polygons(a, b):
    """Union of two consistently oriented polygons sharing one edge
    (traversed u->v in a and v->u in b): the boundary of a from v back to u,
    then b from u back to v."""
    na, nb = len(a), len(b)
    for i in range(na):
        u, v = a[i], a[(i + 1) % na]
        for j in range(nb):
            if b[j] == v and b[(j + 1) % nb] == u:
                pa = [a[(i + 1 + k) % na] for k in range(na)]      # v ... u
                pb = [b[(j + 2 + k) % nb] for k in range(nb - 2)]  # after u, before v
                return pa + pb
    raise ValueError("faces of a duplicated cell pair do not share an edge")


def _block_cells(nx, ny, x0, x1, y0, y1, zlo, zhi, point_id):
    """Hexes of one 2-D block (one cell thick), i fastest then j."""
    xs = np.linspace(x0, x1, nx + 1)
    ys = np.linspace(y0, y1, ny + 1)
    hexes = []
    for j in range(ny):
        for i in range(nx):
            c = [(xs[i], ys[j]), (xs[i + 1], ys[j]), (xs[i + 1], ys[j + 1]), (xs[i], ys[j + 1])]
            lo = [point_id(x, y, zlo) for x, y in c]
            hi = [point_id(x, y, zhi) for x, y in c]
            hexes.append(lo + hi)
    return hexes


def forward_step(nx: int = 120, ny: int = 40) -> PolyMesh:
    """Mach-3 forward-facing step in a 3 x 1 channel, step 0.2 high at x = 0.6.

    Geometry and 3-block topology of tutorials/forwardStep/constant/polyMesh/
    blockMeshDict:18-43 at uniform spacing 3/nx = 1/ny (nx = 3 ny), one cell
    thick in z in [-0.05, 0.05].  Default 120 x 40 -> 4032 cells.
    Patches: inlet, outlet, bottom (symmetryPlane), top (symmetryPlane),
    obstacle, defaultFaces (empty).
    """
    if nx != 3 * ny or ny % 5 != 0:
        raise ValueError("need nx = 3*ny and ny divisible by 5")
    d = 1.0 / ny
    n06 = int(round(0.6 / d))
    n02 = int(round(0.2 / d))
    pts = {}
    plist = []

    def pid(x, y, z):
        k = (int(round(x / d * 2)), int(round(y / d * 2)), 0 if z < 0 else 1)
        if k not in pts:
            pts[k] = len(plist)
            plist.append((x, y, z))
        return pts[k]

    zlo, zhi = -0.05, 0.05
    hexes = []
    hexes += _block_cells(n06, n02, 0.0, 0.6, 0.0, 0.2, zlo, zhi, pid)
    hexes += _block_cells(n06, ny - n02, 0.0, 0.6, 0.2, 1.0, zlo, zhi, pid)
    hexes += _block_cells(nx - n06, ny - n02, 0.6, 3.0, 0.2, 1.0, zlo, zhi, pid)
    points = np.array(plist, dtype=np.float64)
    hexes = np.array(hexes, dtype=np.int64)
    eps = 1e-9

    def classify(cells, lf, cen):
        x, y, z = cen[:, 0], cen[:, 1], cen[:, 2]
        out = np.full(len(cells), -1)
        out[(lf == 4) | (lf == 5)] = 5
        side = out < 0
        out[side & (np.abs(x) < eps)] = 0                      # inlet
        out[side & (np.abs(x - 3.0) < eps)] = 1                # outlet
        out[side & (np.abs(y) < eps)] = 2                      # bottom
        out[side & (np.abs(y - 1.0) < eps)] = 3                # top
        obst = side & (((np.abs(x - 0.6) < eps) & (y < 0.2 + eps)) | ((np.abs(y - 0.2) < eps) & (x > 0.6 - eps)))
        out[obst] = 4
        return out

    patch_defs = [("inlet", PATCH_GENERIC), ("outlet", PATCH_GENERIC), ("bottom", PATCH_SYMMETRY),
                  ("top", PATCH_SYMMETRY), ("obstacle", PATCH_GENERIC), ("defaultFaces", PATCH_EMPTY)]
    return hex_polymesh(points, hexes, classify, patch_defs)


def channel_1d(n: int, length: float = 1.0) -> PolyMesh:
    """A 1-D row of n hex cells along x (block-tridiagonal LDU), used by the
    exact-LU tests.  Patches: left, right, sides (empty)."""
    d = length / n
    xs = np.arange(n + 1) * d
    pts = []
    for z in (-0.5 * d, 0.5 * d):
        for y in (0.0, d):
            for x in xs:
                pts.append((x, y, z))
    points = np.array(pts)
    nx1 = n + 1

    def p(i, j, k):
        return i + nx1 * (j + 2 * k)

    hexes = np.array([[p(i, 0, 0), p(i + 1, 0, 0), p(i + 1, 1, 0), p(i, 1, 0),
                       p(i, 0, 1), p(i + 1, 0, 1), p(i + 1, 1, 1), p(i, 1, 1)] for i in range(n)])

    def classify(cells, lf, cen):
        out = np.full(len(cells), 2)
        out[np.abs(cen[:, 0]) < 1e-12] = 0
        out[np.abs(cen[:, 0] - length) < 1e-12] = 1
        return out

    return hex_polymesh(points, hexes, classify,
                        [("left", PATCH_GENERIC), ("right", PATCH_GENERIC), ("sides", PATCH_EMPTY)])


def box_2d(nx: int, ny: int, lx: float = 1.0, ly: float = 1.0, jitter: float = 0.0, seed: int = 1,
           periodic_like: bool = False) -> PolyMesh:
    """A 2-D box of nx x ny cells, one cell thick, optional seeded interior
    node jitter (fraction of the spacing).  Patches: left, right, bottom,
    top, frontAndBack (empty)."""
    dx, dy = lx / nx, ly / ny
    rng = np.random.default_rng(seed)
    X, Y = np.meshgrid(np.arange(nx + 1) * dx, np.arange(ny + 1) * dy, indexing="ij")
    if jitter > 0:
        J = rng.uniform(-jitter, jitter, size=(2, nx + 1, ny + 1))
        inner = np.zeros_like(X, dtype=bool)
        inner[1:-1, 1:-1] = True
        X = X + np.where(inner, J[0] * dx, 0.0)
        Y = Y + np.where(inner, J[1] * dy, 0.0)
    pts = np.zeros((2, ny + 1, nx + 1, 3))
    for k, z in enumerate((-0.05, 0.05)):
        pts[k, :, :, 0] = X.T
        pts[k, :, :, 1] = Y.T
        pts[k, :, :, 2] = z
    points = pts.reshape(-1, 3)

    def p(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    hexes = np.stack([p(ii, jj, 0), p(ii + 1, jj, 0), p(ii + 1, jj + 1, 0), p(ii, jj + 1, 0),
                      p(ii, jj, 1), p(ii + 1, jj, 1), p(ii + 1, jj + 1, 1), p(ii, jj + 1, 1)], axis=1)

    def classify(cells, lf, cen):
        out = np.full(len(cells), -1)
        out[(lf == 4) | (lf == 5)] = 4
        out[lf == 0] = 0
        out[lf == 1] = 1
        out[lf == 2] = 2
        out[lf == 3] = 3
        return out

    return hex_polymesh(points, hexes, classify,
                        [("left", PATCH_GENERIC), ("right", PATCH_GENERIC), ("bottom", PATCH_GENERIC),
                         ("top", PATCH_GENERIC), ("frontAndBack", PATCH_EMPTY)])


# --------------------------------------------------------------------------- polyMesh ASCII I/O
_HEADER_RE = re.compile(r"FoamFile\s*\{.*?\}", re.S)


def _strip(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return _HEADER_RE.sub("", text, count=1)


def _read_list_body(path: str) -> str:
    with open(path, "r") as fh:
        text = _strip(fh.read())
    m = re.search(r"(\d+)\s*\(", text)
    if not m:
        raise ValueError(f"no list in {path}")
    return text[m.end():text.rfind(")")], int(m.group(1))


def read_polymesh(case_dir: str) -> PolyMesh:
    """Read an ASCII foam polyMesh (constant/polyMesh/{points,faces,owner,
    neighbour,boundary}), the format of tutorials/RAE2822/constant/polyMesh."""
    pm = os.path.join(case_dir, "constant", "polyMesh") if os.path.isdir(os.path.join(case_dir, "constant")) else case_dir
    body, n = _read_list_body(os.path.join(pm, "points"))
    pts = np.array(re.findall(r"\(([^()]*)\)", body)[:n], dtype=object)
    points = np.array([list(map(float, s.split())) for s in pts], dtype=np.float64)
    body, nf = _read_list_body(os.path.join(pm, "faces"))
    entries = re.findall(r"(\d+)\s*\(([^()]*)\)", body)
    offs = [0]
    fpts = []
    for cnt, lst in entries:
        v = list(map(int, lst.split()))
        assert len(v) == int(cnt)
        fpts.extend(v)
        offs.append(len(fpts))
    body, no = _read_list_body(os.path.join(pm, "owner"))
    owner = np.array(body.split(), dtype=np.int32)
    body, nn = _read_list_body(os.path.join(pm, "neighbour"))
    neighbour = np.array(body.split(), dtype=np.int32)
    with open(os.path.join(pm, "boundary")) as fh:
        btxt = _strip(fh.read())
    patches = []
    for name, blk in re.findall(r"(\w+)\s*\{([^{}]*)\}", btxt):
        t = re.search(r"\btype\s+(\w+)", blk).group(1)
        nfaces = int(re.search(r"nFaces\s+(\d+)", blk).group(1))
        start = int(re.search(r"startFace\s+(\d+)", blk).group(1))
        patches.append((name, _PATCH_TYPE_NAMES.get(t, PATCH_GENERIC), start, nfaces))
    n_cells = int(max(owner.max(), neighbour.max() if len(neighbour) else 0)) + 1
    return PolyMesh(points=points, face_offsets=np.array(offs, dtype=np.int32),
                    face_points=np.array(fpts, dtype=np.int32), owner=owner, neighbour=neighbour,
                    n_cells=n_cells, patches=patches)


def polymesh_from_npz(path: str) -> PolyMesh:
    """A polyMesh stored as arrays (points, face_offsets, face_points, owner,
    neighbour, n_cells, patch_names, patch_meta [type, start, size])."""
    with np.load(path, allow_pickle=False) as z:
        patches = [(str(n), int(t), int(s), int(c)) for n, (t, s, c) in zip(z["patch_names"], z["patch_meta"])]
        return PolyMesh(points=np.ascontiguousarray(z["points"]), face_offsets=np.ascontiguousarray(z["face_offsets"]),
                        face_points=np.ascontiguousarray(z["face_points"]), owner=np.ascontiguousarray(z["owner"]),
                        neighbour=np.ascontiguousarray(z["neighbour"]), n_cells=int(z["n_cells"]), patches=patches)


def write_polymesh(mesh: PolyMesh, out_dir: str) -> None:
    """Write the mesh as an ASCII foam polyMesh."""
    os.makedirs(out_dir, exist_ok=True)

    def header(cls, obj):
        return ("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
                f"    class       {cls};\n    location    \"constant/polyMesh\";\n    object      {obj};\n}}\n\n")

    with open(os.path.join(out_dir, "points"), "w") as fh:
        fh.write(header("vectorField", "points"))
        fh.write(f"{len(mesh.points)}\n(\n")
        for p in mesh.points:
            fh.write(f"({p[0]:.17g} {p[1]:.17g} {p[2]:.17g})\n")
        fh.write(")\n")
    with open(os.path.join(out_dir, "faces"), "w") as fh:
        fh.write(header("faceList", "faces"))
        fh.write(f"{mesh.n_faces}\n(\n")
        for f in range(mesh.n_faces):
            v = mesh.face_points[mesh.face_offsets[f]:mesh.face_offsets[f + 1]]
            fh.write(f"{len(v)}(" + " ".join(map(str, v)) + ")\n")
        fh.write(")\n")
    for name, arr in (("owner", mesh.owner), ("neighbour", mesh.neighbour)):
        with open(os.path.join(out_dir, name), "w") as fh:
            fh.write(header("labelList", name))
            fh.write(f"{len(arr)}\n(\n" + "\n".join(map(str, arr)) + "\n)\n")
    inv = {v: k for k, v in _PATCH_TYPE_NAMES.items() if k != "symmetry"}
    with open(os.path.join(out_dir, "boundary"), "w") as fh:
        fh.write(header("polyBoundaryMesh", "boundary"))
        fh.write(f"{len(mesh.patches)}\n(\n")
        for name, t, start, size in mesh.patches:
            fh.write(f"    {name}\n    {{\n        type            {inv[t]};\n"
                     f"        nFaces          {size};\n        startFace       {start};\n    }}\n")
        fh.write(")\n")


# --------------------------------------------------------------------------- 3-D swept wing
def naca4_thickness(x: np.ndarray, t: float) -> np.ndarray:
    """Half-thickness of the analytic symmetric 4-digit section with a closed
    trailing edge (last coefficient -0.1036), x in [0, 1]."""
    return 5 * t * (0.2969 * np.sqrt(np.clip(x, 0, 1)) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3
                    - 0.1036 * x ** 4)


def _stretch(n: int, first: float, total: float) -> np.ndarray:
    """n+1 nodes on [0, total] with the first interval 'first', geometric growth
    (ratio found by bisection)."""
    if n < 1:
        raise ValueError(f"_stretch: {n} intervals")
    if first * n >= total:
        return np.linspace(0.0, total, n + 1)
    lo, hi = 1.0 + 1e-12, 2.0
    while (hi ** n - 1) / (hi - 1) * first < total:
        hi *= 1.5
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if (r ** n - 1) / (r - 1) * first > total:
            hi = r
        else:
            lo = r
    r = 0.5 * (lo + hi)
    d = first * r ** np.arange(n)
    s = np.concatenate([[0.0], np.cumsum(d)])
    return s * (total / s[-1])


def naca4_slope(x: np.ndarray, t: float) -> np.ndarray:
    """d/dx of naca4_thickness (x > 0)."""
    return 5 * t * (0.2969 * 0.5 / np.sqrt(x) - 0.1260 - 2 * 0.3516 * x + 3 * 0.2843 * x ** 2 - 4 * 0.1036 * x ** 3)


def _smoothstep(x):
    x = np.clip(x, 0.0, 1.0)
    return x * x * (3.0 - 2.0 * x)


def swept_wing(n_wake: int = 48, n_surf: int = 128, n_normal: int = 96, n_span_wing: int = 44,
               n_span_out: int = 16, thickness: float = 0.10, sweep_deg: float = 30.0, taper: float = 0.56,
               semi_span: float = 1.0, root_chord: float = 1.0, far: float = 12.0, first_cell: float = 2.0e-6,
               n_layers: int = 32, layer_ratio: float = 1.15, normal_blend: float = 0.08, tip_blend: int = 4,
               tip_first: float | None = None, tip_round: float = 0.08, tip_power: float = 1.0,
               wake_growth: float = 0.01, order: str = "ijk") -> PolyMesh:
    """Structured C-H hex mesh around a semi-span swept, tapered wing.

    Each spanwise station is a C-mesh about the local section (analytic
    symmetric 4-digit thickness, 10 %, closed trailing edge), wrapped from the
    downstream far field under the wake, round the section and back over the
    wake; the wake cut is closed by sharing points.  Beyond the tip the
    section is a zero-thickness sheet in the chord plane, so the cells above
    and below the sheet share faces (internal).  The tip is rounded: over
    the outer tip_round of the span the section thickness falls as
    (1 - s^2)^tip_power (s from 0 at span - tip_round to 1 at the tip
    station), so the wing surface closes at the tip station itself and no
    cell is a wedge between the surface and the sheet.

    Wall-normal lines leave the section along the outward normal of the
    local (thinned) section, turn to the full-thickness section's normal
    above the boundary layer (over 4 layer thicknesses; this spreads the
    leading-edge fan of the thin tip sections over several cells) and bend
    smoothly (over the first normal_blend of the way out) into the straight
    line to their far-field point, so the boundary-layer cells are
    orthogonal to the wall; at the tip station the sheet's lines leave it
    straight up or down (the leading edge's upstream), so no first cell
    lies along the sheet.  The far-field points are spaced evenly in the
    section's cosine parameter (a half circle of radius far about the
    trailing edge), and over the wake their spacing grows from the last
    section spacing.  The first cell is first_cell high (root chords, a y+
    of about 1 at Re 1.2e7), n_layers grow at layer_ratio, then the spacing
    stretches to the far field.  Over the wake the first spacing grows with
    the distance s behind the trailing edge, first_cell + wake_growth s (at
    most far/400), joined to the wall distribution by a smooth maximum, so
    the wake cells have aspect ratios of tens rather than the boundary
    layer's thousands.  Beyond the tip the sheet has no wall: its lines take
    the full-thickness fan round a virtual leading edge and the first
    spacing grows over tip_blend stations to tip_first (default four times
    the tip chord's leading-edge spacing).  The tip station's leading edge
    is a 180-degree corner of the sheet (the C-H topology's collapsed tip):
    the few cells at it are the most non-orthogonal of the mesh (86 deg at
    the default size; every cell volume and wall-normal distance positive).
    Cells are numbered i (round the C) fastest, then j (off the wall), then k
    (span) with order="ijk"; order="tiles" numbers them by 4 x 4 x 4 tiles
    (64 cells, one wavefront), the tiles in Morton order, i fastest inside a
    tile — a renumbering (as foam's renumberMesh is) that keeps every
    cell's six neighbours within a few thousand cells of it.
    Patches: wing (wall), symmetry (symmetryPlane, z = 0),
    farfield (patch).  Default size 352 x 96 x 60 = 2 027 520 cells.
    """
    ni = 2 * n_wake + 2 * n_surf
    nj, nk = n_normal, n_span_wing + n_span_out
    # chordwise distribution on the section (cosine clustering at LE and TE)
    beta = np.linspace(0.0, np.pi, n_surf + 1)
    xs = 0.5 * (1 - np.cos(beta))                   # 0 (LE) -> 1 (TE)
    # wake distribution from TE to the downstream boundary
    xw = 1.0 + _stretch(n_wake, xs[-1] - xs[-2], far)
    # far-field spacing over the wake: from the last section arc spacing
    sw = _stretch(n_wake, far * 0.5 * np.pi / n_surf, far)
    # i index: 0..n_wake wake lower (far -> TE), then lower surface TE -> LE,
    # upper surface LE -> TE, wake upper TE -> far
    x_lower = np.concatenate([xw[::-1], xs[::-1][1:]])          # wake(far..TE), surface(TE..LE)
    x_upper = np.concatenate([xs[1:], xw[1:]])                  # surface(LE..TE), wake(TE..far)
    xc = np.concatenate([x_lower, x_upper])                     # ni+1 values
    bc = np.concatenate([np.full(n_wake + 1, np.pi), beta[::-1][1:], beta[1:], np.full(n_wake, np.pi)])
    so = np.concatenate([sw[::-1], np.zeros(n_surf), np.zeros(n_surf), sw[1:]])
    side = np.concatenate([-np.ones(len(x_lower)), np.ones(len(x_upper))])
    on_surf = (xc <= 1.0 + 1e-12)
    assert len(xc) == ni + 1
    xcs = np.clip(xc, 0.0, 1.0)
    slope_full = np.where(xcs > 0, naca4_slope(np.where(xcs > 0, xcs, 1.0), thickness), 0.0)
    le = on_surf & (xcs == 0.0)
    # the closed trailing edge is a concave corner of the flow region (the
    # surface normals there tilt aft by ~7 deg, the wake's are vertical): turn
    # the normals to vertical over the rear 40 % of the chord, so that the
    # normal lines there diverge instead of crossing
    g = np.where(on_surf, _smoothstep((xcs - 0.6) / 0.4), 1.0)

    def section_normals(th):
        """Unit outward normals (chord-independent) of the section scaled to
        th of the full thickness: the leading edge points upstream, the wake
        lines straight up/down.  At th = 0 (the tip station and the sheet
        beyond) every point but the leading edge leaves the chord plane
        straight up or down, so no first cell lies along the sheet."""
        nx_ = np.where(on_surf, -th * slope_full, 0.0)
        ny_ = side.copy()
        nx_[le], ny_[le] = -1.0, 0.0
        nn = np.hypot(nx_, ny_)
        nx_, ny_ = nx_ / nn, ny_ / nn
        nx_, ny_ = (1.0 - g) * nx_, (1.0 - g) * ny_ + g * side
        nn = np.hypot(nx_, ny_)
        return nx_ / nn, ny_ / nn
    # wall-normal distribution (distance / far): n_layers geometric layers
    # from first_cell, then stretched to the far field
    def normal_dist(first):
        d_layers = first * layer_ratio ** np.arange(n_layers)
        s_in = np.concatenate([[0.0], np.cumsum(d_layers)])
        s_out = _stretch(nj - n_layers, d_layers[-1] * layer_ratio, far - s_in[-1])
        return np.concatenate([s_in, s_in[-1] + s_out[1:]]) / far
    t_wall = normal_dist(first_cell)

    def dists(first):
        """Distributions ([nj + 1, ni + 1]) whose first spacing is first[i]
        >= first_cell: a smooth maximum (cube norm) of the uniform spacing
        first[i] and the wall distribution, rescaled to end at the far field,
        so the spacing passes continuously from first[i] to the wall
        distribution's outer levels."""
        J = np.arange(nj + 1, dtype=np.float64)[:, None]
        lin = J * (np.asarray(first, dtype=np.float64)[None, :] / far)
        t = np.cbrt(lin ** 3 + t_wall[:, None] ** 3)
        return t / t[-1:, :]
    # first spacing per point: first_cell on the wing; over the wake it grows
    # with the distance s downstream of the trailing edge (first_cell +
    # wake_growth s, at most far/400), so the wake cells have aspect ratios of
    # tens instead of the boundary layer's thousands
    s_wake = np.where(on_surf, 0.0, (xc - 1.0) * root_chord)
    f_base = np.minimum(first_cell + wake_growth * s_wake, far / 400.0)
    f_base = np.maximum(f_base, first_cell)
    t_base = dists(f_base)
    # spanwise stations: half uniform, half sine (finer towards the tip)
    th_sp = np.linspace(0, np.pi / 2, n_span_wing + 1)
    zw = semi_span * 0.5 * (np.sin(th_sp) + th_sp / (np.pi / 2))
    zw[-1] = semi_span
    zo = semi_span + _stretch(n_span_out, zw[-1] - zw[-2], far)[1:]
    zk = np.concatenate([zw, zo])
    if tip_first is None:
        tip_first = 4.0 * taper * root_chord * xs[1]
    # beyond the tip the sheet has no wall: the first spacing there is at
    # least tip_first, blended in over tip_blend stations
    t_tip = dists(np.maximum(f_base, tip_first))
    tan_sw = np.tan(np.deg2rad(sweep_deg))
    nfx, nfy = section_normals(1.0)
    t_fan = 4.0 * t_wall[n_layers]
    P = np.zeros((nk + 1, nj + 1, ni + 1, 3))
    for k, z in enumerate(zk):
        zz = min(z, semi_span)
        chord = root_chord * (1 - (1 - taper) * zz / semi_span)
        x_le = zz * tan_sw
        beyond = z > semi_span + 1e-12
        if z >= semi_span - 1e-12:
            th = 0.0
        elif z > semi_span - tip_round:
            sr = (z - (semi_span - tip_round)) / tip_round
            th = (1.0 - sr * sr) ** tip_power
        else:
            th = 1.0
        nlx, nly = section_normals(th)
        b = _smoothstep((k - n_span_wing) / tip_blend) if beyond else 0.0
        t_norm = (1.0 - b) * t_base + b * t_tip
        # inner line: section surface (x in chord units) or the wake/sheet line
        xi = x_le + chord * np.minimum(xc, 1.0) + np.where(xc > 1.0, (xc - 1.0) * root_chord, 0.0)
        yi = np.where(on_surf, side * th * chord * naca4_thickness(np.minimum(xc, 1.0), thickness), 0.0)
        # outer line: a half circle of radius far about the local trailing
        # edge for the section points (psi = beta/2 from the upstream point
        # to the top/bottom), straight lines y = +-far above and below the wake
        x_te = x_le + chord
        psi = 0.5 * bc
        xo = np.where(on_surf, x_te - far * np.cos(psi), x_te + so)
        yo = np.where(on_surf, side * far * np.sin(psi), side * far)
        dx, dy = xo - xi, yo - yi
        D = np.hypot(dx, dy)
        for j, t in enumerate(t_norm):  # t: [ni + 1] at level j
            w = _smoothstep(t / normal_blend)
            # the local section's normals in the boundary layer, turning to
            # the full-thickness section's (the leading-edge fan spread over
            # several cells) above it; beyond the tip, where the first
            # spacing grows, the full-thickness fan from the sheet up
            v = np.maximum(_smoothstep(t / t_fan), b)
            nx_, ny_ = (1.0 - v) * nlx + v * nfx, (1.0 - v) * nly + v * nfy
            nn = np.hypot(nx_, ny_)
            nx_, ny_ = nx_ / nn, ny_ / nn
            P[k, j, :, 0] = xi + t * ((1.0 - w) * D * nx_ + w * dx)
            P[k, j, :, 1] = yi + t * ((1.0 - w) * D * ny_ + w * dy)
            P[k, j, :, 2] = z
    points = P.reshape(-1, 3)
    nI1, nJ1 = ni + 1, nj + 1

    def pid(i, j, k):
        return i + nI1 * (j + nJ1 * k)

    # merge the wake cut (j = 0, wake lower i <-> wake upper ni - i) and the
    # chord plane beyond the tip (j = 0, lower surface i <-> upper surface ni - i)
    remap = np.arange(points.shape[0], dtype=np.int64)
    for k in range(nk + 1):
        beyond = zk[k] >= semi_span - 1e-12   # the tip station is a sheet too
        for i in range(0, ni // 2 + 1):
            mirror = ni - i
            if mirror == i:
                continue
            is_wake = xc[i] >= 1.0 - 1e-12   # the closed trailing edge and the wake
            if is_wake or beyond:
                remap[pid(mirror, 0, k)] = pid(i, 0, k)
    # the leading-edge point beyond the tip is shared already (i = ni/2)
    ii, jj, kk = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    ii, jj, kk = ii.transpose(2, 1, 0).ravel(), jj.transpose(2, 1, 0).ravel(), kk.transpose(2, 1, 0).ravel()
    hexes = np.stack([pid(ii, jj, kk), pid(ii + 1, jj, kk), pid(ii + 1, jj + 1, kk), pid(ii, jj + 1, kk),
                      pid(ii, jj, kk + 1), pid(ii + 1, jj, kk + 1), pid(ii + 1, jj + 1, kk + 1),
                      pid(ii, jj + 1, kk + 1)], axis=1)
    hexes = remap[hexes]
    if order == "tiles":
        hexes = hexes[tile_order(ii, jj, kk)]
    elif order != "ijk":
        raise ValueError(f"swept_wing: order {order!r} (ijk or tiles)")
    # compact the point list
    used = np.unique(hexes)
    newid = np.full(points.shape[0], -1, dtype=np.int64)
    newid[used] = np.arange(len(used))
    hexes = newid[hexes]
    points = points[used]

    def classify(cells, lf, cen):
        out = np.full(len(cells), 2)          # far field by default
        out[lf == 2] = 0                      # j = 0 faces left unpaired: the wing
        out[(lf == 4) & (np.abs(cen[:, 2]) < 1e-12)] = 1  # z = 0 symmetry plane
        return out

    mesh = hex_polymesh(points, hexes, classify,
                        [("wing", PATCH_WALL), ("symmetry", PATCH_SYMMETRY), ("farfield", PATCH_GENERIC)])
    return mesh


def _morton3(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """Interleave the bits of three non-negative int arrays (a lowest)."""
    out = np.zeros(a.shape, dtype=np.int64)
    for bit in range(21):
        out |= ((a >> bit) & 1) << (3 * bit)
        out |= ((b >> bit) & 1) << (3 * bit + 1)
        out |= ((c >> bit) & 1) << (3 * bit + 2)
    return out


def tile_order(ii: np.ndarray, jj: np.ndarray, kk: np.ndarray, t: int = 4) -> np.ndarray:
    """Permutation of structured cells (i, j, k) into t x t x t tiles, the
    tiles in Morton order and the cells i fastest, then j, then k inside a
    tile: new cell n is old cell perm[n]."""
    key_t = _morton3(ii // t, jj // t, kk // t)
    key_c = (ii % t) + t * ((jj % t) + t * (kk % t))
    return np.lexsort((key_c, key_t))


def _face_triangles(mesh: PolyMesh, faces: np.ndarray):
    """Triangles (c, p_k, p_k+1) of the given faces, c the mean of the face's
    points (summed in point order, as the device does), one per edge: arrays
    [nT,3] of corner coordinates and the owning face of each triangle.  A
    warped quad is followed on both diagonals."""
    off = mesh.face_offsets
    A, B, Cc, owner = [], [], [], []
    for f in faces:
        pts = mesh.points[mesh.face_points[off[f]:off[f + 1]]]
        n = len(pts)
        c = pts[0].copy()
        for k in range(1, n):
            c = c + pts[k]
        c = c / n
        for k in range(n):
            A.append(c); B.append(pts[k]); Cc.append(pts[(k + 1) % n]); owner.append(f)
    return np.array(A), np.array(B), np.array(Cc), np.array(owner)


def _point_triangle_distance(P, A, B, Cc):
    """Exact distance from points P [n,3] to triangles (A,B,C) [n,3] each
    (Ericson, Real-Time Collision Detection, 5.1.5, closest point on a
    triangle), vectorised."""
    ab, ac, ap = B - A, Cc - A, P - A
    d1 = np.einsum("ij,ij->i", ab, ap); d2 = np.einsum("ij,ij->i", ac, ap)
    bp = P - B
    d3 = np.einsum("ij,ij->i", ab, bp); d4 = np.einsum("ij,ij->i", ac, bp)
    cp = P - Cc
    d5 = np.einsum("ij,ij->i", ab, cp); d6 = np.einsum("ij,ij->i", ac, cp)
    va = d3 * d6 - d5 * d4
    vb = d5 * d2 - d1 * d6
    vc = d1 * d4 - d3 * d2
    with np.errstate(divide="ignore", invalid="ignore"):
        denom = 1.0 / (va + vb + vc)
        v = vb * denom
        w = vc * denom
        Q = A + ab * v[:, None] + ac * w[:, None]                      # interior
        # vertex and edge regions
        t_ab = d1 / (d1 - d3)
        t_ac = d2 / (d2 - d6)
        t_bc = (d4 - d3) / ((d4 - d3) + (d5 - d6))
    regA = (d1 <= 0) & (d2 <= 0)
    regB = (d3 >= 0) & (d4 <= d3)
    regC = (d6 >= 0) & (d5 <= d6)
    regAB = (vc <= 0) & (d1 >= 0) & (d3 <= 0)
    regAC = (vb <= 0) & (d2 >= 0) & (d6 <= 0)
    regBC = (va <= 0) & ((d4 - d3) >= 0) & ((d5 - d6) >= 0)
    Q = np.where(regBC[:, None], B + (Cc - B) * t_bc[:, None], Q)
    Q = np.where(regAC[:, None], A + ac * t_ac[:, None], Q)
    Q = np.where(regAB[:, None], A + ab * t_ab[:, None], Q)
    Q = np.where(regC[:, None], Cc, Q)
    Q = np.where(regB[:, None], B, Q)
    Q = np.where(regA[:, None], A, Q)
    return np.linalg.norm(P - Q, axis=1)


def cell_centres_approx(mesh: PolyMesh) -> np.ndarray:
    """Area-weighted mean of each cell's face centroids (vertex averages):
    a cell-centre estimate for wall-distance work."""
    off = mesh.face_offsets
    nF = mesh.n_faces
    cnt = np.diff(off)
    fp = mesh.face_points
    fid = np.repeat(np.arange(nF), cnt)
    fc = np.zeros((nF, 3))
    np.add.at(fc, fid, mesh.points[fp])
    fc /= cnt[:, None]
    C = np.zeros((mesh.n_cells, 3))
    w = np.zeros(mesh.n_cells)
    np.add.at(C, mesh.owner, fc)
    np.add.at(w, mesh.owner, 1.0)
    np.add.at(C, mesh.neighbour, fc[:mesh.n_internal])
    np.add.at(w, mesh.neighbour, 1.0)
    return C / w[:, None]


def wall_distance(mesh: PolyMesh, wall_patches, centres: np.ndarray | None = None, k: int = 12) -> np.ndarray:
    """Distance of every cell centre to the nearest face of the wall patches
    (the wallDist d_ of the Spalart-Allmaras models).  Exact point-to-polygon
    distance over the k wall faces whose centroids are nearest (KD tree)."""
    from scipy.spatial import cKDTree
    C = cell_centres_approx(mesh) if centres is None else np.asarray(centres, dtype=np.float64)
    faces = []
    for name in wall_patches:
        _n, _t, start, size = mesh.patches[mesh.patch_index(name)]
        faces.extend(range(start, start + size))
    faces = np.array(faces, dtype=np.int64)
    if len(faces) == 0:
        raise ValueError("wall_distance: no wall faces")
    A, B, Cc, fown = _face_triangles(mesh, faces)
    off = mesh.face_offsets
    cen = np.array([mesh.points[mesh.face_points[off[f]:off[f + 1]]].mean(axis=0) for f in faces])
    tri_of_face = {}
    for t_i, f in enumerate(fown):
        tri_of_face.setdefault(f, []).append(t_i)
    maxt = max(len(v) for v in tri_of_face.values())
    # [n_wall_faces, maxt] triangle table (short faces repeat their first triangle)
    table = np.array([[tri_of_face[f][t] if t < len(tri_of_face[f]) else tri_of_face[f][0] for t in range(maxt)]
                      for f in faces], dtype=np.int64)
    kk = min(k, len(faces))
    _, nearest = cKDTree(cen).query(C, k=kk)
    nearest = nearest.reshape(len(C), kk)
    d = np.full(len(C), np.inf)
    for j in range(kk):
        for t in range(maxt):
            tri = table[nearest[:, j], t]
            d = np.minimum(d, _point_triangle_distance(C, A[tri], B[tri], Cc[tri]))
    return d


def hicks_henne(xhat: np.ndarray, n_modes: int, width: float = 3.0) -> np.ndarray:
    """Hicks-Henne bump functions sin(pi xhat^m_i)^width on xhat in [0, 1],
    peaks at t_i = (i + 1)/(n_modes + 1) (m_i = ln 0.5 / ln t_i); [n_modes, len(xhat)]."""
    x = np.clip(np.asarray(xhat, dtype=np.float64), 0.0, 1.0)
    out = np.zeros((n_modes, x.size))
    for i in range(n_modes):
        t = (i + 1.0) / (n_modes + 1.0)
        m = np.log(0.5) / np.log(t)
        out[i] = np.sin(np.pi * x ** m) ** width
    return out


def bump_channel(nx: int = 48, ny: int = 24, lx: float = 2.0, ly: float = 1.0, bump=(0.5, 1.5),
                 amplitudes=None, node_disp=None) -> PolyMesh:
    """A 2-D channel (one cell thick) whose bottom wall carries a bump
    y_b(x) = sum_i a_i HH_i((x - x0)/(x1 - x0)) over x in bump (Hicks-Henne
    modes), or explicit wall-normal displacements node_disp[nx + 1] of the
    bottom nodes; every grid line moves by y_b (1 - y/ly), so the top stays
    flat.  Patches: inlet (x = 0), outlet (x = lx), bottom (wall), top
    (symmetryPlane), frontAndBack (empty)."""
    dx, dy = lx / nx, ly / ny
    x = np.arange(nx + 1) * dx
    yb = np.zeros(nx + 1)
    if amplitudes is not None and len(amplitudes):
        a = np.asarray(amplitudes, dtype=np.float64)
        xh = (x - bump[0]) / (bump[1] - bump[0])
        inside = (xh > 0.0) & (xh < 1.0)
        yb = np.where(inside, a @ hicks_henne(xh, len(a)), 0.0)
    if node_disp is not None:
        yb = yb + np.asarray(node_disp, dtype=np.float64)
    X, Y = np.meshgrid(x, np.arange(ny + 1) * dy, indexing="ij")
    Y = Y + yb[:, None] * (1.0 - Y / ly)
    pts = np.zeros((2, ny + 1, nx + 1, 3))
    for k, z in enumerate((-0.05, 0.05)):
        pts[k, :, :, 0] = X.T
        pts[k, :, :, 1] = Y.T
        pts[k, :, :, 2] = z
    points = pts.reshape(-1, 3)

    def p(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    hexes = np.stack([p(ii, jj, 0), p(ii + 1, jj, 0), p(ii + 1, jj + 1, 0), p(ii, jj + 1, 0),
                      p(ii, jj, 1), p(ii + 1, jj, 1), p(ii + 1, jj + 1, 1), p(ii, jj + 1, 1)], axis=1)

    def classify(cells, lf, cen):
        out = np.full(len(cells), -1)
        out[(lf == 4) | (lf == 5)] = 4
        out[lf == 0] = 0
        out[lf == 1] = 1
        out[lf == 2] = 2
        out[lf == 3] = 3
        return out

    return hex_polymesh(points, hexes, classify,
                        [("inlet", PATCH_GENERIC), ("outlet", PATCH_GENERIC), ("bottom", PATCH_WALL),
                         ("top", PATCH_SYMMETRY), ("frontAndBack", PATCH_EMPTY)])


# ------------------------------------------------------------------ general cells
PRISM_FACES = [[0, 2, 1], [3, 4, 5], [0, 1, 4, 3], [1, 2, 5, 4], [2, 0, 3, 5]]


def cells_polymesh(points: np.ndarray, cells: list, classify, patch_defs) -> PolyMesh:
    """Build a polyMesh from cells given as lists of outward-oriented faces
    (point-label lists of 3 or 4 points): faces shared by two cells become
    internal faces (owner = lower cell label, owner's orientation), ordered by
    owner then neighbour; the rest are boundary faces, classified by
    classify(cell_ids, local_face_ids, face_centres) and grouped by patch."""
    fc, fl, fp = [], [], []
    for ci, faces in enumerate(cells):
        for li, f in enumerate(faces):
            fc.append(ci); fl.append(li)
            fp.append(list(f) + [-1] * (4 - len(f)))
    return faces_polymesh(points, np.array(fc, dtype=np.int64), np.array(fl, dtype=np.int64),
                          np.array(fp, dtype=np.int64), len(cells), classify, patch_defs)


def faces_polymesh(points: np.ndarray, fc: np.ndarray, fl: np.ndarray, fp: np.ndarray, n_cells: int, classify,
                   patch_defs) -> PolyMesh:
    """cells_polymesh on arrays: face k belongs to cell fc[k] as its local
    face fl[k] with the outward-oriented points fp[k] (-1 padded to 4)."""
    key = np.sort(np.where(fp < 0, np.iinfo(np.int64).max, fp), axis=1)
    order = np.lexsort((key[:, 3], key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    pi = np.nonzero(same)[0]
    a, b = order[pi], order[pi + 1]
    ca, cb = fc[a], fc[b]
    own_slot = np.where(ca < cb, a, b)
    own_cell = np.minimum(ca, cb)
    nei_cell = np.maximum(ca, cb)
    io = np.lexsort((nei_cell, own_cell))
    own_slot, own_cell, nei_cell = own_slot[io], own_cell[io], nei_cell[io]
    paired = np.zeros(len(fc), dtype=bool)
    paired[a] = True
    paired[b] = True
    bs = np.nonzero(~paired)[0]
    npts = (fp >= 0).sum(axis=1)
    if len(bs):
        fpb = fp[bs]
        pts = points[np.where(fpb < 0, fpb[:, :1], fpb)]          # padding repeats a real point
        cen = (pts * (fpb >= 0)[:, :, None]).sum(axis=1) / npts[bs][:, None]
    else:
        cen = np.zeros((0, 3))
    bpatch = np.asarray(classify(fc[bs], fl[bs], cen), dtype=np.int64)
    if np.any(bpatch < 0):
        raise ValueError("unclassified boundary faces")
    ob = np.lexsort((fl[bs], fc[bs], bpatch))
    bs, bpatch = bs[ob], bpatch[ob]
    slots = np.concatenate([own_slot, bs])
    sizes = npts[slots]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    fps = fp[slots]
    fpts = fps[fps >= 0].astype(np.int32)                            # row-major: each face's points in order
    owner = np.concatenate([own_cell, fc[bs]]).astype(np.int32)
    patches = []
    start = len(own_cell)
    for k, (name, ptype) in enumerate(patch_defs):
        size = int(np.count_nonzero(bpatch == k))
        patches.append((name, ptype, start, size))
        start += size
    return PolyMesh(points=np.ascontiguousarray(points, dtype=np.float64), face_offsets=offsets, face_points=fpts,
                    owner=owner, neighbour=nei_cell.astype(np.int32), n_cells=int(n_cells), patches=patches)


def _outward(points, faces, apex):
    """Reverse the faces [k, npts] whose normal (first three points) points
    towards the cell point apex[k] (a point strictly inside or on the other
    side), so that every face points out of its cell."""
    a, b, c = points[faces[:, 0]], points[faces[:, 1]], points[faces[:, 2]]
    inward = np.einsum("ij,ij->i", np.cross(b - a, c - a), apex - a) > 0
    out = faces.copy()
    out[inward] = out[inward][:, ::-1]
    return out


def tet_prism_box(nx: int = 6, ny: int = 5, nz: int = 4, prism_layers: int = 1, lx: float = 0.01,
                  ly: float = 0.008, lz: float = 0.006, first: float = 0.0, jitter: float = 0.0,
                  seed: int = 1) -> PolyMesh:
    """A 3-D box of tetrahedra over prism layers: the structured nx x ny x nz
    hex grid (z layers graded from a first height 'first' when given), the
    lowest prism_layers layers each hex cut into two triangular prisms along
    the (i, j)-(i+1, j+1) diagonal of its horizontal faces, every hex above
    into the six tetrahedra about its (0,0,0)-(1,1,1) diagonal (Kuhn's
    subdivision: neighbouring hexes cut their shared faces alike, and its
    horizontal cuts match the prisms' triangles).  Interior points jittered by
    up to 'jitter' of the local spacing (seeded).  Patches: bottom (wall, z = 0),
    top, left, right, front, back (generic)."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, lx, nx + 1)
    y = np.linspace(0.0, ly, ny + 1)
    if first > 0.0:
        r = _stretch(nz, first, lz) if nz > 1 else np.array([0.0, lz])
        z = np.asarray(r, dtype=np.float64)
    else:
        z = np.linspace(0.0, lz, nz + 1)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    P = np.stack([X, Y, Z], axis=-1)
    if jitter > 0.0:
        dx, dy = lx / nx, ly / ny
        dz = np.diff(z).min()
        inner = np.zeros(P.shape[:3], dtype=bool)
        inner[1:-1, 1:-1, 1:-1] = True
        d = rng.uniform(-1.0, 1.0, P.shape) * jitter * np.array([dx, dy, dz])
        P = P + d * inner[..., None]
    points = P.reshape(-1, 3)
    pid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    corner = lambda a, b, c: pid(I + a, J + b, K + c)
    prism = K < prism_layers
    cells_f = []   # (faces [k, 4] padded, cell id per face, local id)
    # prisms: (000, 100, 110) and (000, 110, 010) at the bottom, the same on top
    ncell = 0
    fc_l, fl_l, fp_l = [], [], []
    pr = np.nonzero(prism)[0]
    for tri in ((corner(0, 0, 0), corner(1, 0, 0), corner(1, 1, 0)), (corner(0, 0, 0), corner(1, 1, 0), corner(0, 1, 0))):
        a, b, c = (t[pr] for t in tri)
        up = pid(0, 0, 1) - pid(0, 0, 0)
        A, B, Cc = a + up, b + up, c + up
        cen = (points[a] + points[b] + points[c] + points[A] + points[B] + points[Cc]) / 6.0
        faces = [np.stack([a, b, c, -np.ones_like(a)], 1), np.stack([A, B, Cc, -np.ones_like(a)], 1),
                 np.stack([a, b, B, A], 1), np.stack([b, c, Cc, B], 1), np.stack([c, a, A, Cc], 1)]
        ids = ncell + np.arange(len(a))
        for li, f in enumerate(faces):
            ff = _outward(points, f, cen)
            fc_l.append(ids); fl_l.append(np.full(len(a), li)); fp_l.append(ff)
        ncell += len(a)
    # tetrahedra: 000 -> 111 through each permutation of the axes
    te = np.nonzero(~prism)[0]
    for perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
        steps = [np.zeros(3, int)]
        for ax in perm:
            v = steps[-1].copy()
            v[ax] = 1
            steps.append(v)
        v = [corner(*s)[te] for s in steps]
        cen = sum(points[q] for q in v) / 4.0
        faces = [np.stack([v[0], v[1], v[2], -np.ones_like(v[0])], 1), np.stack([v[0], v[1], v[3], -np.ones_like(v[0])], 1),
                 np.stack([v[0], v[2], v[3], -np.ones_like(v[0])], 1), np.stack([v[1], v[2], v[3], -np.ones_like(v[0])], 1)]
        ids = ncell + np.arange(len(te))
        for li, f in enumerate(faces):
            ff = _outward(points, f, cen)
            fc_l.append(ids); fl_l.append(np.full(len(te), li)); fp_l.append(ff)
        ncell += len(te)
    fc = np.concatenate(fc_l).astype(np.int64)
    fl = np.concatenate(fl_l).astype(np.int64)
    fp = np.concatenate(fp_l).astype(np.int64)
    tol = 1e-12 * max(lx, ly, lz)

    def classify(cells_, lf, cen):
        out = np.full(len(cells_), -1)
        out[np.abs(cen[:, 2]) < tol] = 0
        out[np.abs(cen[:, 2] - lz) < tol] = 1
        out[np.abs(cen[:, 0]) < tol] = 2
        out[np.abs(cen[:, 0] - lx) < tol] = 3
        out[np.abs(cen[:, 1]) < tol] = 4
        out[np.abs(cen[:, 1] - ly) < tol] = 5
        return out

    return faces_polymesh(points, fc, fl, fp, ncell, classify,
                          [("bottom", PATCH_WALL), ("top", PATCH_GENERIC), ("left", PATCH_GENERIC),
                           ("right", PATCH_GENERIC), ("front", PATCH_GENERIC), ("back", PATCH_GENERIC)])


def naca4_blunt(x: np.ndarray, t: float) -> np.ndarray:
    """Half-thickness of the symmetric 4-digit section with its open (blunt)
    trailing edge (last coefficient -0.1015)."""
    return 5 * t * (0.2969 * np.sqrt(np.clip(x, 0, 1)) - 0.1260 * x - 0.3516 * x ** 2 + 0.2843 * x ** 3
                    - 0.1015 * x ** 4)


def _zip_rings(ia: np.ndarray, ua: np.ndarray, ib: np.ndarray, ub: np.ndarray) -> list:
    """Triangles of the strip between two closed rings of point labels ia, ib
    carrying increasing parameters ua, ub in [0, 1) along the rings (both start
    at parameter 0): the two point sequences are merged by parameter, so every
    point of both rings is used and the strip shares the rings' edges."""
    na, nb = len(ia), len(ib)
    uaw = np.concatenate([ua, [1.0 + ua[0]]])
    ubw = np.concatenate([ub, [1.0 + ub[0]]])
    tris = []
    i = j = 0
    while i < na or j < nb:
        if j >= nb or (i < na and uaw[i + 1] <= ubw[j + 1]):
            tris.append((ia[i % na], ia[(i + 1) % na], ib[j % nb]))
            i += 1
        else:
            tris.append((ia[i % na], ib[(j + 1) % nb], ib[j % nb]))
            j += 1
    return tris


def _polygon_cell(poly2d: list, nb2: int, z_off: int) -> list:
    """Outward faces of the prism/hex extruded from a counter-clockwise 2-D
    polygon of point labels (bottom labels l, top labels l + z_off)."""
    v = list(poly2d)
    k = len(v)
    faces = [[v[i] for i in range(k)][::-1], [v[i] + z_off for i in range(k)]]
    for i in range(k):
        a, b = v[i], v[(i + 1) % k]
        faces.append([a, b, b + z_off, a + z_off])
    return faces


def airfoil_loop(n_surf: int = 400, thickness: float = 0.12):
    """The wall loop of hybrid_airfoil: points [ns, 2] in the loop's
    (counter-clockwise) order and each point's kind (0 lower surface, 1 upper
    surface, 2 blunt-base)."""
    beta = np.linspace(0.0, np.pi, n_surf // 2 + 1)
    xs = 0.5 * (1 - np.cos(beta))
    yt = naca4_blunt(xs, thickness)
    lower = np.stack([xs[::-1], -yt[::-1]], axis=1)          # TE -> LE (lower)
    upper = np.stack([xs[1:], yt[1:]], axis=1)                # LE -> TE (upper)
    h_te = yt[-1]
    n_base = max(2, int(np.ceil(2 * h_te / (xs[-1] - xs[-2]))))
    base = np.stack([np.ones(n_base - 1), np.linspace(h_te, -h_te, n_base + 1)[1:-1]], axis=1)
    loop = np.concatenate([lower, upper, base], axis=0)
    kind = np.concatenate([np.zeros(len(lower), int), np.ones(len(upper), int), np.full(len(base), 2)])
    # counter-clockwise (signed area > 0)
    area = 0.5 * np.sum(loop[:, 0] * np.roll(loop[:, 1], -1) - np.roll(loop[:, 0], -1) * loop[:, 1])
    if area < 0:
        loop, kind = loop[::-1].copy(), kind[::-1].copy()
    return loop, kind


def hybrid_airfoil(n_surf: int = 400, n_layers: int = 40, first_cell: float = 2.0e-6, ratio: float = 1.15,
                   far: float = 50.0, thickness: float = 0.12, outer_ratio: float = 1.15, layer_aspect: float = 1.8,
                   jitter: float = 0.2, split_aspect: float = 2.5, seed: int = 1, depth: float = 0.1,
                   wall_disp=None) -> PolyMesh:
    """Hybrid mesh about a 12 % symmetric 4-digit section (blunt trailing edge):
    an O-grid marched off the wall along smoothed normals, n_layers quad
    layers at 'ratio' from 'first_cell', then layers growing at 'outer_ratio'
    up to 'layer_aspect' times the median spacing along the ring and
    following it from there (the ring points moved along the ring towards uniform arc length by at
    most 0.3 of the layer height per layer, and beyond 5 chords pulled
    radially towards a circle) until the ring lies outside a circle of radius 'far' chords, the
    last ring projected onto that circle.  Outside the wall layers every
    O-grid quad whose aspect ratio is within 'split_aspect' is split into two
    triangles (diagonal alternating with the ring index; a split rectangle of
    aspect a has face non-orthogonality acos(2a / (1 + a^2))), the others
    stay quads; interior nodes outside the wall layers are jittered by up to
    'jitter' of the smaller of the local tangential and normal spacings,
    scaled down by the local aspect ratio (slivers are not jittered: their
    non-orthogonality would follow the height variation along them)
    (seeded).  Extruded one cell thick: hexes and prisms.  Patches: airfoil
    (wall), farfield (patch), frontAndBack (empty).  The defaults give about
    2e5 cells.

    wall_disp [ns] (airfoil_loop order): a displacement of each wall point
    along its wall-layer line (the direction from the wall point to the next
    ring's point, taken from the undisplaced mesh), carried into the wall
    layers with a weight falling linearly in arc length from 1 at the wall to
    0 at ring n_layers; the rest of the mesh and its topology are unchanged,
    so the displaced mesh is a linear function of wall_disp."""
    rng = np.random.default_rng(seed)
    loop, _ = airfoil_loop(n_surf, thickness)
    ns = len(loop)
    ctr = np.array([0.5, 0.0])
    rings = [loop]
    heights = []
    cur = loop.copy()
    h = first_cell
    k = 0
    while True:
        prv, nxt = np.roll(cur, 1, axis=0), np.roll(cur, -1, axis=0)
        tang = nxt - prv
        nrm = np.stack([tang[:, 1], -tang[:, 0]], axis=1)
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        for _ in range(4 if k < n_layers else 8):
            nrm = 0.5 * nrm + 0.25 * (np.roll(nrm, 1, axis=0) + np.roll(nrm, -1, axis=0))
            nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        new = cur + h * nrm
        if k >= n_layers:
            # move the points along the ring a little towards uniform arc length
            seg = np.linalg.norm(np.roll(new, -1, axis=0) - new, axis=1)
            sa = np.concatenate([[0.0], np.cumsum(seg)])
            shift = 0.05 * (np.arange(ns) * sa[-1] / ns - sa[:-1])
            tgt = sa[:-1] + np.clip(shift, -0.3 * h, 0.3 * h)    # shear of a layer <= 0.3
            closed = np.concatenate([new, new[:1]], axis=0)
            new = np.stack([np.interp(tgt, sa, closed[:, 0]), np.interp(tgt, sa, closed[:, 1])], axis=1)
            # beyond 5 chords pull the ring towards a circle about the section
            d = new - ctr
            rad = np.linalg.norm(d, axis=1)
            if rad.mean() > 5.0:
                new = ctr + d * ((rad + 0.1 * (rad.mean() - rad)) / rad)[:, None]
        heights.append(h)
        k += 1
        if np.min(np.linalg.norm(new - ctr, axis=1)) >= far:
            d = new - ctr
            new = ctr + far * d / np.linalg.norm(d, axis=1)[:, None]
            rings.append(new)
            break
        rings.append(new)
        cur = new
        if k < n_layers:
            h *= ratio
        else:
            # grow until the layer height reaches 'layer_aspect' times the
            # median spacing along the ring, then follow that spacing
            med = np.median(np.linalg.norm(np.roll(new, -1, axis=0) - new, axis=1))
            h = min(h * outer_ratio, max(h, layer_aspect * med))
    n_rings = len(rings)
    P = np.stack(rings, axis=0)                                # [ring, i, 2]
    # jitter the interior nodes outside the wall layers
    for kk in range(n_layers + 1, n_rings - 1):
        r = P[kk]
        sp = np.linalg.norm(np.roll(r, -1, axis=0) - r, axis=1)
        tang = np.minimum(sp, np.roll(sp, 1))
        nrml = np.minimum(np.linalg.norm(P[kk + 1] - r, axis=1), np.linalg.norm(r - P[kk - 1], axis=1))
        loc = np.minimum(tang, nrml)
        iso = loc / np.maximum(tang, nrml)      # full jitter on isotropic cells, none on slivers
        ang = rng.uniform(0.0, 2.0 * np.pi, ns)
        mag = jitter * loc * iso * np.sqrt(rng.uniform(0.0, 1.0, ns))
        P[kk] = r + mag[:, None] * np.stack([np.cos(ang), np.sin(ang)], axis=1)
    pts2 = [P[kk] for kk in range(n_rings)]
    lab = lambda kk, i: kk * ns + (i % ns)
    cells2d = []
    for kk in range(n_rings - 1):
        for i in range(ns):
            q = [lab(kk, i), lab(kk, i + 1), lab(kk + 1, i + 1), lab(kk + 1, i)]
            if kk >= n_layers:
                a = np.linalg.norm(P[kk, (i + 1) % ns] - P[kk, i]) / max(np.linalg.norm(P[kk + 1, i] - P[kk, i]), 1e-300)
                if 1.0 / split_aspect <= a <= split_aspect:
                    if (kk + i) % 2 == 0:
                        cells2d += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
                    else:
                        cells2d += [[q[0], q[1], q[3]], [q[1], q[2], q[3]]]
                    continue
            cells2d.append(q)
    P2 = np.concatenate(pts2, axis=0)
    # counter-clockwise orientation of every 2-D cell
    for ci, c in enumerate(cells2d):
        q = P2[c]
        ar = 0.5 * np.sum(q[:, 0] * np.roll(q[:, 1], -1) - np.roll(q[:, 0], -1) * q[:, 1])
        if ar < 0:
            cells2d[ci] = c[::-1]
        elif ar == 0:
            raise ValueError("degenerate 2-D cell")
    nP = len(P2)
    points = np.zeros((2 * nP, 3))
    points[:nP, :2] = P2
    points[nP:, :2] = P2
    points[:nP, 2] = -0.5 * depth
    points[nP:, 2] = 0.5 * depth
    cells = [_polygon_cell(c, len(c), nP) for c in cells2d]

    def classify(cells_, lf, cen):
        out = np.full(len(cells_), -1)
        zf = np.abs(np.abs(cen[:, 2]) - 0.5 * depth) < 1e-12
        out[zf] = 2
        rr = np.linalg.norm(cen[:, :2] - ctr, axis=1)
        out[(~zf) & (rr > 0.99 * far)] = 1
        out[(~zf) & (rr < 2.0)] = 0
        return out

    mesh = cells_polymesh(points, cells, classify,
                          [("airfoil", PATCH_WALL), ("farfield", PATCH_GENERIC), ("frontAndBack", PATCH_EMPTY)])
    if wall_disp is not None:
        mesh = displace_airfoil_wall(mesh, wall_disp, n_layers)
    return mesh


def displace_airfoil_wall(mesh: PolyMesh, wall_disp, n_layers: int) -> PolyMesh:
    """hybrid_airfoil's mesh with its wall points displaced: point i of the
    wall loop (airfoil_loop order, len(wall_disp) points) moves by
    wall_disp[i] along its wall-layer line (the unit vector from the wall point
    to the next ring's point of the given mesh), and the points of that line
    up to ring n_layers by wall_disp[i] times a weight falling linearly in arc
    length from 1 at the wall to 0 at ring n_layers.  The topology is kept, so
    the points are a linear function of wall_disp."""
    d = np.asarray(wall_disp, dtype=np.float64)
    ns = len(d)
    nP = len(mesh.points) // 2
    if nP % ns:
        raise ValueError("wall_disp does not match the mesh's wall loop")
    P = mesh.points[:nP, :2].reshape(-1, ns, 2)
    L = min(n_layers, P.shape[0] - 1)
    seg = np.linalg.norm(np.diff(P[:L + 1], axis=0), axis=2)
    arc = np.concatenate([np.zeros((1, ns)), np.cumsum(seg, axis=0)], axis=0)
    fade = 1.0 - arc / arc[-1][None, :]
    line = P[1] - P[0]
    line = line / np.linalg.norm(line, axis=1)[:, None]
    delta = ((d[None, :, None] * fade[:, :, None]) * line[None, :, :]).reshape(-1, 2)
    pts = mesh.points.copy()
    pts[:(L + 1) * ns, :2] += delta
    pts[nP:nP + (L + 1) * ns, :2] += delta
    return replace(mesh, points=pts)
