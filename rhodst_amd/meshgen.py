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
from dataclasses import dataclass, field

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
    all_faces = np.concatenate([internal_faces, bfaces], axis=0)
    owner = np.concatenate([own_cell, bcell]).astype(np.int32)
    patches = []
    start = internal_faces.shape[0]
    for pi, (name, ptype) in enumerate(patch_defs):
        size = int(np.count_nonzero(bpatch == pi))
        patches.append((name, ptype, start, size))
        start += size
    nF = all_faces.shape[0]
    return PolyMesh(points=np.ascontiguousarray(points, dtype=np.float64),
                    face_offsets=(4 * np.arange(nF + 1)).astype(np.int32),
                    face_points=all_faces.reshape(-1).astype(np.int32),
                    owner=owner, neighbour=nei_cell.astype(np.int32),
                    n_cells=nC, patches=patches)


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
