"""Writing foam-extend time directories and force-coefficient logs from a case.

A converged (or intermediate) state leaves the device as ASCII volFields in
`<case>/<time>/` (p, T, U, rho and, with a turbulence model, nuTilda and
mut), in the format of the reference's tutorials (e.g. tutorials/RAE2822/0/p:
FoamFile header, dimensions, internalField, boundaryField), so the result can
be opened by foam tools and ParaView next to the polyMesh that
meshgen.write_polymesh writes.  Boundary entries are `calculated` with the
patch values the device holds (`empty` patches stay `empty`).  The force log
follows the columns of the forceCoeffs function object
(RAE2822/system/forces:19-38): Time, Cd, Cl.
"""
from __future__ import annotations

import os
import re

import numpy as np

from .meshgen import PATCH_EMPTY, PolyMesh

# SI dimension sets [kg m s K mol A cd]
DIMENSIONS = {
    "p": "[1 -1 -2 0 0 0 0]", "T": "[0 0 0 1 0 0 0]", "U": "[0 1 -1 0 0 0 0]", "rho": "[1 -3 0 0 0 0 0]",
    "nuTilda": "[0 2 -1 0 0 0 0]", "mut": "[1 -1 -1 0 0 0 0]",
}

# device fields: (cell field id, boundary field id or None, vector?)
_FIELDS = {"rho": (0, 16, False), "p": (3, 18, False), "T": (4, 19, False), "U": (5, 17, True),
           "nuTilda": (29, 33, False), "mut": (30, 34, False)}


def _header(cls: str, obj: str, location: str) -> str:
    return ("FoamFile\n{\n    version     2.0;\n    format      ascii;\n"
            f"    class       {cls};\n    location    \"{location}\";\n    object      {obj};\n}}\n\n")


def _list(values: np.ndarray, vector: bool) -> str:
    kind = "vector" if vector else "scalar"
    if vector:
        body = "\n".join(f"({v[0]:.17g} {v[1]:.17g} {v[2]:.17g})" for v in values)
    else:
        body = "\n".join(f"{v:.17g}" for v in values)
    return f"nonuniform List<{kind}>\n{len(values)}\n(\n{body}\n)"


def write_field(case_dir: str, time_name: str, name: str, mesh: PolyMesh, cell_values, boundary_values=None,
                dimensions: str | None = None) -> str:
    """Write one volScalarField / volVectorField.  boundary_values holds the
    values of the non-empty boundary faces in patch order (the compacted
    boundary of the device), or None for zero-gradient-like copies of the
    face cells."""
    vals = np.asarray(cell_values, dtype=np.float64)
    vector = vals.ndim == 2
    if vector and vals.shape[1] != 3:
        raise ValueError("vector fields are [n_cells][3]")
    if len(vals) != mesh.n_cells:
        raise ValueError(f"{name}: {len(vals)} values for {mesh.n_cells} cells")
    out_dir = os.path.join(case_dir, time_name)
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, name)
    bvals = None if boundary_values is None else np.asarray(boundary_values, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write(_header("volVectorField" if vector else "volScalarField", name, time_name))
        fh.write(f"dimensions      {dimensions or DIMENSIONS.get(name, '[0 0 0 0 0 0 0]')};\n\n")
        fh.write(f"internalField   {_list(vals, vector)};\n\nboundaryField\n{{\n")
        off = 0
        for pname, ptype, start, size in mesh.patches:
            fh.write(f"    {pname}\n    {{\n")
            if ptype == PATCH_EMPTY:
                fh.write("        type            empty;\n")
            else:
                if bvals is not None:
                    pv = bvals[off:off + size]
                else:
                    pv = vals[mesh.owner[start:start + size]]
                off += size
                fh.write("        type            calculated;\n")
                fh.write(f"        value           {_list(pv, vector)};\n")
            fh.write("    }\n")
        fh.write("}\n")
    return path


def write_time(case_dir: str, time_name: str, case, fields=None) -> list:
    """Write the state of a DeviceCase (or an object with the same get()) as
    `<case_dir>/<time_name>/{p,T,U,rho[,nuTilda,mut]}`; returns the paths."""
    mesh = case.spec.mesh
    if fields is None:
        fields = ["p", "T", "U", "rho"]
        if getattr(case.spec, "turbulence", None) is not None:
            fields += ["nuTilda", "mut"]
    paths = []
    for name in fields:
        cid, bid, vector = _FIELDS[name]
        cv = case.get(cid)
        bv = case.get(bid) if bid is not None else None
        if vector:
            cv = np.asarray(cv).reshape(-1, 3)
            if bv is not None:
                bv = np.asarray(bv).reshape(-1, 3)
        paths.append(write_field(case_dir, time_name, name, mesh, cv, bv))
    return paths


def read_field(path: str):
    """The internal field and the per-patch values of a volField written by
    write_field (nonuniform or uniform entries): (cells, {patch: values or
    None for empty})."""
    with open(path) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    vector = "volVectorField" in text.split("}", 1)[0]

    def parse_value(s):
        s = s.strip()
        m = re.match(r"uniform\s+(.*)", s, re.S)
        if m:
            v = m.group(1).strip().strip("()")
            return np.array(list(map(float, v.split()))) if vector else float(v)
        m = re.match(r"nonuniform\s+List<\w+>\s*(\d+)\s*\((.*)\)\s*$", s, re.S)
        if not m:
            raise ValueError(f"cannot parse field value in {path}")
        n = int(m.group(1))
        if vector:
            rows = re.findall(r"\(([^()]*)\)", m.group(2))
            arr = np.array([list(map(float, r.split())) for r in rows])
        else:
            arr = np.array(m.group(2).split(), dtype=np.float64)
        if len(arr) != n:
            raise ValueError(f"{path}: {len(arr)} values, list says {n}")
        return arr

    im = re.search(r"internalField\s+(.*?);\s*boundaryField", text, re.S)
    cells = parse_value(im.group(1))
    bf = text[text.index("boundaryField"):]
    patches = {}
    for name, blk in re.findall(r"(\w+)\s*\{([^{}]*)\}", bf):
        t = re.search(r"\btype\s+(\w+)", blk).group(1)
        vm = re.search(r"\bvalue\s+(.*?);", blk, re.S)
        patches[name] = None if t == "empty" or vm is None else parse_value(vm.group(1))
    return cells, patches


class ForceLog:
    """forceCoeffs-style coefficient log: one line per call of append()."""

    def __init__(self, path: str):
        self.path = path
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        with open(path, "w") as fh:
            fh.write("# Time          Cd              Cl\n")

    def append(self, time, cd: float, cl: float) -> None:
        with open(self.path, "a") as fh:
            fh.write(f"{time:<15} {cd:.10e} {cl:.10e}\n")

    @staticmethod
    def read(path: str) -> np.ndarray:
        return np.loadtxt(path, comments="#", ndmin=2)
