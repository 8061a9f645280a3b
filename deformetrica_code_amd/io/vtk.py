"""ASCII legacy VTK POLYDATA and UNSTRUCTURED_GRID: the meshes, polylines and volume meshes
the reference's landmark objects hold.

The reference reads them with VTK's vtkPolyDataReader (DeformableObjectReader.cxx:58-145)
into vtkPoints of the file's declared type: ``POINTS n float`` coordinates are single
precision values, which the objects then hold as doubles.  ``read_vtk`` keeps that rounding
so that a template read here starts from the same numbers.  Attributes (POINT_DATA,
CELL_DATA, FIELD) are skipped.  Binary legacy files are refused.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_CELL_SECTIONS = ("VERTICES", "LINES", "POLYGONS", "TRIANGLE_STRIPS", "CELLS")
_TYPES = {"float": np.float32, "double": np.float64, "int": np.int32, "long": np.int64,
          "short": np.int16, "unsigned_char": np.uint8, "char": np.int8,
          "unsigned_int": np.uint32, "unsigned_short": np.uint16, "unsigned_long": np.uint64}


@dataclass
class PolyData:
    points: np.ndarray                       # (n, 3) float64
    cells: dict = field(default_factory=dict)  # section -> list of index arrays

    def triangles(self):
        """(F, 3) int32 vertex indices of the polygons, which must all be triangles."""
        polys = self.cells.get("POLYGONS", [])
        if any(len(c) != 3 for c in polys):
            raise ValueError("the mesh has polygons that are not triangles")
        return np.array(polys, dtype=np.int32).reshape(-1, 3)

    def segments(self):
        """(S, 2) int32 vertex indices of the polyline segments (each LINES cell split into
        consecutive pairs)."""
        segs = [(c[i], c[i + 1]) for c in self.cells.get("LINES", []) for i in range(len(c) - 1)]
        return np.array(segs, dtype=np.int32).reshape(-1, 2)

    def volume_cells(self, dimension):
        """(F, dimension + 1) int32 cells of a volume mesh: the triangles of a 2D mesh, or the
        tetrahedra (unstructured-grid CELLS of 4 points) of a 3D one."""
        key = "CELLS" if "CELLS" in self.cells else "POLYGONS"
        cl = self.cells.get(key, [])
        if any(len(c) != dimension + 1 for c in cl):
            raise ValueError("Not an admissible cell!")
        return np.array(cl, dtype=np.int32).reshape(-1, dimension + 1)

    def coordinates(self, dimension):
        """The first ``dimension`` coordinates of each point, as the reference's 2D objects
        take them."""
        return np.ascontiguousarray(self.points[:, :dimension])


def read_vtk(path):
    with open(path) as f:
        lines = f.read().splitlines()
    if len(lines) < 4 or not lines[0].startswith("# vtk DataFile"):
        raise ValueError(f"{path}: not a legacy VTK file")
    if lines[2].strip().upper() != "ASCII":
        raise ValueError(f"{path}: only ASCII legacy VTK files are read")
    if lines[3].split()[:2] not in (["DATASET", "POLYDATA"], ["DATASET", "UNSTRUCTURED_GRID"]):
        raise ValueError(f"{path}: not a POLYDATA or UNSTRUCTURED_GRID dataset")
    tok = " ".join(lines[4:]).split()
    i = 0
    points = None
    cells = {}
    while i < len(tok):
        key = tok[i].upper()
        if key == "POINTS":
            n, typ = int(tok[i + 1]), tok[i + 2].lower()
            if typ not in _TYPES:
                raise ValueError(f"{path}: point type {typ!r}")
            vals = np.array(tok[i + 3:i + 3 + 3 * n], dtype=np.float64)
            if vals.size != 3 * n:
                raise ValueError(f"{path}: {n} points declared, fewer found")
            points = vals.astype(_TYPES[typ]).astype(np.float64).reshape(n, 3)
            i += 3 + 3 * n
        elif key in _CELL_SECTIONS:
            n, size = int(tok[i + 1]), int(tok[i + 2])
            flat = np.array(tok[i + 3:i + 3 + size], dtype=np.int64)
            if flat.size != size:
                raise ValueError(f"{path}: {key} shorter than declared")
            out, j = [], 0
            for _ in range(n):
                k = int(flat[j])
                out.append(flat[j + 1:j + 1 + k].astype(np.int32))
                j += 1 + k
            if j != size:
                raise ValueError(f"{path}: {key} size mismatch")
            cells[key] = out
            i += 3 + size
        elif key == "CELL_TYPES":
            i += 2 + int(tok[i + 1])
        elif key in ("POINT_DATA", "CELL_DATA", "FIELD", "METADATA"):
            break
        else:
            raise ValueError(f"{path}: unexpected token {tok[i]!r}")
    if points is None:
        raise ValueError(f"{path}: no POINTS section")
    for key, cl in cells.items():
        for c in cl:
            if c.size and (c.min() < 0 or c.max() >= points.shape[0]):
                raise ValueError(f"{path}: {key} index out of range")
    return PolyData(points, cells)


def write_vtk(path, points, triangles=None, lines=None, title="vtk output"):
    """Writes points (n x 2 or n x 3, padded with zeros to 3) and triangles or polyline
    segments as ASCII legacy POLYDATA, coordinates in double precision."""
    pts = np.asarray(points, dtype=np.float64)
    if pts.shape[1] < 3:
        pts = np.hstack([pts, np.zeros((pts.shape[0], 3 - pts.shape[1]))])
    out = ["# vtk DataFile Version 3.0", title, "ASCII", "DATASET POLYDATA",
           f"POINTS {pts.shape[0]} double"]
    out += [" ".join("%.17g" % v for v in row) for row in pts]
    for key, cl in (("POLYGONS", triangles), ("LINES", lines)):
        if cl is None:
            continue
        cl = np.asarray(cl, dtype=np.int64)
        out.append(f"{key} {cl.shape[0]} {cl.shape[0] * (cl.shape[1] + 1)}")
        out += [f"{cl.shape[1]} " + " ".join(str(v) for v in row) for row in cl]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
