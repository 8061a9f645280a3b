"""ASCII legacy VTK POLYDATA and UNSTRUCTURED_GRID: the meshes, polylines and volume meshes
the reference's landmark objects hold.

The reference reads them with VTK's vtkPolyDataReader (DeformableObjectReader.cxx:58-145)
into vtkPoints of the file's declared type: ``POINTS n float`` coordinates are single
precision values, which the objects then hold as doubles.  ``read_vtk`` keeps that rounding
so that a template read here starts from the same numbers.  The point attributes SCALARS,
VECTORS and NORMALS are kept (PointCloud takes its weights from the point scalars); any
other attribute ends the read.  Binary legacy files are refused.
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
    point_data: list = field(default_factory=list)  # (kind, name, (n, ncomp) float64)
    point_data_complete: bool = True         # False when a point attribute was not parsed

    def point_weights(self):
        """One weight per point as PointCloud::UpdatePointWeights (PointCloud.cxx:203-226)
        takes them: the point scalars when the point data hold one component in all, else 1
        (a scalar count other than the point count cannot occur in a legacy file)."""
        n = self.points.shape[0]
        if not self.point_data_complete:
            raise ValueError("point attributes this reader does not parse")
        ncomp = sum(a.shape[1] for _, _, a in self.point_data)
        if ncomp == 1 and self.point_data[0][0] == "SCALARS":
            return self.point_data[0][2][:, 0].copy()
        return np.ones(n)

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
    point_data = []
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
        elif key == "POINT_DATA":
            i = _point_data(path, tok, i, point_data)
            break
        elif key in ("CELL_DATA", "FIELD", "METADATA"):
            break
        else:
            raise ValueError(f"{path}: unexpected token {tok[i]!r}")
    if points is None:
        raise ValueError(f"{path}: no POINTS section")
    for key, cl in cells.items():
        for c in cl:
            if c.size and (c.min() < 0 or c.max() >= points.shape[0]):
                raise ValueError(f"{path}: {key} index out of range")
    complete = not (point_data and point_data[-1] is None)
    return PolyData(points, cells, [a for a in point_data if a is not None], complete)


def _point_data(path, tok, i, out):
    """Parse the POINT_DATA section from tok[i]; appends (kind, name, values) per attribute
    and None when an attribute this reader does not parse ends it."""
    n = int(tok[i + 1])
    i += 2
    while i < len(tok):
        key = tok[i].upper()
        if key == "SCALARS":
            name, typ = tok[i + 1], tok[i + 2].lower()
            i += 3
            ncomp = 1
            if i < len(tok) and tok[i].isdigit():
                ncomp = int(tok[i])
                i += 1
            if i < len(tok) and tok[i].upper() == "LOOKUP_TABLE":
                i += 2
        elif key in ("VECTORS", "NORMALS"):
            name, typ, ncomp = tok[i + 1], tok[i + 2].lower(), 3
            i += 3
        else:
            if key != "CELL_DATA":
                out.append(None)
            return i
        if typ not in _TYPES:
            raise ValueError(f"{path}: attribute type {typ!r}")
        vals = np.array(tok[i:i + n * ncomp], dtype=np.float64)
        if vals.size != n * ncomp:
            raise ValueError(f"{path}: {key} {name} shorter than declared")
        out.append((key, name, vals.astype(_TYPES[typ]).astype(np.float64).reshape(n, ncomp)))
        i += n * ncomp
    return i


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
