"""Mesh and magnetisation files.

* ``read_vtk``: legacy VTK unstructured grids in ASCII, both the 4.x layout
  (``CELLS n size`` with a vertex count before each cell) and the 5.1 layout
  (``OFFSETS`` / ``CONNECTIVITY`` arrays), as the reference's example meshes
  use (VTKReader.cpp:119-181 reads them through the VTK library).  Only
  tetrahedra (cell type 10) are kept; a ``CELL_DATA`` integer scalar, when
  present, is returned as the element material.
* ``write_vtu`` / ``read_vtu``: ASCII VTK XML unstructured grids carrying the
  point vectors ``Magnetization``, the point scalars ``Material`` and the field
  value ``timeInPs`` (MeshWriter.cpp:174-196; VTKReader.cpp:151-176 reads the
  magnetisation and time back).
* ``write_dense`` / ``read_dense``: the ``.dns`` file of a dense matrix (two
  int32 dimensions, then the values column-major as float64), the format
  BEMprocessing.cpp:448-480 caches the boundary matrix in.
"""
from __future__ import annotations

import xml.etree.ElementTree as ET

import numpy as np

from .mesh import Mesh

VTK_TETRA = 10


def _tokens(lines, start, count, dtype):
    out = []
    i = start
    while len(out) < count:
        out.extend(lines[i].split())
        i += 1
    return np.array(out[:count], dtype=dtype), i


def read_vtk(path: str, scale: float = 1e-9):
    """Returns (Mesh, element_material or None).  A path ending in ``.xz`` is
    read through lzma."""
    if path.endswith(".xz"):
        import lzma
        with lzma.open(path, "rt") as f:
            lines = [ln.strip() for ln in f]
    else:
        with open(path) as f:
            lines = [ln.strip() for ln in f]
    if not lines or not lines[0].startswith("# vtk DataFile"):
        raise ValueError(f"{path}: not a legacy VTK file")
    if lines[2].upper() != "ASCII":
        raise ValueError(f"{path}: only ASCII legacy VTK is read")
    xyz = None
    conn = None
    types = None
    mat = None
    i = 3
    while i < len(lines):
        ln = lines[i]
        if not ln:
            i += 1
            continue
        key = ln.split()
        head = key[0].upper()
        if head == "POINTS":
            n = int(key[1])
            v, i = _tokens(lines, i + 1, 3 * n, np.float64)
            xyz = v.reshape(n, 3)
            continue
        if head == "CELLS":
            a, b = int(key[1]), int(key[2])
            nxt = lines[i + 1].split()[0].upper() if i + 1 < len(lines) else ""
            if nxt == "OFFSETS":  # 5.1: OFFSETS (a values) then CONNECTIVITY (b values)
                off, j = _tokens(lines, i + 2, a, np.int64)
                while not lines[j].upper().startswith("CONNECTIVITY"):
                    j += 1
                cv, i = _tokens(lines, j + 1, b, np.int64)
                conn = [cv[off[k]:off[k + 1]] for k in range(a - 1)]
            else:  # 4.x: a cells, b integers (count followed by ids)
                cv, i = _tokens(lines, i + 1, b, np.int64)
                conn = []
                p = 0
                for _ in range(a):
                    c = int(cv[p])
                    conn.append(cv[p + 1:p + 1 + c])
                    p += 1 + c
            continue
        if head == "CELL_TYPES":
            nc = int(key[1])
            types, i = _tokens(lines, i + 1, nc, np.int64)
            continue
        if head == "CELL_DATA":
            nc = int(key[1])
            j = i + 1
            while j < len(lines) and not lines[j].upper().startswith("SCALARS"):
                j += 1
            if j < len(lines):
                j += 1
                if lines[j].upper().startswith("LOOKUP_TABLE"):
                    j += 1
                mat, i = _tokens(lines, j, nc, np.float64)
                mat = mat.astype(np.int64)
                continue
            i = j
            continue
        i += 1
    if xyz is None or conn is None or types is None:
        raise ValueError(f"{path}: POINTS, CELLS and CELL_TYPES are required")
    keep = [k for k in range(len(conn)) if types[k] == VTK_TETRA and len(conn[k]) == 4]
    fel = np.array([conn[k] for k in keep], dtype=np.int32).reshape(-1, 4)
    if mat is not None:
        mat = mat[keep]
    return Mesh(np.ascontiguousarray(xyz), np.ascontiguousarray(fel), scale), mat


def _array_xml(parent, name, data, ncomp=1, dtype="Float64"):
    e = ET.SubElement(parent, "DataArray", type=dtype, Name=name, NumberOfComponents=str(ncomp), format="ascii")
    flat = np.asarray(data).reshape(-1)
    fmt = "%.17g" if dtype.startswith("Float") else "%d"
    e.text = " ".join(fmt % v for v in flat)
    return e


def write_vtu(path: str, mesh: Mesh, m3n=None, time_ps: float | None = None, node_material=None):
    """m3n: (3n,) SoA magnetisation (written as n x 3 'Magnetization')."""
    n, nt = mesh.n, mesh.ntet
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="0.1", byte_order="LittleEndian")
    grid = ET.SubElement(root, "UnstructuredGrid")
    if time_ps is not None:
        fd = ET.SubElement(grid, "FieldData")
        e = _array_xml(fd, "timeInPs", [time_ps])
        e.set("NumberOfTuples", "1")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(n), NumberOfCells=str(nt))
    pd = ET.SubElement(piece, "PointData")
    if m3n is not None:
        pd.set("Vectors", "Magnetization")
        _array_xml(pd, "Magnetization", np.asarray(m3n, np.float64).reshape(3, n).T, 3)
    if node_material is not None:
        pd.set("Scalars", "Material")
        _array_xml(pd, "Material", np.asarray(node_material, np.int32), 1, "Int32")
    pts = ET.SubElement(piece, "Points")
    _array_xml(pts, "Points", mesh.xyz, 3)
    cells = ET.SubElement(piece, "Cells")
    _array_xml(cells, "connectivity", mesh.fel, 1, "Int64")
    _array_xml(cells, "offsets", 4 * np.arange(1, nt + 1), 1, "Int64")
    _array_xml(cells, "types", np.full(nt, VTK_TETRA), 1, "UInt8")
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def read_vtu(path: str, scale: float = 1e-9):
    """Returns (Mesh, m3n or None, time_ps or None)."""
    root = ET.parse(path).getroot()
    grid = root.find("UnstructuredGrid")
    piece = grid.find("Piece")

    def arr(parent, name, dtype):
        for e in parent.iter("DataArray"):
            if e.get("Name") == name:
                if e.get("format", "ascii") != "ascii":
                    raise ValueError(f"{path}: only ASCII data arrays are read")
                return np.array((e.text or "").split(), dtype=dtype)
        return None

    xyz = arr(piece.find("Points"), "Points", np.float64).reshape(-1, 3)
    cells = piece.find("Cells")
    conn = arr(cells, "connectivity", np.int64)
    off = arr(cells, "offsets", np.int64)
    typ = arr(cells, "types", np.int64)
    start = np.concatenate([[0], off[:-1]])
    keep = (typ == VTK_TETRA) & (off - start == 4)
    fel = np.stack([conn[start[keep] + k] for k in range(4)], axis=1).astype(np.int32)
    m = None
    pd = piece.find("PointData")
    if pd is not None:
        mm = arr(pd, "Magnetization", np.float64)
        if mm is not None:
            m = np.ascontiguousarray(mm.reshape(-1, 3).T.reshape(-1))
    t = None
    fd = grid.find("FieldData")
    if fd is not None:
        tv = arr(fd, "timeInPs", np.float64)
        if tv is not None and tv.size:
            t = float(tv[0])
    return Mesh(np.ascontiguousarray(xyz), np.ascontiguousarray(fel), scale), m, t


def write_dense(path: str, B: np.ndarray):
    B = np.asarray(B, np.float64)
    with open(path, "wb") as f:
        np.array(B.shape, dtype=np.int32).tofile(f)
        np.asfortranarray(B).T.reshape(-1).tofile(f)  # column-major values


def read_dense(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        r, c = np.fromfile(f, dtype=np.int32, count=2)
        v = np.fromfile(f, dtype=np.float64, count=int(r) * int(c))
    if v.size != int(r) * int(c):
        raise ValueError(f"{path}: truncated dense matrix")
    return np.asfortranarray(v.reshape(int(c), int(r)).T)
