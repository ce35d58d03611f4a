"""Mesh and magnetisation files: legacy VTK (4.x and 5.1 layouts), VTK XML
snapshots with Magnetization / timeInPs, and the .dns dense-matrix file."""
import os

import numpy as np
import pytest

from tetmag_amd import mesh as M
from tetmag_amd import meshio as IO

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SP4 = os.path.join(GOLDEN, "sp4_.vtk.xz")  # examples/example2/sp4_.vtk
SP_STT = os.path.join(GOLDEN, "sp_stt.vtk.xz")  # examples/example3/sp_stt.vtk


def _write_legacy(path, mesh, layout, mat=None):
    n, nt = mesh.n, mesh.ntet
    with open(path, "w") as f:
        f.write(f"# vtk DataFile Version {'5.1' if layout == 5 else '4.2'}\ntest\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {n} double\n")
        for p in mesh.xyz:
            f.write(" ".join(repr(float(v)) for v in p) + "\n")
        if layout == 5:
            f.write(f"CELLS {nt + 1} {4 * nt}\nOFFSETS vtktypeint64\n")
            f.write(" ".join(str(4 * k) for k in range(nt + 1)) + "\n")
            f.write("CONNECTIVITY vtktypeint64\n")
            f.write(" ".join(str(v) for v in mesh.fel.reshape(-1)) + "\n")
        else:
            f.write(f"CELLS {nt + 1} {5 * nt + 4}\n")
            for c in mesh.fel:
                f.write("4 " + " ".join(str(v) for v in c) + "\n")
            f.write("3 0 1 2\n")  # a triangle, dropped on reading
        f.write(f"CELL_TYPES {nt + (0 if layout == 5 else 1)}\n")
        f.write("\n".join(["10"] * nt) + ("\n" if layout == 5 else "\n5\n"))
        if mat is not None:
            f.write(f"CELL_DATA {nt + (0 if layout == 5 else 1)}\nSCALARS material int 1\nLOOKUP_TABLE default\n")
            f.write(" ".join(str(v) for v in list(mat) + ([] if layout == 5 else [0])) + "\n")


@pytest.mark.parametrize("layout", [4, 5])
def test_legacy_vtk_round_trip(tmp_path, layout):
    mesh = M.kuhn_cuboid(4, 3, 2, 2.5, 0.1, 3)
    mat = np.arange(mesh.ntet) % 3
    p = tmp_path / f"m{layout}.vtk"
    _write_legacy(p, mesh, layout, mat)
    got, gm = IO.read_vtk(str(p))
    assert np.array_equal(got.xyz, mesh.xyz)
    assert np.array_equal(got.fel, mesh.fel)
    assert np.array_equal(gm, mat)


def test_vtu_snapshot_round_trip(tmp_path):
    mesh = M.kuhn_cuboid(3, 3, 2, 2.0, 0.1, 5)
    m = M.init_random_unit(mesh.n, 9)
    p = tmp_path / "snap.vtu"
    IO.write_vtu(str(p), mesh, m, time_ps=12.5, node_material=np.zeros(mesh.n, int))
    got, gm, t = IO.read_vtu(str(p))
    assert np.array_equal(got.xyz, mesh.xyz) and np.array_equal(got.fel, mesh.fel)
    assert np.array_equal(gm, m) and t == 12.5


def test_dense_matrix_file(tmp_path):
    B = np.random.default_rng(1).normal(size=(5, 7))
    p = tmp_path / "b.dns"
    IO.write_dense(str(p), B)
    raw = np.fromfile(str(p), dtype=np.int32, count=2)
    assert list(raw) == [5, 7]
    vals = np.fromfile(str(p), dtype=np.float64, offset=8)
    assert np.array_equal(vals, B.T.reshape(-1))  # column-major after the header
    assert np.array_equal(IO.read_dense(str(p)), B)


def test_reference_example_mesh():
    """The muMAG SP4 mesh of the reference's example 2 (legacy VTK 5.1)."""
    mesh, _ = IO.read_vtk(SP4)
    assert mesh.n == 16602 and mesh.ntet == 48569
    from tetmag_amd.autograd import element_volumes  # noqa: F401  (torch import is fine here)
    x = mesh.xyz
    f = mesh.fel
    vol = np.einsum("ij,ij->i", x[f[:, 1]] - x[f[:, 0]], np.cross(x[f[:, 2]] - x[f[:, 0]], x[f[:, 3]] - x[f[:, 0]]))
    assert np.all(np.abs(vol) > 0)
    assert abs(np.abs(vol).sum() / 6 - 500 * 125 * 3) < 1e-6 * 500 * 125 * 3
    bnd = M.boundary(mesh)
    assert bnd.bnx > 0 and bnd.nbel > 0


def test_reference_stt_example_mesh():
    """The spin-transfer-torque mesh of the reference's example 3 (legacy VTK 5.1)."""
    mesh, _ = IO.read_vtk(SP_STT)
    assert mesh.n == 10269 and mesh.ntet == 39246
    x, f = mesh.xyz, mesh.fel
    vol = np.einsum("ij,ij->i", x[f[:, 1]] - x[f[:, 0]], np.cross(x[f[:, 2]] - x[f[:, 0]], x[f[:, 3]] - x[f[:, 0]]))
    assert np.all(np.abs(vol) > 0)
