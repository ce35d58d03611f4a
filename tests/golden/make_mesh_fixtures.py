"""Writes the reference's example meshes as xz-compressed fixtures.

The reference ships two tetrahedral meshes as legacy VTK 5.1 ASCII data
files: examples/example2/sp4_.vtk (muMAG standard problem 4, 16,602 nodes,
48,569 tetrahedra) and examples/example3/sp_stt.vtk (the spin-transfer-torque
example, 10,269 nodes, 39,246 tetrahedra).  They are data, not code; the
GPU tests read them from tests/golden/ because the reference tree is not
present where those tests run.  Run from the repository root:

    python tests/golden/make_mesh_fixtures.py [/root/reference]
"""
import lzma
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = {"sp4_.vtk.xz": "examples/example2/sp4_.vtk", "sp_stt.vtk.xz": "examples/example3/sp_stt.vtk"}


def main(ref_root="/root/reference"):
    for out, rel in FILES.items():
        with open(os.path.join(ref_root, rel), "rb") as f:
            data = f.read()
        with lzma.open(os.path.join(HERE, out), "wb", preset=9 | lzma.PRESET_EXTREME) as f:
            f.write(data)
        print(out, len(data), "bytes ->", os.path.getsize(os.path.join(HERE, out)))


if __name__ == "__main__":
    main(*sys.argv[1:])
