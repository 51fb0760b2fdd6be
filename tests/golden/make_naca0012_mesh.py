"""Extract the reference's airfoil mesh (``naca0012.vtk``, the input of
``src/airfoil.cpp`` after ``tools/vtk2CGNS``) as a compact fixture.

The file is a Gmsh legacy-VTK unstructured grid: 1395 points, and among its
2992 cells 2450 triangles (VTK type 5) plus the boundary lines and vertices.
The fixture keeps the point coordinates and the triangles' point indices,
which is all the solver reads from it (edges and boundary classes are
rebuilt from the triangles, as ``tools/vtk2CGNS.cpp:112-199`` and
``src/load_mesh.cpp:147-172`` do).  The tests write the fixture back out as a
triangle-only legacy-VTK file and read it with the library's reader, so the
reference scenario also runs where the reference tree is absent.

Run in a container where the reference tree is mounted at /root/reference:

    python tests/golden/make_naca0012_mesh.py
"""
import os

import numpy as np

REF = os.environ.get("DG_REFERENCE_ROOT", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def read_legacy_vtk(path):
    tok = open(path).read().split()
    i = tok.index("POINTS")
    npts = int(tok[i + 1])
    pts = np.array(tok[i + 3:i + 3 + 3 * npts], dtype=np.float64).reshape(npts, 3)
    i = tok.index("CELLS")
    ncells = int(tok[i + 1])
    k = i + 3
    cells = []
    for _ in range(ncells):
        n = int(tok[k])
        cells.append([int(v) for v in tok[k + 1:k + 1 + n]])
        k += 1 + n
    i = tok.index("CELL_TYPES")
    types = [int(v) for v in tok[i + 2:i + 2 + ncells]]
    tris = np.array([c for c, t in zip(cells, types) if t == 5], dtype=np.int32)
    return pts, tris


def main():
    pts, tris = read_legacy_vtk(os.path.join(REF, "naca0012.vtk"))
    assert pts.shape == (1395, 3) and tris.shape == (2450, 3)
    out = os.path.join(HERE, "naca0012_mesh.npz")
    np.savez_compressed(out, points=pts, triangles=tris)
    print(f"wrote {out}: {pts.shape[0]} points, {tris.shape[0]} triangles")


if __name__ == "__main__":
    main()
