"""Extract the reference's vortex mesh (``vortex/grid.vtk``, the input of
``vortex/src/vortex.cpp`` after ``vortex/tools``' CGNS conversion) as a
compact fixture.

The file is a Gmsh legacy-VTK unstructured grid of the box [0,10] x [-5,5]:
342 points, and among its cells 618 triangles (VTK type 5) plus the boundary
lines and vertices.  The fixture keeps the point coordinates and the
triangles' point indices; the solver rebuilds edges from the triangles
(``tools/vtk2CGNS.cpp:112-199``) and the vortex case gives every boundary
edge the exact state (``vortex/src/vortex.cpp:250-260``).  The tests write
the fixture back out as a triangle-only legacy-VTK file and read it with the
library's reader.

Run in a container where the reference tree is mounted at /root/reference:

    python tests/golden/make_vortex_grid_mesh.py
"""
import os

import numpy as np

from make_naca0012_mesh import read_legacy_vtk

REF = os.environ.get("DG_REFERENCE_ROOT", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    pts, tris = read_legacy_vtk(os.path.join(REF, "vortex", "grid.vtk"))
    assert pts.shape == (342, 3) and tris.shape == (618, 3), (pts.shape, tris.shape)
    out = os.path.join(HERE, "vortex_grid_mesh.npz")
    np.savez_compressed(out, points=pts, triangles=tris)
    print(f"wrote {out}: {pts.shape[0]} points, {tris.shape[0]} triangles")


if __name__ == "__main__":
    main()
