import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests of the HIP path")


@pytest.fixture(scope="session")
def dg():
    return importlib.import_module("airfoil-op2-fem_amd")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def ops3(dg):
    return dg.generate_operators(3)


def _write_triangle_vtk(npz_name, path, title):
    """A fixture mesh (points, triangles) written as a triangle-only
    legacy-VTK file for the library's reader."""
    d = np.load(os.path.join(ROOT, "tests", "golden", npz_name))
    pts, tris = d["points"], d["triangles"]
    with open(path, "w") as f:
        f.write(f"# vtk DataFile Version 2.0\n{title}\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {len(pts)} double\n")
        for p in pts:  # shortest round-trip decimal form: the same doubles back
            f.write(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r}\n")
        f.write(f"CELLS {len(tris)} {4 * len(tris)}\n")
        for t in tris:
            f.write(f"3 {t[0]} {t[1]} {t[2]}\n")
        f.write(f"CELL_TYPES {len(tris)}\n")
        f.write("5\n" * len(tris))
    return str(path)


@pytest.fixture(scope="session")
def naca_vtk(tmp_path_factory):
    """The reference's airfoil mesh (tests/golden/naca0012_mesh.npz, made by
    make_naca0012_mesh.py from naca0012.vtk) written back out as a
    triangle-only legacy-VTK file for the library's reader."""
    return _write_triangle_vtk("naca0012_mesh.npz", tmp_path_factory.mktemp("mesh") / "naca0012.vtk",
                               "naca0012")


@pytest.fixture(scope="session")
def vortex_vtk(tmp_path_factory):
    """The reference's vortex mesh (tests/golden/vortex_grid_mesh.npz, made
    by make_vortex_grid_mesh.py from vortex/grid.vtk), likewise."""
    return _write_triangle_vtk("vortex_grid_mesh.npz", tmp_path_factory.mktemp("mesh") / "grid.vtk", "grid")


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
