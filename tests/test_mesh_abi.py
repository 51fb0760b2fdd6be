"""Host-side checks: mesh generators and connectivity against the oracle's
restatements, and the C ABI exports (no device calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
VTK = "/root/reference/naca0012.vtk"


def _areas(cxy):
    return 0.5 * ((cxy[:, 1, 0] - cxy[:, 0, 0]) * (cxy[:, 2, 1] - cxy[:, 0, 1])
                  - (cxy[:, 2, 0] - cxy[:, 0, 0]) * (cxy[:, 1, 1] - cxy[:, 0, 1]))


def test_periodic_grid_matches_restatement(dg):
    m = dg.periodic_grid(9, 7, seed=11, jitter=0.1)
    cells, cxy = oracle.periodic_grid_np(9, 7, seed=11, jitter=0.1)
    assert np.array_equal(m.cells, cells)
    assert np.array_equal(m.cell_xy, cxy)
    assert m.ncells == 2 * 9 * 7 and m.nedges == 3 * 9 * 7 and m.nbedges == 0
    assert np.all(_areas(m.cell_xy) > 0)


@pytest.mark.parametrize("kind", ["periodic", "ogrid", "vtk"])
def test_edges_follow_vtk2cgns_contract(dg, kind):
    if kind == "periodic":
        m = dg.periodic_grid(6, 5, seed=3)
    elif kind == "ogrid":
        m = dg.naca0012_ogrid(24, 6)
    else:
        if not os.path.exists(VTK):
            pytest.skip("reference mesh file not present")
        m = dg.read_vtk(VTK)
    e = oracle.build_edges(m.cells)
    for k in ("edge2cells", "edgeNum", "edge2nodes", "bedge2cells", "bedgeNum", "bedge2nodes"):
        assert np.array_equal(getattr(m, k), e[k]), k
    # every (cell, local edge) is covered exactly once
    slots = np.zeros((m.ncells, 3), dtype=int)
    np.add.at(slots, (m.edge2cells[:, 0], m.edgeNum[:, 0]), 1)
    np.add.at(slots, (m.edge2cells[:, 1], m.edgeNum[:, 1]), 1)
    np.add.at(slots, (m.bedge2cells, m.bedgeNum), 1)
    assert np.all(slots == 1)
    assert np.all(_areas(m.cell_xy) > 0)


def test_ogrid_boundaries(dg):
    m = dg.naca0012_ogrid(32, 8, stretch=1.08, rfar=20.0)
    assert m.ncells == 2 * 32 * 8
    assert np.array_equal(np.bincount(m.bedge_type, minlength=4), [0, 0, 32, 32])
    # wall nodes lie on the NACA0012 profile of naca0012.geo:14
    wall = m.bedge2nodes[m.bedge_type == 2].ravel()
    x, y = m.coords[wall, 0], m.coords[wall, 1]
    t = 0.594689181 * (0.298222773 * np.sqrt(x) - 0.127125232 * x - 0.357907906 * x ** 2
                       + 0.291984971 * x ** 3 - 0.105174696 * x ** 4)
    # the polynomial leaves a 5e-8 gap at x=1; the trailing edge point is closed
    assert np.allclose(np.abs(y), t, atol=1e-7)
    far = m.bedge2nodes[m.bedge_type == 3].ravel()
    assert np.allclose(np.hypot(m.coords[far, 0] - 0.5, m.coords[far, 1]), 20.0)


def test_reference_vtk_mesh_classification(dg):
    if not os.path.exists(VTK):
        pytest.skip("reference mesh file not present")
    m = dg.read_vtk(VTK)
    # counts from the file's CELL_TYPES; classes from src/load_mesh.cpp:147-172
    assert (m.ncells, m.nnodes) == (2450, 1395)
    assert m.nedges + m.nbedges == (3 * 2450 + m.nbedges) // 2
    t = m.bedge_type
    c = m.coords
    for b in range(m.nbedges):
        (x1, y1), (x2, y2) = c[m.bedge2nodes[b, 0]], c[m.bedge2nodes[b, 1]]
        inner = all(-1 < v < 2 for v in (x1, x2)) and all(-1 < v < 1 for v in (y1, y2))
        expect = 2 if inner or y1 == y2 else (0 if x1 < 0 else 1)
        assert t[b] == expect


def test_mesh_fixture_reads_like_the_reference_file(dg, naca_vtk):
    """The committed airfoil-mesh fixture, read through the VTK reader, is the
    mesh the reference file gives (same points, cells, edges and boundary
    classes); checked against the file itself where the reference is
    mounted."""
    m = dg.read_vtk(naca_vtk)
    assert (m.ncells, m.nnodes) == (2450, 1395)
    assert set(np.unique(m.bedge_type)) <= {0, 1, 2}
    assert np.all(_areas(m.cell_xy.reshape(-1, 3, 2)) > 0)
    if os.path.exists(VTK):
        r = dg.read_vtk(VTK)
        for k in ("cells", "cell_xy", "edge2cells", "edgeNum", "bedge2cells", "bedgeNum", "bedge_type"):
            assert np.array_equal(getattr(m, k), getattr(r, k)), k


def test_mesh_errors_are_reported(dg):
    with pytest.raises(dg.DGError):
        dg.periodic_grid(2, 2)
    with pytest.raises(dg.DGError):
        dg.read_vtk("/nonexistent/mesh.vtk")


def test_abi_exports_every_declared_symbol(dg):
    hdr = open(os.path.join(ROOT, "include", "dg_euler.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(dg_\w+)\s*\(", hdr, re.M)))
    assert len(names) > 30
    lib = ctypes.CDLL(dg.lib_path)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _header_prototypes():
    """{name: parameter count} of every function declared in dg_euler.h."""
    hdr = open(os.path.join(ROOT, "include", "dg_euler.h")).read()
    hdr = re.sub(r"/\*.*?\*/", " ", hdr, flags=re.S)
    protos = {}
    for m in re.finditer(r"^\s*(?:int|void|const char \*)\s*(dg_\w+)\s*\(([^;]*?)\)\s*;", hdr, re.M | re.S):
        args = m.group(2).strip()
        protos[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return protos


def test_ctypes_signatures_match_the_header(dg):
    """Every ctypes binding declares as many arguments as the C prototype it
    binds (a short list would let ctypes convert the tail by default rules),
    and every header function has a binding."""
    from importlib import import_module
    sigs = import_module("airfoil-op2-fem_amd._lib")._SIGS
    protos = _header_prototypes()
    assert len(protos) > 60
    wrong = {n: (len(sigs[n][0]), protos[n]) for n in sigs if n in protos and len(sigs[n][0]) != protos[n]}
    assert not wrong, wrong
    assert set(sigs) <= set(protos), set(sigs) - set(protos)
    unbound = sorted(set(protos) - set(sigs))
    assert not unbound, unbound


def test_op2_interface_exports_every_declared_function(dg):
    """Every non-template function of include/op2_dg.h (C++ linkage) is
    defined in the library, and the OP2 driver built against it links."""
    import shutil
    import subprocess
    hdr = open(os.path.join(ROOT, "include", "op2_dg.h")).read()
    names = set(re.findall(r"^(?:void|op_set|op_map|op_dat|op_arg|int)\s+(op_\w+)\s*\(", hdr, re.M))
    templates = set(re.findall(r"template <class T>\s*\n\s*\w+\s+(op_\w+)\s*\(", hdr))
    assert templates == {"op_decl_dat", "op_decl_const2", "op_arg_gbl", "op_fetch_data"}
    names = sorted(names - templates)
    assert "op_par_loop_euler_rhs" in names and len(names) >= 20
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", "-C", dg.lib_path], capture_output=True, text=True).stdout
    defined = set(re.findall(r"\s(op_\w+)\(", out))
    missing = [n for n in names if n not in defined]
    assert not missing, missing
    assert os.path.exists(os.path.join(ROOT, "tests", "op2", "airfoil_op2_driver"))


def test_bad_order_rejected_without_device(dg):
    with pytest.raises(dg.DGError):
        dg.Context(order=0)
    with pytest.raises(dg.DGError):
        dg.Context(order=7)
