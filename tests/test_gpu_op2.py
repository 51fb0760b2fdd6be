"""The OP2-compatible host interface (include/op2_dg.h): the airfoil driver
built against it, in the source form of src/airfoil.cpp (op_decl_const,
op_par_loop(kernel, "name", ...)) and in the translated form of
src/airfoil_op.cpp (op_decl_const2, op_par_loop_<kernel>), runs the
reference's loop sequence on the GPU and reproduces the oracle's restatement
of that sequence.

Bar: relative L2 per conserved variable <= 1e-12 after 10 steps at the
reference's order (N=4, its literal table size), dt = 1/dt1 every step.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVERS = {"translated": os.path.join(ROOT, "tests", "op2", "airfoil_op2_driver"),
           "source": os.path.join(ROOT, "tests", "op2", "airfoil_seq_form_driver")}


def rel_per_var(a, b):
    a = a.reshape(-1, 4)
    b = b.reshape(-1, 4)
    return [np.linalg.norm(a[:, v] - b[:, v]) / np.linalg.norm(b[:, v]) for v in range(4)]


@pytest.mark.parametrize("form", ["source", "translated"])
def test_op2_driver_matches_oracle(dg, orc, tmp_path, form):
    DRIVER = DRIVERS[form]
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (make -C airfoil-op2-fem_amd)")
    ncirc, nrad, order, iters, alpha = 24, 6, 4, 10, 3.0
    out = tmp_path / "q.bin"
    print(f"op2 driver ({form} form): ogrid {ncirc}x{nrad} N={order} {iters} steps", flush=True)
    r = subprocess.run([DRIVER, "ogrid", str(ncirc), str(nrad), str(order), str(iters), str(alpha), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    for loop in ("get_neighbour_q", "get_bedge_q", "euler_rhs", "set_workingQ", "update_Q", "calc_dt"):
        assert loop in r.stdout  # op_timing_output lists every loop
    mesh = dg.naca0012_ogrid(ncirc, nrad, 1.15, 15.0)
    Np = (order + 1) * (order + 2) // 2
    raw = np.fromfile(out, dtype=np.float64)
    q_gpu, t_gpu = raw[:-2].reshape(mesh.ncells, Np, 4), raw[-2]
    o = orc.Oracle(dg.generate_operators(order), orc.freestream(1.4, 0.4, alpha))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    ph = o.phys
    q0 = np.broadcast_to(np.array([ph["bc_r"], ph["bc_r"] * ph["bc_u"], ph["bc_r"] * ph["bc_v"], ph["bc_e"]]),
                         (mesh.ncells, Np, 4))
    st = o.new_state(case, q0)
    t_ref, _ = o.run(case, st, iters, cfl=1.0)
    assert max(rel_per_var(q_gpu, st["q"])) <= 1e-12
    assert abs(t_gpu - t_ref) <= 1e-12 * t_ref


def test_op2_vortex_driver_matches_oracle(dg, orc, tmp_path, vortex_vtk):
    """vortex/src/vortex.cpp's source form over the same interface: the
    reference's vortex mesh (grid.vtk through the VTK reader; no boundary
    classes declared, so every boundary edge takes the exact vortex state),
    N=4, the vortex at its reference constants advected at speed 1, dt =
    0.1 / dt1 of the initial state kept fixed, 200 steps, l2_error at the
    end.  Against the oracle's fixed-dt run of the same sequence: <= 1e-12
    relative L2 per variable and the same time; the l2_error sums against
    the squared errors of the driver's state from the oracle's exact state,
    each square within 2|e|d + d^2 for an exact-state difference
    d <= 1e-13 (|exact| + 1)."""
    import math
    DRIVER = os.path.join(ROOT, "tests", "op2", "vortex_seq_form_driver")
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (make -C airfoil-op2-fem_amd)")
    order, iters = 4, 200
    out = tmp_path / "q.bin"
    print(f"op2 vortex driver: grid.vtk N={order} {iters} steps", flush=True)
    r = subprocess.run([DRIVER, vortex_vtk, str(order), str(iters), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    for loop in ("set_ic", "neighbour_zero", "get_neighbour_q", "get_bedge_q", "euler_rhs", "set_workingQ",
                 "update_Q", "calc_dt", "l2_error"):
        assert loop in r.stdout  # op_timing_output lists every loop
    mesh = dg.read_vtk(vortex_vtk)
    mesh.set_bedge_types(dg.BC_EXACT)
    Np = (order + 1) * (order + 2) // 2
    raw = np.fromfile(out, dtype=np.float64)
    q_gpu, (t_gpu, dt_gpu), err_gpu = raw[:-6].reshape(mesh.ncells, Np, 4), raw[-6:-4], raw[-4:]
    o = orc.Oracle(dg.generate_operators(order), orc.freestream(1.4, 1.0 / math.sqrt(1.4), 0.0))
    o.set_exact_state(1.0, 5.0, 5.0, 0.0, 0.0)
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    g = case["geo"]
    q0 = o.vortex_state(1.0, 5.0, 5.0, 0.0, 0.0, g["x"], g["y"])
    dt_ref = 0.1 / o.calc_dt_all(case, q0)
    assert abs(dt_gpu - dt_ref) <= 1e-14 * dt_ref, (dt_gpu, dt_ref)
    st = o.new_state(case, q0)
    t_ref = o.run_dt(case, st, iters, dt_gpu)
    errs = rel_per_var(q_gpu, st["q"])
    print(f"op2 vortex driver: rel L2 per variable vs oracle {errs}, l2_error sums {err_gpu}", flush=True)
    assert max(errs) <= 1e-12, errs
    assert abs(t_gpu - t_ref) <= 1e-13 * t_ref
    exact = o.vortex_state(1.0, 5.0, 5.0, 0.0, t_gpu, g["x"], g["y"])
    e = (q_gpu - exact).reshape(-1, 4)
    d = 1e-13 * (np.abs(exact).reshape(-1, 4) + 1.0)
    bound = (2 * np.abs(e) * d + d * d).sum(axis=0) + 1e-12 * (e * e).sum(axis=0)
    assert np.all(np.abs(err_gpu - (e * e).sum(axis=0)) <= bound), (err_gpu, (e * e).sum(axis=0), bound)
