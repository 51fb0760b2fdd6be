"""ctypes binding of libdgeuler.so (the C ABI declared in include/dg_euler.h)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libdgeuler.so")


class DGError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _load():
    if not os.path.exists(lib_path):
        raise ImportError(
            f"{lib_path} is not built; run `make -C {_HERE}` (or __graft_entry__.build()). "
            "The DG kernels have no CPU fallback.")
    return ctypes.CDLL(lib_path)


lib = _load()

c_int, c_double, c_void_p, c_char_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
P = ctypes.POINTER
c_uint64 = ctypes.c_uint64

_SIGS = {
    "dg_ctx_create": ([c_int, P(c_void_p)], c_int),
    "dg_ctx_destroy": ([c_void_p], None),
    "dg_last_error": ([c_void_p], c_char_p),
    "dg_order": ([c_void_p, P(c_int), P(c_int), P(c_int)], c_int),
    "dg_set_stream": ([c_void_p, c_void_p], c_int),
    "dg_set_max_blocks": ([c_void_p, c_int], c_int),
    "dg_generate_operators": ([c_int] + [c_void_p] * 10, c_int),
    "dg_set_const": ([c_void_p, c_char_p, c_int, c_void_p], c_int),
    "dg_get_const": ([c_void_p, c_char_p, c_int, c_void_p], c_int),
    "dg_set_freestream": ([c_void_p, c_double, c_double, c_double, c_double, c_double], c_int),
    "dg_decl_mesh": ([c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                      c_void_p, c_void_p], c_int),
    "dg_mesh_check": ([c_int, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                       c_void_p], c_int),
    "dg_init_grid": ([c_void_p] + [c_void_p] * 16, c_int),
    "dg_set_ic": ([c_void_p, c_void_p, c_void_p], c_int),
    "dg_set_ic_vortex": ([c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "dg_neighbour_zero": ([c_void_p, c_void_p], c_int),
    "dg_get_neighbour_q": ([c_void_p] + [c_void_p] * 4, c_int),
    "dg_get_bedge_q": ([c_void_p] + [c_void_p] * 4, c_int),
    "dg_get_bedge_q_exact": ([c_void_p, c_void_p, c_void_p, c_double, c_void_p], c_int),
    "dg_set_exact_state": ([c_void_p, c_double, c_double, c_double, c_double, c_double], c_int),
    "dg_l2_error": ([c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p], c_int),
    "dg_l2_error_mass": ([c_void_p] + [c_void_p] * 7 + [c_double, c_void_p], c_int),
    "dg_mesh_rect": ([c_int, c_int, c_double, c_double, c_double, c_double, c_uint64, c_double, c_int,
                      P(c_void_p)], c_int),
    "dg_euler_rhs": ([c_void_p] + [c_void_p] * 10, c_int),
    "dg_rhs_fused": ([c_void_p] + [c_void_p] * 9, c_int),
    "dg_set_workingQ": ([c_void_p, c_double, c_int, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "dg_update_Q": ([c_void_p, c_double] + [c_void_p] * 5, c_int),
    "dg_calc_dt": ([c_void_p, c_void_p, c_void_p, P(c_double)], c_int),
    "dg_mesh_periodic": ([c_int, c_int, c_double, c_double, c_double, c_double, c_uint64,
                          c_double, P(c_void_p)], c_int),
    "dg_mesh_ogrid": ([c_int, c_int, c_double, c_double, P(c_void_p)], c_int),
    "dg_mesh_read_vtk": ([c_char_p, P(c_void_p)], c_int),
    "dg_mesh_destroy": ([c_void_p], None),
    "dg_mesh_sizes": ([c_void_p, P(c_int), P(c_int), P(c_int), P(c_int)], c_int),
    "dg_mesh_get": ([c_void_p, c_char_p, c_void_p], c_int),
    "dg_mesh_set_bedge_types": ([c_void_p, c_int, c_void_p], c_int),
    "dg_solver_create": ([c_void_p, c_void_p, P(c_void_p)], c_int),
    "dg_solver_destroy": ([c_void_p], None),
    "dg_solver_set_vortex": ([c_void_p, c_double, c_double, c_double, c_double, c_double], c_int),
    "dg_solver_set_q": ([c_void_p, c_void_p], c_int),
    "dg_solver_get_q": ([c_void_p, c_void_p], c_int),
    "dg_solver_dat": ([c_void_p, c_char_p, P(c_void_p)], c_int),
    "dg_solver_run": ([c_void_p, c_int, c_double, c_int], c_int),
    "dg_solver_run_graph": ([c_void_p, c_int, c_double], c_int),
    "dg_solver_stream": ([c_void_p, P(c_void_p)], c_int),
    "dg_solver_create_batch": ([c_void_p, c_int, c_void_p, c_void_p, c_void_p, P(c_void_p)], c_int),
    "dg_solver_ncases": ([c_void_p, P(c_int)], c_int),
    "dg_solver_case_set_q": ([c_void_p, c_int, c_void_p], c_int),
    "dg_solver_case_get_q": ([c_void_p, c_int, c_void_p], c_int),
    "dg_solver_case_time": ([c_void_p, c_int, P(c_double), P(c_double)], c_int),
    "dg_solver_case_forces": ([c_void_p, c_int, c_double, P(c_double), P(c_double), c_void_p], c_int),
    "dg_solver_time": ([c_void_p, P(c_double), P(c_double)], c_int),
    "dg_solver_reset_time": ([c_void_p], c_int),
    "dg_solver_check": ([c_void_p], c_int),
    "dg_solver_forces": ([c_void_p, c_double, P(c_double), P(c_double), c_void_p], c_int),
    "dg_solver_wall_cp": ([c_void_p, P(c_int), c_void_p, c_void_p], c_int),
    "dg_rhs_tangent": ([c_void_p] + [c_void_p] * 16 + [c_double, c_double, c_void_p, c_void_p], c_int),
    "dg_rhs_adjoint": ([c_void_p, c_void_p, c_void_p, c_double] + [c_void_p] * 18, c_int),
    "dg_rhs_adjoint_noalloc": ([c_void_p, c_void_p, c_void_p, c_double] + [c_void_p] * 17, c_int),
    "dg_sens_create": ([c_void_p, c_int, c_int, c_double, c_double, c_int, c_int, c_double, c_double,
                        P(c_void_p)], c_int),
    "dg_sens_destroy": ([c_void_p], None),
    "dg_sens_record": ([c_void_p, c_void_p, c_int, c_double, P(c_double)], c_int),
    "dg_sens_eval": ([c_void_p, c_double, c_void_p, P(c_double)], c_int),
    "dg_sens_set_alpha": ([c_void_p, c_double], c_int),
    "dg_sens_tangent": ([c_void_p, c_double, c_void_p, P(c_double), P(c_double)], c_int),
    "dg_sens_gradient": ([c_void_p, c_int, P(c_double), P(c_double), c_void_p], c_int),
    "dg_sens_timing": ([c_void_p, P(c_double), P(c_double), P(c_double)], c_int),
    "dg_sens_dts": ([c_void_p, c_int, c_void_p], c_int),
    "dg_sens_gradient_breakdown": ([c_void_p, P(c_double), P(c_double), P(c_double)], c_int),
    "dg_solver_profile": ([c_void_p, c_int], c_int),
    "dg_solver_set_launch": ([c_void_p, c_int, c_int, c_int], c_int),
    "dg_solver_set_geometry": ([c_void_p, c_int], c_int),
    "dg_solver_vertex_fields": ([c_void_p, c_int, P(c_int), c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "dg_solver_write_vtk": ([c_void_p, c_int, c_char_p], c_int),
    "dg_solver_launch_config": ([c_void_p, P(c_int), P(c_int), P(c_int)], c_int),
    "dg_solver_autotune": ([c_void_p, c_int, c_double, c_void_p, P(c_int)], c_int),
    "dg_solver_kernel_time": ([c_void_p, c_char_p, P(c_double), P(c_int)], c_int),
}

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res


def check(rc, ctx=None):
    if rc != 0:
        msg = lib.dg_last_error(ctx)
        raise DGError(rc, msg.decode() if msg else "unknown error")
    return rc
