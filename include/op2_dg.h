/* OP2-compatible host interface for the DG Euler loops of the airfoil
 * driver, on top of the dg_euler.h C ABI (C++ only).
 *
 * Covers the subset of OP2 that the airfoil driver uses, in both of its
 * forms:
 *   - the source form of src/airfoil.cpp (op_seq.h style):
 *     op_decl_const(dim, type, &var) and op_par_loop(kernel, "name", set,
 *     args...) -- the constant's name is taken from the pointer expression
 *     and the loop from its name string, as OP2's translator does; the
 *     kernel identifier is not evaluated, so the seq kernel headers
 *     (src/euler_rhs.h etc., which need CBLAS) are not needed;
 *   - the translated form of src/airfoil_op.cpp: op_decl_const2 and the
 *     nine op_par_loop_* host stubs of src/airfoil_op.cpp:17-98;
 * plus op_init, op_decl_set, op_decl_map, op_decl_dat, op_arg_dat,
 * op_arg_gbl, op_timers, op_get_size, op_fetch_data, op_timing_output and
 * op_exit.  Every loop runs on the MI355X kernels.
 *
 * Semantics follow OP2's device backends: op_decl_dat copies the user's
 * host array to a device mirror; par_loops run on the mirrors and return
 * when the device work is done (as the generated CUDA stubs do); OP_MAX
 * global reductions land in host memory on return; op_fetch_data copies a
 * mirror back.  Errors (an unknown constant name, arguments that do not
 * match the loop) throw std::runtime_error, where OP2 prints and exits.
 *
 * Mesh connectivity for the owner-computes face map is taken from the
 * declared maps: edges -> cells (dimension 2) with its int dat of dimension
 * 2 (edgeNum), and boundary edges -> cells (dimension 1) with the int dats
 * named "bedgeNum" and "bedge_type" as src/airfoil.cpp declares them.  The
 * maps and dats passed to get_neighbour_q / get_bedge_q are checked against
 * that choice on every call and replace it when they differ.
 */
#ifndef OP2_DG_H
#define OP2_DG_H

#ifndef __cplusplus
#error "op2_dg.h is a C++ interface"
#endif

#include <initializer_list>

struct op_set_core;
struct op_map_core;
struct op_dat_core;
typedef op_set_core *op_set;
typedef op_map_core *op_map;
typedef op_dat_core *op_dat;

enum op_access { OP_READ = 0, OP_WRITE = 1, OP_RW = 2, OP_INC = 3, OP_MIN = 4, OP_MAX = 5 };

/* One argument of a par_loop: a dataset (directly, or through a map and
 * map index) or a global (host) value. */
struct op_arg {
  op_dat dat;
  op_map map;
  int idx;
  int dim;
  const char *type;
  op_access acc;
  void *gbl;  /* host pointer of an op_arg_gbl argument, else null */
};

#define OP_ID ((op_map)0)

void op_init(int argc, char **argv, int diags);
void op_exit();
op_set op_decl_set(int size, char const *name);
op_map op_decl_map(op_set from, op_set to, int dim, int *imap, char const *name);
op_dat op_decl_dat_char(op_set set, int dim, char const *type, int type_size, char *data,
                        char const *name);
void op_decl_const_char(int dim, char const *type, int type_size, char *data, char const *name);
op_arg op_arg_dat(op_dat dat, int idx, op_map map, int dim, char const *type, op_access acc);
op_arg op_arg_gbl_char(char *data, int dim, const char *type, int type_size, op_access acc);
void op_timers(double *cpu, double *wall);
int op_get_size(op_set set);
void op_fetch_data_char(op_dat dat, char *dest);
void op_timing_output();

template <class T>
op_dat op_decl_dat(op_set set, int dim, char const *type, T *data, char const *name) {
  return op_decl_dat_char(set, dim, type, (int)sizeof(T), (char *)data, name);
}
template <class T>
void op_decl_const2(char const *name, int dim, char const *type, T *data) {
  op_decl_const_char(dim, type, (int)sizeof(T), (char *)data, name);
}
template <class T>
op_arg op_arg_gbl(T *data, int dim, char const *type, op_access acc) {
  return op_arg_gbl_char((char *)data, dim, type, (int)sizeof(T), acc);
}
template <class T>
void op_fetch_data(op_dat dat, T *dest) {
  op_fetch_data_char(dat, (char *)dest);
}

/* The host stubs of src/airfoil_op.cpp:17-98 */
void op_par_loop_init_grid(char const *, op_set, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg,
                           op_arg, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg,
                           op_arg, op_arg, op_arg);
void op_par_loop_set_ic(char const *, op_set, op_arg, op_arg);
void op_par_loop_neighbour_zero(char const *, op_set, op_arg);
void op_par_loop_calc_dt(char const *, op_set, op_arg, op_arg, op_arg);
void op_par_loop_get_neighbour_q(char const *, op_set, op_arg, op_arg, op_arg, op_arg, op_arg,
                                 op_arg, op_arg, op_arg, op_arg);
void op_par_loop_get_bedge_q(char const *, op_set, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg);
void op_par_loop_euler_rhs(char const *, op_set, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg,
                           op_arg, op_arg, op_arg, op_arg);
void op_par_loop_set_workingQ(char const *, op_set, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg);
void op_par_loop_update_Q(char const *, op_set, op_arg, op_arg, op_arg, op_arg, op_arg, op_arg);

/* The source form of src/airfoil.cpp:183-334.  op_decl_const(dim, type,
 * ptr): the constant's name is the pointer expression without its '&'
 * ("&gam" -> gam, "Drw" -> Drw), as OP2's translator turns it into
 * op_decl_const2("gam", ...); an expression that is not a plain name throws.
 * op_par_loop(kernel, "name", set, args...): the loop named "name" (one of
 * the nine above, with the same arguments), or one of the vortex case's
 * forms (vortex/src/vortex.cpp:195-328), told apart by argument count:
 * set_ic(t, x, y, q, workingQ); get_neighbour_q with the edge's two node
 * coordinates first (11 arguments); get_bedge_q(bedgeNum, two node
 * coordinates, nodeX, nodeY, x, y, t, exteriorQ), the exact vortex state on
 * every boundary edge; l2_error(x, y, q, t, four OP_INC globals, error).
 * The vortex is the one of the constants vortex_x0, vortex_y0, vortex_beta
 * advected at the free-stream speed 1 of vortex/src/set_ic.h:8; a mesh
 * whose boundary edges have no int dat named bedge_type gives every one of
 * them the exact state. */
void op_decl_const_expr(int dim, char const *type, int type_size, char *data, char const *expr);
template <class T>
void op_decl_const_named(int dim, char const *type, T *data, char const *expr) {
  op_decl_const_expr(dim, type, (int)sizeof(T), (char *)data, expr);
}
void op_par_loop_by_name(char const *name, op_set set, std::initializer_list<op_arg> args);
#define op_decl_const(dim, type, ptr) op_decl_const_named((dim), (type), (ptr), #ptr)
#define op_par_loop(kernel, name, set, ...) op_par_loop_by_name((name), (set), {__VA_ARGS__})

#endif /* OP2_DG_H */
