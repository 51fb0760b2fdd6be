// Host-only checks of include/op2_dg.h's source form (no device work): the
// name op_decl_const takes from its pointer expression, and the errors of
// op_par_loop for an unknown loop name or a wrong argument count (either
// form, the airfoil's or the vortex case's), which throw before any loop
// runs.  Prints "source form checks ok".
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/op2_dg.h"

static double gam = 1.4, tab[3] = {1.0, 2.0, 3.0};
static int failures = 0;

template <class F>
static void expect_throw(const char *what, const char *words, F &&f) {
  try {
    f();
    std::fprintf(stderr, "FAILED: %s did not throw\n", what);
    failures++;
  } catch (const std::runtime_error &e) {
    if (std::string(e.what()).find(words) == std::string::npos) {
      std::fprintf(stderr, "FAILED: %s threw '%s'\n", what, e.what());
      failures++;
    }
  }
}

int main(int argc, char **argv) {
  op_init(argc, argv, 0);
  op_set cells = op_decl_set(4, "cells");
  // global arguments only: op_decl_dat would allocate a device mirror
  const op_arg g = op_arg_gbl(&gam, 1, "double", OP_READ);
  try {
    op_decl_const(1, "double", &gam);   // name "gam"
    op_decl_const(3, "double", tab);    // name "tab"
    op_decl_const(1, "double", (&gam)); // parentheses stripped
  } catch (const std::runtime_error &e) {
    std::fprintf(stderr, "FAILED: op_decl_const of a plain name threw '%s'\n", e.what());
    failures++;
  }
  expect_throw("op_decl_const(&tab[1])", "cannot take a constant name", [] { op_decl_const(1, "double", &tab[1]); });
  expect_throw("op_decl_const(tab + 1)", "cannot take a constant name", [] { op_decl_const(1, "double", tab + 1); });
  expect_throw("unknown loop", "no loop named 'flux_limiter'", [&] {
    op_par_loop(flux_limiter, "flux_limiter", cells, g);
  });
  expect_throw("set_ic with one argument", "1 arguments, the loop takes 2", [&] {
    op_par_loop(set_ic, "set_ic", cells, g);
  });
  expect_throw("euler_rhs with two arguments", "the loop takes 10", [&] {
    op_par_loop(euler_rhs, "euler_rhs", cells, g, g);
  });
  expect_throw("get_bedge_q with three arguments", "the loop takes 6 (9 in the vortex case's form)", [&] {
    op_par_loop(get_bedge_q, "get_bedge_q", cells, g, g, g);
  });
  expect_throw("l2_error with one argument", "1 arguments, the loop takes 9", [&] {
    op_par_loop(l2_error, "l2_error", cells, g);
  });
  op_exit();
  std::printf("%s\n", failures ? "source form checks FAILED" : "source form checks ok");
  return failures ? 1 : 0;
}
