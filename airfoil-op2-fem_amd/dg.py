"""Host-side mirror of the reference's DG Euler operator interface.

``Context`` plays the part of OP2's op_init/op_decl_const/op_decl_map
(src/airfoil.cpp:67-201) and exposes one method per op_par_loop of the hot
path, operating on torch CUDA tensors in the op_dat layouts.  ``Mesh``
replaces the CGNS loader (src/load_mesh.cpp) with generated or VTK meshes.
``Solver`` is the native SSPRK3 driver of src/airfoil.cpp:204-334.
"""
import ctypes
import os

import numpy as np

from ._lib import DGError, P, c_double, c_int, c_void_p, check, lib

BC_INFLOW, BC_OUTFLOW, BC_WALL, BC_FARFIELD, BC_EXACT = 0, 1, 2, 3, 4
MODE_FUSED, MODE_REFERENCE = 0, 1
GEOM_NODE, GEOM_CELL = 0, 1


def _np_ptr(a):
    return a.ctypes.data_as(c_void_p)


def generate_operators(order):
    """Order-N DG operator tables (host only): dict of numpy arrays named as
    the reference's constants (r, s, Dr, Ds, Drw, Dsw, LIFT, FMASK) plus the
    reference-element mass matrix and the 1D face mass matrix."""
    n = int(order)
    np_, nfp = (n + 1) * (n + 2) // 2, n + 1
    out = {"r": np.zeros(np_), "s": np.zeros(np_), "Dr": np.zeros((np_, np_)),
           "Ds": np.zeros((np_, np_)), "Drw": np.zeros((np_, np_)), "Dsw": np.zeros((np_, np_)),
           "LIFT": np.zeros((np_, 3 * nfp)), "FMASK": np.zeros(3 * nfp, dtype=np.int32),
           "mass": np.zeros((np_, np_)), "face_mass": np.zeros((nfp, nfp))}
    check(lib.dg_generate_operators(
        n, *[_np_ptr(out[k]) for k in ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK",
                                        "mass", "face_mass")]))
    out.update(N=n, Np=np_, Nfp=nfp)
    return out


# --------------------------------------------------------------------------
# meshes
# --------------------------------------------------------------------------

class Mesh:
    """A triangle mesh with the tools/vtk2CGNS.cpp connectivity contract."""

    def __init__(self, handle):
        self._h = c_void_p(handle)
        nn, nc, ne, nb = c_int(), c_int(), c_int(), c_int()
        check(lib.dg_mesh_sizes(self._h, nn, nc, ne, nb))
        self.nnodes, self.ncells, self.nedges, self.nbedges = nn.value, nc.value, ne.value, nb.value
        self._cache = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.dg_mesh_destroy(h)
            self._h = None

    _SHAPES = {"coords": ("nnodes", (2,), np.float64), "cells": ("ncells", (3,), np.int32),
               "cell_xy": ("ncells", (3, 2), np.float64), "edge2cells": ("nedges", (2,), np.int32),
               "edgeNum": ("nedges", (2,), np.int32), "edge2nodes": ("nedges", (2,), np.int32),
               "bedge2cells": ("nbedges", (), np.int32), "bedgeNum": ("nbedges", (), np.int32),
               "bedge_type": ("nbedges", (), np.int32), "bedge2nodes": ("nbedges", (2,), np.int32)}

    def get(self, field):
        if field not in self._cache:
            cnt, tail, dt = self._SHAPES[field]
            a = np.zeros((getattr(self, cnt),) + tail, dtype=dt)
            if a.size:
                check(lib.dg_mesh_get(self._h, field.encode(), _np_ptr(a)))
            self._cache[field] = a
        return self._cache[field]

    def set_bedge_types(self, types):
        """Boundary class of every boundary edge (DG_BC_* values, one per
        boundary edge), e.g. all BC_EXACT for the vortex case."""
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(types, dtype=np.int32), (self.nbedges,)))
        check(lib.dg_mesh_set_bedge_types(self._h, self.nbedges, _np_ptr(t) if t.size else None))
        self._cache.pop("bedge_type", None)

    def __getattr__(self, name):
        if name in Mesh._SHAPES:
            return self.get(name)
        raise AttributeError(name)


def periodic_grid(nx, ny, x0=0.0, y0=-5.0, lx=10.0, ly=10.0, seed=0, jitter=0.1):
    """Seeded jittered periodic grid of nx*ny quads split into 2*nx*ny CCW triangles."""
    h = c_void_p()
    check(lib.dg_mesh_periodic(int(nx), int(ny), float(x0), float(y0), float(lx), float(ly),
                               int(seed), float(jitter), ctypes.byref(h)))
    return Mesh(h.value)


def rect_grid(nx, ny, x0=0.0, y0=-5.0, lx=10.0, ly=10.0, seed=0, jitter=0.1, bc_type=BC_EXACT):
    """Seeded jittered nx*ny quad grid (not periodic), 2*nx*ny CCW triangles,
    every boundary edge of type bc_type (default: exact vortex state)."""
    h = c_void_p()
    check(lib.dg_mesh_rect(int(nx), int(ny), float(x0), float(y0), float(lx), float(ly), int(seed),
                           float(jitter), int(bc_type), ctypes.byref(h)))
    return Mesh(h.value)


def naca0012_ogrid(ncirc, nrad, stretch=1.08, rfar=20.0):
    """O-grid around the NACA0012 profile: 2*ncirc*nrad CCW triangles."""
    h = c_void_p()
    check(lib.dg_mesh_ogrid(int(ncirc), int(nrad), float(stretch), float(rfar), ctypes.byref(h)))
    return Mesh(h.value)


def mesh_check(cells, edge2cells, edgeNum, bedge2cells, bedgeNum, bedge_type):
    """dg_mesh_check on host arrays: raises DGError (DG_ERR_MESH) with the
    reason when the connectivity is not a valid counter-clockwise mesh
    (cells may be None to skip the orientation check)."""
    def arr(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.int32)
    c, e2c, en, b2c, bn, bt = (arr(a) for a in (cells, edge2cells, edgeNum, bedge2cells, bedgeNum, bedge_type))
    ncells = len(c) if c is not None else int(max(e2c.max() if e2c.size else -1, b2c.max() if b2c.size else -1)) + 1
    check(lib.dg_mesh_check(int(ncells), _np_ptr(c) if c is not None else None, int(e2c.size // 2),
                            _np_ptr(e2c), _np_ptr(en), int(b2c.size), _np_ptr(b2c), _np_ptr(bn), _np_ptr(bt)))


def sweep_cases(ncases=64, seed=3):
    """The batched sweep's seeded uneven O-grid cases: (ncirc, nrad, mach,
    alpha_deg) with 2 * ncirc * nrad in about [8k, 130k] cells, alpha ~
    U[-4, 8] deg, Mach ~ U[0.3, 0.6] (numpy's default_rng(seed))."""
    import math
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ncases):
        cells = rng.uniform(8400, 128000)
        ratio = rng.uniform(1.5, 3.0)
        nrad = max(4, int(round(math.sqrt(cells / (2.0 * ratio)))))
        ncirc = max(16, 2 * int(round(cells / (4.0 * nrad))))
        out.append((ncirc, nrad, float(rng.uniform(0.3, 0.6)), float(rng.uniform(-4.0, 8.0))))
    return out


def read_vtk(path):
    """Legacy-VTK triangle mesh, boundary edges classified as src/load_mesh.cpp:147-172."""
    h = c_void_p()
    check(lib.dg_mesh_read_vtk(str(path).encode(), ctypes.byref(h)))
    return Mesh(h.value)


# --------------------------------------------------------------------------
# context: constants, maps and the per-loop kernels
# --------------------------------------------------------------------------

def _dev(t, name, numel=None):
    """Pointer of a contiguous float64 CUDA tensor (checked)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: expected a device tensor, got {t.device}")
    if t.dtype != torch.float64:
        raise TypeError(f"{name}: expected float64, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: expected {numel} elements, got {t.numel()}")
    return c_void_p(t.data_ptr())


class Context:
    """op_init + op_decl_const + op_decl_map, and the op_par_loop drop-ins."""

    _DBL1 = ("gam", "bc_mach", "bc_alpha", "bc_p", "bc_r", "bc_u", "bc_v", "bc_e", "vortex_x0",
             "vortex_y0", "vortex_beta")

    def __init__(self, order=3, stream=None):
        h = c_void_p()
        check(lib.dg_ctx_create(int(order), ctypes.byref(h)))
        self._h = h
        n, np_, nfp = c_int(), c_int(), c_int()
        check(lib.dg_order(self._h, n, np_, nfp), self._h)
        self.N, self.Np, self.Nfp = n.value, np_.value, nfp.value
        self.NF = 3 * self.Nfp
        self.ncells = None
        if stream is not None:
            self.set_stream(stream)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.dg_ctx_destroy(h)
            self._h = None

    def _chk(self, rc):
        return check(rc, self._h)

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream or a raw hipStream_t integer."""
        ptr = getattr(stream, "cuda_stream", stream)
        self._chk(lib.dg_set_stream(self._h, c_void_p(int(ptr) if ptr else 0)))

    def set_max_blocks(self, max_blocks=0):
        """Cap on the stage kernels' persistent grid (0: occupancy-limited)."""
        self._chk(lib.dg_set_max_blocks(self._h, int(max_blocks)))

    def _dim(self, name):
        if name in self._DBL1:
            return 1
        return {"ones": self.Np, "r": self.Np, "s": self.Np, "Dr": self.Np ** 2,
                "Ds": self.Np ** 2, "Drw": self.Np ** 2, "Dsw": self.Np ** 2,
                "LIFT": self.Np * self.NF, "FMASK": self.NF}.get(name, 1)

    def set_const(self, name, value, dim=None):
        """op_decl_const(dim, type, value) under the reference's constant name."""
        isint = name == "FMASK"
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(value, dtype=np.int32 if isint else np.float64)).ravel())
        d = int(dim) if dim is not None else a.size
        if a.size < d:
            raise ValueError(f"{name}: {a.size} values for dim {d}")
        self._chk(lib.dg_set_const(self._h, name.encode(), d, _np_ptr(a)))

    def get_const(self, name):
        d = self._dim(name)
        a = np.zeros(d, dtype=np.int32 if name == "FMASK" else np.float64)
        self._chk(lib.dg_get_const(self._h, name.encode(), d, _np_ptr(a)))
        return a[0] if d == 1 and name in self._DBL1 else a

    def set_freestream(self, gam=1.4, mach=0.4, alpha=0.0, p=1.0, rho=1.0):
        self._chk(lib.dg_set_freestream(self._h, gam, mach, alpha, p, rho))

    def decl_mesh(self, mesh=None, ncells=None, cells=None, edge2cells=None, edgeNum=None,
                  bedge2cells=None, bedgeNum=None, bedge_type=None):
        """Declare sets and maps, from a Mesh or from raw int32 arrays."""
        if mesh is not None:
            ncells, cells = mesh.ncells, mesh.cells
            edge2cells, edgeNum = mesh.edge2cells, mesh.edgeNum
            bedge2cells, bedgeNum, bedge_type = mesh.bedge2cells, mesh.bedgeNum, mesh.bedge_type

        def arr(a):
            return None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.int32))
        cells, e2c, en = arr(cells), arr(edge2cells), arr(edgeNum)
        b2c, bn, bt = arr(bedge2cells), arr(bedgeNum), arr(bedge_type)
        ne = 0 if e2c is None else e2c.size // 2
        nb = 0 if b2c is None else b2c.size
        ptr = lambda a: None if a is None or a.size == 0 else _np_ptr(a)  # noqa: E731
        self._chk(lib.dg_decl_mesh(self._h, int(ncells), ptr(cells), ne, ptr(e2c), ptr(en), nb,
                                   ptr(b2c), ptr(bn), ptr(bt)))
        self.ncells = int(ncells)

    # ---- shapes of the op_dats ----
    def state_numel(self):
        return self.ncells * self.Np * 4

    # ---- op_par_loop drop-ins (device tensors) ----
    def init_grid(self, cell_xy, d):
        """init_grid over cells; d is a dict of output tensors named as the op_dats."""
        K, Np, NF = self.ncells, self.Np, self.NF
        names = [("nodeX", 3), ("nodeY", 3), ("x", Np), ("y", Np), ("xr", Np), ("yr", Np),
                 ("xs", Np), ("ys", Np), ("rx", Np), ("ry", Np), ("sx", Np), ("sy", Np),
                 ("nx", NF), ("ny", NF), ("fscale", NF)]
        self._chk(lib.dg_init_grid(self._h, _dev(cell_xy, "cell_xy", K * 6),
                                   *[_dev(d[n], n, K * s) for n, s in names]))

    def set_ic(self, q, workingQ):
        n = self.state_numel()
        self._chk(lib.dg_set_ic(self._h, _dev(q, "q", n), _dev(workingQ, "workingQ", n)))

    def set_ic_vortex(self, t, x, y, q, workingQ):
        """vortex/src/set_ic.h over cells at time t (the vortex of set_exact_state)."""
        K, Np, n = self.ncells, self.Np, self.state_numel()
        self._chk(lib.dg_set_ic_vortex(self._h, float(t), _dev(x, "x", K * Np), _dev(y, "y", K * Np),
                                       _dev(q, "q", n), _dev(workingQ, "workingQ", n)))

    def neighbour_zero(self, exteriorQ):
        self._chk(lib.dg_neighbour_zero(self._h, _dev(exteriorQ, "exteriorQ", self.ncells * self.NF * 4)))

    def get_neighbour_q(self, nodeX, nodeY, workingQ, exteriorQ):
        K = self.ncells
        self._chk(lib.dg_get_neighbour_q(self._h, _dev(nodeX, "nodeX", K * 3), _dev(nodeY, "nodeY", K * 3),
                                         _dev(workingQ, "workingQ", self.state_numel()),
                                         _dev(exteriorQ, "exteriorQ", K * self.NF * 4)))

    def get_bedge_q(self, nx, ny, workingQ, exteriorQ):
        K, NF = self.ncells, self.NF
        self._chk(lib.dg_get_bedge_q(self._h, _dev(nx, "nx", K * NF), _dev(ny, "ny", K * NF),
                                     _dev(workingQ, "workingQ", self.state_numel()),
                                     _dev(exteriorQ, "exteriorQ", K * NF * 4)))

    def get_bedge_q_exact(self, x, y, t, exteriorQ):
        """vortex/src/get_bedge_q.h over every boundary edge at time t."""
        K, Np, NF = self.ncells, self.Np, self.NF
        self._chk(lib.dg_get_bedge_q_exact(self._h, _dev(x, "x", K * Np), _dev(y, "y", K * Np), float(t),
                                           _dev(exteriorQ, "exteriorQ", K * NF * 4)))

    def set_exact_state(self, uinf, beta=5.0, x0=5.0, y0=0.0, period_x=0.0):
        """Vortex of the exact-state boundaries and get_bedge_q_exact."""
        self._chk(lib.dg_set_exact_state(self._h, float(uinf), float(beta), float(x0), float(y0),
                                         float(period_x)))

    def l2_error(self, x, y, q, t, error=None, err=None):
        """vortex/src/l2_error.h over cells (vortex/src/vortex.cpp:314-328):
        error (optional) = (q - exact)^2 per node; returns err + the sums of
        those squares over every node, per variable (r, ru, rv, ener)."""
        K, Np = self.ncells, self.Np
        out = np.zeros(4) if err is None else np.array(err, dtype=np.float64).reshape(4)
        ep = None if error is None else _dev(error, "error", K * Np * 4)
        self._chk(lib.dg_l2_error(self._h, _dev(x, "x", K * Np), _dev(y, "y", K * Np),
                                  _dev(q, "q", self.state_numel()), float(t), ep,
                                  out.ctypes.data_as(c_void_p)))
        return out

    def l2_error_mass(self, x, y, xr, yr, xs, ys, q, t):
        """Mass-weighted L2 norm of q - exact per variable:
        sqrt(sum_cells J e^T M e), J at each cell's node 0."""
        K, Np = self.ncells, self.Np
        out = np.zeros(4)
        self._chk(lib.dg_l2_error_mass(self._h, _dev(x, "x", K * Np), _dev(y, "y", K * Np),
                                       _dev(xr, "xr", K * Np), _dev(yr, "yr", K * Np),
                                       _dev(xs, "xs", K * Np), _dev(ys, "ys", K * Np),
                                       _dev(q, "q", self.state_numel()), float(t),
                                       out.ctypes.data_as(c_void_p)))
        return out

    def euler_rhs(self, workingQ, exteriorQ, rx, ry, sx, sy, fscale, nx, ny, qRHS):
        K, Np, NF, n = self.ncells, self.Np, self.NF, self.state_numel()
        self._chk(lib.dg_euler_rhs(self._h, _dev(workingQ, "workingQ", n),
                                   _dev(exteriorQ, "exteriorQ", K * NF * 4),
                                   _dev(rx, "rx", K * Np), _dev(ry, "ry", K * Np),
                                   _dev(sx, "sx", K * Np), _dev(sy, "sy", K * Np),
                                   _dev(fscale, "fscale", K * NF), _dev(nx, "nx", K * NF),
                                   _dev(ny, "ny", K * NF), _dev(qRHS, "qRHS", n)))

    def rhs_fused(self, workingQ, rx, ry, sx, sy, fscale, nx, ny, qRHS):
        K, Np, NF, n = self.ncells, self.Np, self.NF, self.state_numel()
        self._chk(lib.dg_rhs_fused(self._h, _dev(workingQ, "workingQ", n),
                                   _dev(rx, "rx", K * Np), _dev(ry, "ry", K * Np),
                                   _dev(sx, "sx", K * Np), _dev(sy, "sy", K * Np),
                                   _dev(fscale, "fscale", K * NF), _dev(nx, "nx", K * NF),
                                   _dev(ny, "ny", K * NF), _dev(qRHS, "qRHS", n)))

    def rhs_tangent(self, w, wd, geo, geod=None, bcd=(0.0, 0.0), scratch=None, out=None):
        """L'(w) wd + dL/dgeo geod + dL/d(bc_u, bc_v) bcd.  geo/geod: dicts of
        device tensors rx, ry, sx, sy, fscale, nx, ny (geod entries optional)."""
        import torch
        K, Np, NF, n = self.ncells, self.Np, self.NF, self.state_numel()
        scratch = scratch if scratch is not None else torch.empty(K * NF * 4, dtype=torch.float64, device=w.device)
        out = out if out is not None else torch.empty(n, dtype=torch.float64, device=w.device)
        geod = geod or {}
        sizes = dict(rx=K * Np, ry=K * Np, sx=K * Np, sy=K * Np, fscale=K * NF, nx=K * NF, ny=K * NF)
        g = [_dev(geo[k], k, sizes[k]) for k in ("rx", "ry", "sx", "sy", "fscale", "nx", "ny")]
        gd = [(_dev(geod[k], k + "d", sizes[k]) if k in geod else None)
              for k in ("rx", "ry", "sx", "sy", "fscale", "nx", "ny")]
        self._chk(lib.dg_rhs_tangent(self._h, _dev(w, "w", n), _dev(wd, "wd", n), *g, *gd,
                                     float(bcd[0]), float(bcd[1]), _dev(scratch, "scratch", K * NF * 4),
                                     _dev(out, "out", n)))
        return out

    def rhs_adjoint(self, w, mu, scale, geo, acc, gacc, scratch=None):
        """acc += scale L'(w)^T mu; gacc (dict: rx, ry, sx, sy, fscale, nx, ny,
        bcu, bcv) += scale (dL/dgeo)^T mu."""
        import torch
        K, Np, NF, n = self.ncells, self.Np, self.NF, self.state_numel()
        scratch = scratch if scratch is not None else torch.empty(2 * K * NF * 4, dtype=torch.float64, device=w.device)
        sizes = dict(rx=K * Np, ry=K * Np, sx=K * Np, sy=K * Np, fscale=K * NF, nx=K * NF, ny=K * NF,
                     bcu=K * NF, bcv=K * NF)
        g = [_dev(geo[k], k, sizes[k]) for k in ("rx", "ry", "sx", "sy", "fscale", "nx", "ny")]
        ga = [_dev(gacc[k], "g" + k, sizes[k]) for k in ("rx", "ry", "sx", "sy", "fscale", "nx", "ny", "bcu", "bcv")]
        self._chk(lib.dg_rhs_adjoint(self._h, _dev(w, "w", n), _dev(mu, "mu", n), float(scale), *g,
                                     _dev(scratch, "scratch", 2 * K * NF * 4), _dev(acc, "acc", n), *ga))

    def rhs_adjoint_noalloc(self, w, mu, scale, geo, acc, gacc):
        """rhs_adjoint through dg_rhs_adjoint_noalloc: no scratch argument,
        the library's static per-GPU scratch instead (meshes up to
        DG_NOALLOC_SCRATCH_BYTES of slot cotangents)."""
        K, Np, NF, n = self.ncells, self.Np, self.NF, self.state_numel()
        sizes = dict(rx=K * Np, ry=K * Np, sx=K * Np, sy=K * Np, fscale=K * NF, nx=K * NF, ny=K * NF,
                     bcu=K * NF, bcv=K * NF)
        g = [_dev(geo[k], k, sizes[k]) for k in ("rx", "ry", "sx", "sy", "fscale", "nx", "ny")]
        ga = [_dev(gacc[k], "g" + k, sizes[k]) for k in ("rx", "ry", "sx", "sy", "fscale", "nx", "ny", "bcu", "bcv")]
        self._chk(lib.dg_rhs_adjoint_noalloc(self._h, _dev(w, "w", n), _dev(mu, "mu", n), float(scale), *g,
                                             _dev(acc, "acc", n), *ga))

    def set_workingQ(self, dt, stage, q, k1, k2, workingQ):
        n = self.state_numel()
        self._chk(lib.dg_set_workingQ(self._h, float(dt), int(stage), _dev(q, "q", n), _dev(k1, "k1", n),
                                      _dev(k2, "k2", n), _dev(workingQ, "workingQ", n)))

    def update_Q(self, dt, q, rk1, rk2, rk3, workingQ):
        n = self.state_numel()
        self._chk(lib.dg_update_Q(self._h, float(dt), _dev(q, "q", n), _dev(rk1, "rk1", n),
                                  _dev(rk2, "rk2", n), _dev(rk3, "rk3", n),
                                  _dev(workingQ, "workingQ", n)))

    def calc_dt(self, q, fscale, dt1=0.0):
        """OP_MAX reduction: returns max(dt1, per-cell CFL quantities)."""
        v = c_double(dt1)
        self._chk(lib.dg_calc_dt(self._h, _dev(q, "q", self.state_numel()),
                                 _dev(fscale, "fscale", self.ncells * self.NF), ctypes.byref(v)))
        return v.value


# --------------------------------------------------------------------------
# solver
# --------------------------------------------------------------------------

class Solver:
    """Native SSPRK3 driver owning the op_dats of one mesh (src/airfoil.cpp:204-334)."""

    def __init__(self, ctx, mesh):
        self.ctx, self.mesh = ctx, mesh
        h = c_void_p()
        check(lib.dg_solver_create(ctx._h, mesh._h, ctypes.byref(h)), ctx._h)
        self._h = h
        ctx.ncells = mesh.ncells

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.dg_solver_destroy(h)
            self._h = None

    def _chk(self, rc):
        return check(rc, self.ctx._h)

    def set_vortex(self, uinf, beta=5.0, x0=5.0, y0=0.0, period_x=10.0):
        self._chk(lib.dg_solver_set_vortex(self._h, uinf, beta, x0, y0, period_x))

    @property
    def state_shape(self):
        return (self.mesh.ncells, self.ctx.Np, 4)

    def get_q(self):
        q = np.zeros(self.state_shape)
        self._chk(lib.dg_solver_get_q(self._h, _np_ptr(q)))
        return q

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        if q.size != np.prod(self.state_shape):
            raise ValueError("state has the wrong size")
        self._chk(lib.dg_solver_set_q(self._h, _np_ptr(q)))

    def dat_ptr(self, name):
        p = c_void_p()
        self._chk(lib.dg_solver_dat(self._h, name.encode(), ctypes.byref(p)))
        return p.value

    def run(self, nsteps, cfl=1.0, mode=MODE_FUSED, graph=False):
        if graph and mode == MODE_FUSED:
            self._chk(lib.dg_solver_run_graph(self._h, int(nsteps), float(cfl)))
        else:
            self._chk(lib.dg_solver_run(self._h, int(nsteps), float(cfl), int(mode)))

    def stream(self):
        p = c_void_p()
        self._chk(lib.dg_solver_stream(self._h, ctypes.byref(p)))
        return p.value or 0

    def time(self):
        t, dt = c_double(), c_double()
        self._chk(lib.dg_solver_time(self._h, ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    def reset_time(self):
        self._chk(lib.dg_solver_reset_time(self._h))

    def check(self):
        return lib.dg_solver_check(self._h)

    def forces(self, chord=1.0):
        """(C_L, C_D, (Fx, Fy)) from the pressure on the wall faces."""
        cl, cd = c_double(), c_double()
        f = np.zeros(2)
        self._chk(lib.dg_solver_forces(self._h, float(chord), ctypes.byref(cl), ctypes.byref(cd), _np_ptr(f)))
        return cl.value, cd.value, f

    def wall_cp(self):
        """(x, Cp) at the face nodes of every wall face, shape [nwall, Nfp]."""
        n = c_int()
        self._chk(lib.dg_solver_wall_cp(self._h, ctypes.byref(n), None, None))
        x = np.zeros((n.value, self.ctx.Nfp))
        cp = np.zeros((n.value, self.ctx.Nfp))
        if n.value:
            self._chk(lib.dg_solver_wall_cp(self._h, ctypes.byref(n), _np_ptr(x), _np_ptr(cp)))
        return x, cp

    def profile(self, enable=True):
        self._chk(lib.dg_solver_profile(self._h, 1 if enable else 0))

    def vertex_fields(self, case=0):
        """Vertex values as src/save_solution.cpp forms them: dict of
        Density, Pressure, VelocityX, VelocityY arrays [nnodes]."""
        n = c_int()
        self._chk(lib.dg_solver_vertex_fields(self._h, int(case), ctypes.byref(n), None, None, None, None))
        out = {k: np.zeros(n.value) for k in ("Density", "VelocityX", "VelocityY", "Pressure")}
        self._chk(lib.dg_solver_vertex_fields(self._h, int(case), ctypes.byref(n), _np_ptr(out["Density"]),
                                              _np_ptr(out["VelocityX"]), _np_ptr(out["VelocityY"]),
                                              _np_ptr(out["Pressure"])))
        return out

    def write_vtk(self, path, case=0):
        """Mesh and vertex fields as a legacy-VTK ASCII file."""
        self._chk(lib.dg_solver_write_vtk(self._h, int(case), str(path).encode()))

    def set_launch(self, wpb=1, walk=0, max_blocks=0):
        """Launch shape of the matrix-core stage kernel (results unchanged)."""
        self._chk(lib.dg_solver_set_launch(self._h, int(wpb), int(walk), int(max_blocks)))

    def set_geometry(self, mode):
        """GEOM_NODE (default): the fused stage kernels read the per-node
        factors, normals and fscale of init_grid, as the reference does;
        GEOM_CELL: one metric record per cell (faster; leaves the reference
        by init_grid's per-node rounding scatter, ~1e-12 relative)."""
        self._chk(lib.dg_solver_set_geometry(self._h, int(mode)))

    def launch_config(self):
        w, k, b = c_int(), c_int(), c_int()
        self._chk(lib.dg_solver_launch_config(self._h, ctypes.byref(w), ctypes.byref(k), ctypes.byref(b)))
        return dict(wpb=w.value, walk=k.value, max_blocks=b.value)

    def autotune(self, trial_steps=3, cfl=1.0):
        """Times every launch shape on this solver's cases (state restored
        afterwards), keeps the fastest; returns the table of
        (wpb, walk, waves_per_cu, ms_per_step) rows."""
        tab = np.zeros((32, 4))
        n = c_int()
        self._chk(lib.dg_solver_autotune(self._h, int(trial_steps), float(cfl), _np_ptr(tab), ctypes.byref(n)))
        return [dict(wpb=int(r[0]), walk=int(r[1]), waves_per_cu=int(r[2]), ms_per_step=float(r[3]))
                for r in tab[:n.value]]

    def kernel_time(self, name):
        ms, n = c_double(), c_int()
        self._chk(lib.dg_solver_kernel_time(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value


class Sensitivity:
    """dC_L/d(alpha, bump amplitudes) of the NACA0012 O-grid case by the
    discrete adjoint, with tangent-linear and frozen-dt evaluations."""

    def __init__(self, ctx, ncirc, nrad, stretch, rfar, nbu, nbl, mach, alpha):
        self.ctx, self.nb = ctx, int(nbu) + int(nbl)
        h = c_void_p()
        check(lib.dg_sens_create(ctx._h, int(ncirc), int(nrad), float(stretch), float(rfar), int(nbu),
                                 int(nbl), float(mach), float(alpha), ctypes.byref(h)), ctx._h)
        self._h = h
        self.nsteps = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib.dg_sens_destroy(h)
            self._h = None

    def _chk(self, rc):
        return check(rc, self.ctx._h)

    def _b(self, b):
        if b is None:
            return None, None
        a = np.ascontiguousarray(b, dtype=np.float64)
        if a.size != self.nb:
            raise ValueError(f"expected {self.nb} bump values")
        return a, _np_ptr(a)

    def record(self, bumps=None, nsteps=400, cfl=1.0):
        a, p = self._b(bumps)
        cl = c_double()
        self._chk(lib.dg_sens_record(self._h, p, int(nsteps), float(cfl), ctypes.byref(cl)))
        self.nsteps = int(nsteps)
        return cl.value

    def set_alpha(self, alpha):
        self._chk(lib.dg_sens_set_alpha(self._h, float(alpha)))

    def eval(self, alpha, bumps=None):
        a, p = self._b(bumps)
        cl = c_double()
        self._chk(lib.dg_sens_eval(self._h, float(alpha), p, ctypes.byref(cl)))
        return cl.value

    def tangent(self, dalpha=0.0, dbumps=None):
        a, p = self._b(dbumps)
        cl, dcl = c_double(), c_double()
        self._chk(lib.dg_sens_tangent(self._h, float(dalpha), p, ctypes.byref(cl), ctypes.byref(dcl)))
        return cl.value, dcl.value

    def gradient(self, ckpt_every=20):
        cl, da = c_double(), c_double()
        db = np.zeros(self.nb)
        self._chk(lib.dg_sens_gradient(self._h, int(ckpt_every), ctypes.byref(cl), ctypes.byref(da),
                                       _np_ptr(db) if self.nb else None))
        return cl.value, da.value, db

    def timing(self):
        f, g, t = c_double(), c_double(), c_double()
        self._chk(lib.dg_sens_timing(self._h, ctypes.byref(f), ctypes.byref(g), ctypes.byref(t)))
        out = dict(forward_ms=f.value, gradient_ms=g.value, tangent_ms=t.value)
        self._chk(lib.dg_sens_gradient_breakdown(self._h, ctypes.byref(f), ctypes.byref(g), ctypes.byref(t)))
        out.update(gradient_forward_ms=f.value, gradient_reverse_ms=g.value, gradient_host_ms=t.value)
        return out

    def dts(self):
        d = np.zeros(self.nsteps)
        self._chk(lib.dg_sens_dts(self._h, self.nsteps, _np_ptr(d)))
        return d


class BatchSolver(Solver):
    """Many independent cases (uneven meshes, per-case Mach and angle) run by
    one persistent kernel launch per RK stage."""

    def __init__(self, ctx, meshes, mach, alpha):
        self.ctx, self.meshes = ctx, list(meshes)
        n = len(self.meshes)
        hs = (c_void_p * n)(*[m._h.value for m in self.meshes])
        ma = np.ascontiguousarray(mach, dtype=np.float64)
        al = np.ascontiguousarray(alpha, dtype=np.float64)
        if ma.size != n or al.size != n:
            raise ValueError("one Mach number and one angle per case")
        h = c_void_p()
        check(lib.dg_solver_create_batch(ctx._h, n, ctypes.cast(hs, c_void_p), _np_ptr(ma), _np_ptr(al),
                                         ctypes.byref(h)), ctx._h)
        self._h = h
        self.ncases = n
        self.mesh = None

    def case_q(self, k):
        q = np.zeros((self.meshes[k].ncells, self.ctx.Np, 4))
        self._chk(lib.dg_solver_case_get_q(self._h, int(k), _np_ptr(q)))
        return q

    def set_case_q(self, k, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._chk(lib.dg_solver_case_set_q(self._h, int(k), _np_ptr(q)))

    def case_time(self, k):
        t, dt = c_double(), c_double()
        self._chk(lib.dg_solver_case_time(self._h, int(k), ctypes.byref(t), ctypes.byref(dt)))
        return t.value, dt.value

    def case_forces(self, k, chord=1.0):
        cl, cd = c_double(), c_double()
        f = np.zeros(2)
        self._chk(lib.dg_solver_case_forces(self._h, int(k), float(chord), ctypes.byref(cl),
                                            ctypes.byref(cd), _np_ptr(f)))
        return cl.value, cd.value, f

    @property
    def state_shape(self):
        return (sum(m.ncells for m in self.meshes), self.ctx.Np, 4)


__all__ = ["Context", "Mesh", "Solver", "BatchSolver", "DGError", "generate_operators", "periodic_grid",
           "naca0012_ogrid", "read_vtk", "P", "c_void_p"]
