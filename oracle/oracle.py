"""Oracle: ctypes front end of the C restatement (oracle/dg_oracle.c).

TEST INFRASTRUCTURE ONLY.  Used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the shipped library never
imports it.

Every set-level helper here drives the per-element restatements in the
order of the reference's seq backend (src/seq/*_seqkernel.cpp): cells,
edges and boundary edges in index order.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return ctypes.CDLL(LIB_PATH)


lib = _load()
D = ctypes.c_double
I = ctypes.c_int
VP = ctypes.c_void_p


class Consts(ctypes.Structure):
    _fields_ = [("N", I), ("Np", I), ("Nfp", I),
                ("r", VP), ("s", VP), ("Dr", VP), ("Ds", VP), ("Drw", VP), ("Dsw", VP), ("LIFT", VP),
                ("FMASK", VP),
                ("gam", D), ("bc_mach", D), ("bc_alpha", D), ("bc_p", D), ("bc_r", D), ("bc_u", D),
                ("bc_v", D), ("bc_e", D)]


class MeshS(ctypes.Structure):
    _fields_ = [("ncells", I), ("nedges", I), ("nbedges", I),
                ("edge2cells", VP), ("edgeNum", VP), ("bedge2cells", VP), ("bedgeNum", VP),
                ("bedge_type", VP), ("nodeX", VP), ("nodeY", VP), ("rx", VP), ("ry", VP),
                ("sx", VP), ("sy", VP), ("nx", VP), ("ny", VP), ("fscale", VP),
                ("x", VP), ("y", VP), ("exact", VP)]


class StateS(ctypes.Structure):
    _fields_ = [("q", VP), ("workingQ", VP), ("exteriorQ", VP), ("rk1", VP), ("rk2", VP), ("rk3", VP)]


def _p(a):
    return a.ctypes.data_as(VP)


def freestream(gam=1.4, mach=0.4, alpha=0.0, p=1.0, rho=1.0):
    """src/airfoil.cpp:79-99."""
    import math
    a = math.sqrt(gam * p / rho) * mach
    if alpha != 0.0:
        u = math.sin((math.pi / 2.0) - (alpha * math.pi / 180.0)) * a
        v = math.cos((math.pi / 2.0) - (alpha * math.pi / 180.0)) * a
        e = p / (gam - 1.0) + 0.5 * rho * (u * u + v * v)
    else:
        u, v = a, 0.0
        e = p / (gam - 1.0) + 0.5 * rho * (u * u)
    return dict(gam=gam, bc_mach=mach, bc_alpha=alpha, bc_p=p, bc_r=rho, bc_u=u, bc_v=v, bc_e=e)


class Oracle:
    """Constants + set-level drivers at one order."""

    def __init__(self, ops, phys=None):
        self.ops = {k: (np.ascontiguousarray(v, dtype=np.int32) if k == "FMASK"
                        else np.ascontiguousarray(v, dtype=np.float64))
                    for k, v in ops.items() if k in ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK")}
        self.N = int(ops["N"])
        self.Np = (self.N + 1) * (self.N + 2) // 2
        self.Nfp = self.N + 1
        self.NF = 3 * self.Nfp
        self.phys = dict(freestream())
        if phys:
            self.phys.update(phys)
        self.exact = None
        self._mk()

    def _mk(self):
        o = self.ops
        c = Consts()
        c.N, c.Np, c.Nfp = self.N, self.Np, self.Nfp
        for k in ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK"):
            setattr(c, k, _p(o[k]))
        for k, v in self.phys.items():
            setattr(c, k, float(v))
        self.c = c

    def set_phys(self, **kw):
        self.phys.update(kw)
        self._mk()

    # ---- per-element restatements ----
    def euler_flux(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        F, G = np.zeros(4), np.zeros(4)
        rho, u, v, p = D(), D(), D(), D()
        lib.or_euler_flux(ctypes.byref(self.c), _p(q), _p(F), _p(G), ctypes.byref(rho),
                          ctypes.byref(u), ctypes.byref(v), ctypes.byref(p))
        return F, G, (rho.value, u.value, v.value, p.value)

    def roe(self, nx, ny, fscale, q, pQ):
        flux = np.zeros(4 * self.NF)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (nx, ny, fscale, q, pQ)]
        lib.or_roe(ctypes.byref(self.c), _p(flux), *[_p(a) for a in args])
        return flux.reshape(self.NF, 4)

    def euler_rhs(self, q, exteriorQ, rx, ry, sx, sy, fscale, nx, ny):
        ext = np.array(exteriorQ, dtype=np.float64).ravel().copy()
        out = np.zeros(4 * self.Np)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (rx, ry, sx, sy, fscale, nx, ny)]
        qq = np.ascontiguousarray(q, dtype=np.float64)
        lib.or_euler_rhs(ctypes.byref(self.c), _p(qq), _p(ext), *[_p(a) for a in args], _p(out))
        return out.reshape(self.Np, 4), ext

    def get_neighbour_q(self, edgeNum, xL, yL, xR, yR, qL, qR, exL, exR):
        en = np.ascontiguousarray(edgeNum, dtype=np.int32)
        exL = np.array(exL, dtype=np.float64).copy()
        exR = np.array(exR, dtype=np.float64).copy()
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (xL, yL, xR, yR, qL, qR)]
        lib.or_get_neighbour_q(ctypes.byref(self.c), _p(en), *[_p(a) for a in args], _p(exL), _p(exR))
        return exL, exR

    def get_bedge_q(self, btype, bnum, nx, ny, q, exQ):
        t = np.array([btype], dtype=np.int32)
        n = np.array([bnum], dtype=np.int32)
        ex = np.array(exQ, dtype=np.float64).copy()
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (nx, ny, q)]
        lib.or_get_bedge_q(ctypes.byref(self.c), _p(t), _p(n), *[_p(a) for a in args], _p(ex))
        return ex

    def set_workingQ(self, dt, stage, q, k1, k2):
        w = np.zeros(4 * self.Np)
        d, s = D(dt), I(stage)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (q, k1, k2)]
        lib.or_set_workingQ(ctypes.byref(self.c), ctypes.byref(d), ctypes.byref(s), *[_p(a) for a in args], _p(w))
        return w

    def update_Q(self, dt, q, k1, k2, k3):
        qq = np.array(q, dtype=np.float64).copy()
        w = np.zeros(4 * self.Np)
        d = D(dt)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (k1, k2, k3)]
        lib.or_update_Q(ctypes.byref(self.c), ctypes.byref(d), _p(qq), *[_p(a) for a in args], _p(w))
        return qq, w

    def calc_dt(self, q, fscale, dt1=0.0):
        v = D(dt1)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        fs = np.ascontiguousarray(fscale, dtype=np.float64)
        lib.or_calc_dt(ctypes.byref(self.c), _p(qq), _p(fs), ctypes.byref(v))
        return v.value

    def init_grid(self, n0, n1, n2):
        Np, NF = self.Np, self.NF
        outs = [np.zeros(3), np.zeros(3)] + [np.zeros(Np) for _ in range(10)] + [np.zeros(NF) for _ in range(3)]
        ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (n0, n1, n2)]
        lib.or_init_grid(ctypes.byref(self.c), *[_p(a) for a in ins], *[_p(a) for a in outs])
        keys = ["nodeX", "nodeY", "x", "y", "xr", "yr", "xs", "ys", "rx", "ry", "sx", "sy", "nx", "ny", "fscale"]
        return dict(zip(keys, outs))

    def set_ic(self):
        q, w = np.zeros(4 * self.Np), np.zeros(4 * self.Np)
        lib.or_set_ic(ctypes.byref(self.c), _p(q), _p(w))
        return q, w

    # ---- set-level ----
    def init_grid_all(self, cell_xy):
        K = cell_xy.shape[0]
        Np, NF = self.Np, self.NF
        cxy = np.ascontiguousarray(cell_xy, dtype=np.float64)
        shapes = [("nodeX", 3), ("nodeY", 3), ("x", Np), ("y", Np), ("xr", Np), ("yr", Np), ("xs", Np),
                  ("ys", Np), ("rx", Np), ("ry", Np), ("sx", Np), ("sy", Np), ("nx", NF), ("ny", NF),
                  ("fscale", NF)]
        g = {k: np.zeros((K, s)) for k, s in shapes}
        lib.or_init_grid_all(ctypes.byref(self.c), K, _p(cxy), *[_p(g[k]) for k, _ in shapes])
        # the cell's J: init_grid's value at node 0 (src/init_grid.h, same
        # operation order), so that fscale * J is init_grid's face sJ
        g["J"] = -g["xs"][:, 0] * g["yr"][:, 0] + g["xr"][:, 0] * g["ys"][:, 0]
        return g

    def make_case(self, cells, cell_xy, edge2cells=None, edgeNum=None, bedge2cells=None, bedgeNum=None,
                  bedge_type=None):
        """Geometry + connectivity for the set-level drivers."""
        K = cell_xy.shape[0]
        if edge2cells is None:
            e = build_edges(cells)
            edge2cells, edgeNum = e["edge2cells"], e["edgeNum"]
            bedge2cells, bedgeNum = e["bedge2cells"], e["bedgeNum"]
            bedge_type = np.zeros(len(bedge2cells), dtype=np.int32)
        case = dict(K=K, geo=self.init_grid_all(cell_xy),
                    edge2cells=np.ascontiguousarray(edge2cells, dtype=np.int32).reshape(-1, 2),
                    edgeNum=np.ascontiguousarray(edgeNum, dtype=np.int32).reshape(-1, 2),
                    bedge2cells=np.ascontiguousarray(bedge2cells, dtype=np.int32).ravel(),
                    bedgeNum=np.ascontiguousarray(bedgeNum, dtype=np.int32).ravel(),
                    bedge_type=np.ascontiguousarray(bedge_type, dtype=np.int32).ravel())
        return case

    @staticmethod
    def use_cell_geometry(case):
        """Replaces a case's per-node geometry, in place, by the per-cell form
        of the fused solver's DG_GEOM_CELL mode: rx, ry, sx, sy at every node
        = their node-0 values, and on face f the normal and fscale of the
        vectors (-sx, -sy), (rx + sx, ry + sy), (-rx, -ry), which are
        src/init_grid.h:80-104's face vectors divided by J."""
        g = case["geo"]
        met = [g[k][:, 0].copy() for k in ("rx", "ry", "sx", "sy")]
        for k, v in zip(("rx", "ry", "sx", "sy"), met):
            g[k][:] = v[:, None]
        rx, ry, sx, sy = met
        nfp = g["nx"].shape[1] // 3
        for f, (a, b) in enumerate([(-sx, -sy), (rx + sx, ry + sy), (-rx, -ry)]):
            L = np.sqrt(a * a + b * b)
            sl = slice(f * nfp, (f + 1) * nfp)
            g["fscale"][:, sl] = L[:, None]
            g["nx"][:, sl] = (a / L)[:, None]
            g["ny"][:, sl] = (b / L)[:, None]
        return case

    def _mesh_struct(self, case):
        g = case["geo"]
        m = MeshS()
        m.ncells, m.nedges, m.nbedges = case["K"], case["edge2cells"].shape[0], case["bedge2cells"].size
        m.edge2cells, m.edgeNum = _p(case["edge2cells"]), _p(case["edgeNum"])
        m.bedge2cells, m.bedgeNum, m.bedge_type = _p(case["bedge2cells"]), _p(case["bedgeNum"]), _p(case["bedge_type"])
        for k in ("nodeX", "nodeY", "rx", "ry", "sx", "sy", "nx", "ny", "fscale", "x", "y"):
            setattr(m, k, _p(g[k]))
        if np.any(case["bedge_type"] == 4):
            if self.exact is None:
                raise ValueError("boundary type 4 needs set_exact_state()")
            m.exact = _p(self.exact)
        return m

    def set_exact_state(self, uinf, beta=5.0, x0=5.0, y0=0.0, period_x=0.0):
        """Vortex of the exact-state boundaries (type 4, vortex/src/get_bedge_q.h)."""
        self.exact = np.array([uinf, beta, x0, y0, period_x], dtype=np.float64)

    def get_bedge_q_exact(self, bedgeNum, x, y, t, exteriorQ):
        """vortex/src/get_bedge_q.h:1-40 on one boundary edge; returns the new exteriorQ."""
        ext = np.ascontiguousarray(exteriorQ, dtype=np.float64).copy()
        xx = np.ascontiguousarray(x, dtype=np.float64)
        yy = np.ascontiguousarray(y, dtype=np.float64)
        lib.or_get_bedge_q_exact(ctypes.byref(self.c), ctypes.byref(I(int(bedgeNum))), _p(xx), _p(yy), D(t),
                                 _p(self.exact), _p(ext))
        return ext

    def new_state(self, case, q0):
        K, Np, NF = case["K"], self.Np, self.NF
        st = {"q": np.ascontiguousarray(q0, dtype=np.float64).reshape(K, Np, 4).copy()}
        st["workingQ"] = st["q"].copy()
        st["exteriorQ"] = np.zeros((K, NF, 4))
        for k in ("rk1", "rk2", "rk3"):
            st[k] = np.zeros((K, Np, 4))
        return st

    def run(self, case, state, nsteps, cfl=1.0, nthreads=1, ktime=None):
        """src/airfoil.cpp:231-334: returns (t, dt_next); state updated in place."""
        m = self._mesh_struct(case)
        s = StateS(*[_p(state[k]) for k in ("q", "workingQ", "exteriorQ", "rk1", "rk2", "rk3")])
        dtn = D()
        kt = np.zeros(6) if ktime is None else ktime
        lib.or_run.restype = D
        t = lib.or_run(ctypes.byref(self.c), ctypes.byref(m), ctypes.byref(s), int(nsteps), D(cfl),
                       int(nthreads), ctypes.byref(dtn), _p(kt))
        return t, dtn.value

    def run_dt(self, case, state, nsteps, dt, nthreads=1):
        """nsteps SSPRK3 steps of fixed size dt from time 0; returns the final time."""
        m = self._mesh_struct(case)
        s = StateS(*[_p(state[k]) for k in ("q", "workingQ", "exteriorQ", "rk1", "rk2", "rk3")])
        lib.or_run_dt.restype = D
        return lib.or_run_dt(ctypes.byref(self.c), ctypes.byref(m), ctypes.byref(s), int(nsteps), D(dt),
                             int(nthreads))

    def calc_dt_all(self, case, q, nthreads=1):
        m = self._mesh_struct(case)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        lib.or_calc_dt_all.restype = D
        return lib.or_calc_dt_all(ctypes.byref(self.c), ctypes.byref(m), _p(qq), D(0.0), int(nthreads))

    def rhs_all(self, case, w, t=0.0):
        """One pass of get_neighbour_q + get_bedge_q (and the exact-state loop
        of vortex/src/get_bedge_q.h at time t for type-4 edges) + euler_rhs
        over the mesh."""
        K, Np, NF = case["K"], self.Np, self.NF
        g = case["geo"]
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(K, Np, 4)
        ext = np.zeros((K, NF, 4))
        e2c, en = case["edge2cells"], case["edgeNum"]
        for e in range(e2c.shape[0]):
            L, R = e2c[e]
            ext[L], ext[R] = [a.reshape(NF, 4) for a in self.get_neighbour_q(
                en[e], g["nodeX"][L], g["nodeY"][L], g["nodeX"][R], g["nodeY"][R], w[L], w[R], ext[L], ext[R])]
        for b in range(case["bedge2cells"].size):
            c = case["bedge2cells"][b]
            if case["bedge_type"][b] == 4:
                ext[c] = self.get_bedge_q_exact(case["bedgeNum"][b], g["x"][c], g["y"][c], t,
                                                ext[c]).reshape(NF, 4)
                continue
            ext[c] = self.get_bedge_q(case["bedge_type"][b], case["bedgeNum"][b], g["nx"][c], g["ny"][c],
                                      w[c], ext[c]).reshape(NF, 4)
        rhs = np.zeros((K, Np, 4))
        ext_after = np.zeros_like(ext)
        for k in range(K):
            r, ex = self.euler_rhs(w[k], ext[k], g["rx"][k], g["ry"][k], g["sx"][k], g["sy"][k],
                                   g["fscale"][k], g["nx"][k], g["ny"][k])
            rhs[k] = r
            ext_after[k] = ex.reshape(NF, 4)
        return rhs, ext

    def wall_forces(self, case, q, facew):
        """Pressure force (Fx, Fy) on the wall (type 2) boundary faces:
        face quadrature with sJ = fscale * J (init_grid's J of the cell)."""
        sel = case["bedge_type"] == 2
        wc = np.ascontiguousarray(case["bedge2cells"][sel], dtype=np.int32)
        wf = np.ascontiguousarray(case["bedgeNum"][sel], dtype=np.int32)
        g = case["geo"]
        J = np.ascontiguousarray(g["J"])
        out = np.zeros(2)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (q, g["nx"], g["ny"], g["fscale"], J, facew)]
        lib.or_wall_forces(ctypes.byref(self.c), int(wc.size), _p(wc), _p(wf), *[_p(a) for a in arrs], _p(out))
        return out

    def wall_cp(self, case, q):
        """Pressure coefficient (p - p_inf) / (rho_inf |V_inf|^2 / 2) at the
        face nodes of the wall (type 2) boundary faces, in boundary-edge
        order: [nwall * Nfp]."""
        sel = case["bedge_type"] == 2
        wc = np.ascontiguousarray(case["bedge2cells"][sel], dtype=np.int32)
        wf = np.ascontiguousarray(case["bedgeNum"][sel], dtype=np.int32)
        ph = self.phys
        qdyn = 0.5 * ph["bc_r"] * (ph["bc_u"] ** 2 + ph["bc_v"] ** 2)
        out = np.zeros(wc.size * self.Nfp)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        lib.or_wall_cp(ctypes.byref(self.c), int(wc.size), _p(wc), _p(wf), _p(qq), D(qdyn), _p(out))
        return out

    def lift_drag(self, case, q, facew, chord=1.0):
        """(C_L, C_D) of the wall pressure force: the force of wall_forces
        projected on the free-stream normal / direction, over
        rho_inf |V_inf|^2 / 2 times the chord."""
        import math
        fx, fy = self.wall_forces(case, q, facew)
        ph = self.phys
        qdyn = 0.5 * ph["bc_r"] * (ph["bc_u"] ** 2 + ph["bc_v"] ** 2)
        al = math.radians(ph["bc_alpha"])
        return ((fy * math.cos(al) - fx * math.sin(al)) / (qdyn * chord),
                (fx * math.cos(al) + fy * math.sin(al)) / (qdyn * chord))

    def vortex_state(self, uinf, beta, x0, y0, t, x, y, period_x=0.0):
        x = np.asarray(x, dtype=np.float64)
        y = np.asarray(y, dtype=np.float64)
        out = np.zeros(x.shape + (4,))
        q4 = np.zeros(4)
        gam = self.phys["gam"]
        for idx in np.ndindex(x.shape):
            lib.or_vortex_state(D(gam), D(uinf), D(beta), D(x0), D(y0), D(t), D(x[idx]), D(y[idx]),
                                D(period_x), _p(q4))
            out[idx] = q4
        return out

    def vortex_l2(self, x, y, J, mass, q, uinf, beta, x0, y0, t, period_x=0.0):
        K = x.shape[0]
        err = np.zeros(4)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, J, mass, q)]
        lib.or_vortex_l2(ctypes.byref(self.c), K, *[_p(a) for a in arrs], D(self.phys["gam"]), D(uinf),
                         D(beta), D(x0), D(y0), D(t), D(period_x), _p(err))
        return err


def vertex_fields(cells, q, gam, nnodes, order):
    """src/save_solution.cpp:13-47 for order N: every cell, in order, writes
    its corner nodes (0, N, Np-1 for its vertices 0, 1, 2) into the vertex
    arrays, so the last cell containing a vertex wins; vertices of no cell
    stay 0.  Returns (rho, velX, velY, p)."""
    Np = (order + 1) * (order + 2) // 2
    q = np.asarray(q, dtype=np.float64).reshape(-1, Np, 4)
    rho, vx, vy, p = (np.zeros(nnodes) for _ in range(4))
    corner = (0, order, Np - 1)
    for i, cl in enumerate(np.asarray(cells).reshape(-1, 3)):
        for t in range(3):
            n, qq = int(cl[t]), q[i, corner[t]]
            rho[n] = qq[0]
            vx[n] = qq[1] / rho[n]
            vy[n] = qq[2] / rho[n]
            p[n] = (gam - 1) * (qq[3] - 0.5 * (qq[1] * vx[n] + qq[2] * vy[n]))
    return rho, vx, vy, p


def build_edges(cells):
    """tools/vtk2CGNS.cpp:112-199 edge contract (C restatement)."""
    cells = np.ascontiguousarray(cells, dtype=np.int32)
    K = cells.shape[0]
    n = 3 * K
    e2c, en, e2n = np.zeros(2 * n, np.int32), np.zeros(2 * n, np.int32), np.zeros(2 * n, np.int32)
    b2c, bn, b2n = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(2 * n, np.int32)
    ne, nb = I(), I()
    rc = lib.or_build_edges(K, _p(cells), ctypes.byref(ne), _p(e2c), _p(en), _p(e2n), ctypes.byref(nb),
                            _p(b2c), _p(bn), _p(b2n))
    if rc != 0:
        raise ValueError(f"or_build_edges failed ({rc})")
    ne, nb = ne.value, nb.value
    return dict(edge2cells=e2c[:2 * ne].reshape(ne, 2), edgeNum=en[:2 * ne].reshape(ne, 2),
                edge2nodes=e2n[:2 * ne].reshape(ne, 2), bedge2cells=b2c[:nb], bedgeNum=bn[:nb],
                bedge2nodes=b2n[:2 * nb].reshape(nb, 2))


def colour_edges(ncells, edge2cells):
    e2c = np.ascontiguousarray(edge2cells, dtype=np.int32)
    ne = e2c.shape[0]
    ptr, lst = np.zeros(9, np.int32), np.zeros(max(ne, 1), np.int32)
    nc = lib.or_colour_edges(int(ncells), ne, _p(e2c), _p(ptr), _p(lst))
    return nc, ptr[:nc + 1], lst[:ne]


def splitmix64_stream(seed):
    """splitmix64 (Steele, Lea & Flood 2014) generator of uniforms in [0, 1)."""
    mask = (1 << 64) - 1
    state = seed & mask
    while True:
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        yield (z >> 11) * (1.0 / 9007199254740992.0)


def periodic_grid_np(nx, ny, x0=0.0, y0=-5.0, lx=10.0, ly=10.0, seed=0, jitter=0.1):
    """Independent restatement of the jittered periodic grid (cells, cell_xy)."""
    hx, hy = lx / nx, ly / ny
    g = splitmix64_stream(seed)
    coords = np.zeros((ny, nx, 2))
    for j in range(ny):
        for i in range(nx):
            dx = (2.0 * next(g) - 1.0) * jitter * hx
            dy = (2.0 * next(g) - 1.0) * jitter * hy
            seam = i == 0 or j == 0
            coords[j, i] = (x0 + i * hx + (0 if seam else dx), y0 + j * hy + (0 if seam else dy))
    cells, cxy = [], []

    def vert(i, j):
        c = coords[j % ny, i % nx].copy()
        if i >= nx:
            c[0] += lx
        if j >= ny:
            c[1] += ly
        return (j % ny) * nx + (i % nx), c
    for j in range(ny):
        for i in range(nx):
            for tri in (((i, j), (i + 1, j), (i + 1, j + 1)), ((i, j), (i + 1, j + 1), (i, j + 1))):
                vs = [vert(*v) for v in tri]
                cells.append([v[0] for v in vs])
                cxy.append([v[1] for v in vs])
    return np.array(cells, dtype=np.int32), np.array(cxy)


def ref_lib():
    """The reference's own BLAS-free kernel headers, compiled (None if absent)."""
    if not os.path.exists(REF_LIB_PATH):
        return None
    return ctypes.CDLL(REF_LIB_PATH)
