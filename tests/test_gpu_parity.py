"""Parity of the HIP path with the oracle (FP64).

Tolerances (relative, per conserved variable unless stated):
  * pure copies (get_neighbour_q): bit-exact;
  * single kernels: <= 1e-13 (the GPU contracts multiply-adds into FMAs,
    the oracle rounds every product);
  * 100 SSPRK3 steps of the isentropic vortex (64x64 quads, 8192 cells,
    N=3, CFL 0.5): <= 1e-12 relative L2, the bound stated for this path.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

GAM = 1.4
UINF = 0.5 * math.sqrt(GAM)


def oracle_threads():
    """Host threads for the oracle's OpenMP loops: the process's CPU share."""
    import os
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n) or n)))


def mass_l2(J, M, d):
    """sqrt(sum_k J_k d_k^T M d_k) per variable of d [K][Np][4]."""
    d = np.asarray(d).reshape(len(J), M.shape[0], 4)
    return np.sqrt(np.einsum("k,kiv,ij,kjv->v", J, d, M, d))


def assert_l2_errors_agree(J, M, q_gpu, q_orc, e_gpu, e_orc):
    """The two L2 errors against the analytic state differ by at most the L2
    norm of the two states' difference (triangle inequality in the same
    mass-weighted norm), up to the rounding of the sums."""
    bound = mass_l2(J, M, np.asarray(q_gpu) - np.asarray(q_orc))
    diff = np.abs(np.asarray(e_gpu) - np.asarray(e_orc))
    assert np.all(diff <= bound * (1 + 1e-9) + 1e-14 * np.asarray(e_orc)), (diff, bound, e_gpu, e_orc)


def device_l2(dg, mesh, g, q, t, dev, x0=5.0, y0=0.0, period=10.0):
    """The library's own mass-weighted L2 error against the analytic vortex
    (dg_l2_error_mass) of the state q on the oracle's geometry g."""
    c = dg.Context(3)
    c.set_freestream(GAM, 0.5, 0.0, 1.0, 1.0)
    c.decl_mesh(mesh)
    c.set_exact_state(UINF, 5.0, x0, y0, period)
    d = [T(g[k], dev) for k in ("x", "y", "xr", "yr", "xs", "ys")]
    return c.l2_error_mass(*d, T(q, dev), t)


def assert_device_l2_agrees(J, M, q, e_dev, e_host):
    """The device's error of q and the oracle's differ by the two exact
    states' difference (exp / pow and contracted multiply-adds, within 1e-13
    of the state's unit scale) in the mass norm, and by the sums' rounding."""
    tol = 1e-13 * (np.abs(np.asarray(q)) + 1.0)
    bound = mass_l2(J, np.abs(M), tol) + 1e-12 * np.asarray(e_host)
    assert np.all(np.abs(np.asarray(e_dev) - np.asarray(e_host)) <= bound), (e_dev, e_host, bound)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def rel_per_var(a, b):
    a = np.asarray(a).reshape(-1, 4)
    b = np.asarray(b).reshape(-1, 4)
    return [float(np.linalg.norm(a[:, v] - b[:, v]) / np.linalg.norm(b[:, v])) for v in range(4)]


@pytest.fixture(scope="module")
def dev():
    assert torch.cuda.is_available(), "HIP device required"
    return torch.device("cuda:0")


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def smooth_state(x, y):
    k = 2 * np.pi / 10
    rho = 1 + 0.2 * np.sin(k * x) * np.cos(k * y)
    u = 0.5 + 0.1 * np.cos(k * x + 0.3)
    v = -0.2 + 0.1 * np.sin(k * y - 0.2)
    p = 1 + 0.1 * np.cos(k * (x - y))
    return np.stack([rho, rho * u, rho * v, p / (GAM - 1) + 0.5 * rho * (u * u + v * v)], axis=-1)


def setup_case(dg, orc, mesh, order=3, alpha=0.0, mach=0.4):
    ctx = dg.Context(order)
    ctx.set_freestream(GAM, mach, alpha, 1.0, 1.0)
    ctx.decl_mesh(mesh)
    o = orc.Oracle(dg.generate_operators(order), orc.freestream(GAM, mach, alpha))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    return ctx, o, case


@pytest.mark.parametrize("kind", ["periodic", "ogrid", "vortex724"])
def test_init_grid(dg, orc, dev, kind):
    """init_grid on the device gives the oracle's bits (every product and sum
    rounded on its own in the BLAS loop order, IEEE division and sqrt), on a
    small periodic grid, an O-grid (thin wall cells far from the origin in
    cell sizes) and the 1,048,352-cell bench mesh."""
    if kind == "periodic":
        mesh = dg.periodic_grid(6, 5, seed=2)
    elif kind == "ogrid":
        mesh = dg.naca0012_ogrid(64, 32, 1.08, 20.0)
    else:
        mesh = dg.periodic_grid(724, 724, 0.0, -5.0, 10.0, 10.0, seed=0, jitter=0.1)
    ctx, o, case = setup_case(dg, orc, mesh)
    K, Np, NF = mesh.ncells, ctx.Np, ctx.NF
    shapes = dict(nodeX=3, nodeY=3, x=Np, y=Np, xr=Np, yr=Np, xs=Np, ys=Np, rx=Np, ry=Np, sx=Np, sy=Np,
                  nx=NF, ny=NF, fscale=NF)
    d = {k: torch.zeros(K * s, dtype=torch.float64, device=dev) for k, s in shapes.items()}
    ctx.init_grid(T(mesh.cell_xy, dev), d)
    torch.cuda.synchronize()
    for k in shapes:
        ref = case["geo"][k].ravel()
        got = d[k].cpu().numpy()
        assert np.array_equal(got, ref), (k, np.abs(got - ref).max())


@pytest.mark.parametrize("kind", ["periodic", "ogrid"])
def test_stage_kernels_match_oracle(dg, orc, dev, kind):
    if kind == "periodic":
        mesh = dg.periodic_grid(6, 6, seed=5)
    else:
        mesh = dg.naca0012_ogrid(24, 6, stretch=1.3, rfar=8.0)
    ctx, o, case = setup_case(dg, orc, mesh, alpha=3.0)
    g = case["geo"]
    K, Np, NF = mesh.ncells, ctx.Np, ctx.NF
    if kind == "periodic":
        w = smooth_state(g["x"], g["y"])
    else:
        rng = np.random.default_rng(9)
        ph = o.phys
        qinf = np.array([ph["bc_r"], ph["bc_r"] * ph["bc_u"], ph["bc_r"] * ph["bc_v"], ph["bc_e"]])
        w = qinf * (1 + 0.02 * rng.normal(size=(K, Np, 4)))
    geo = {k: T(g[k], dev) for k in ("nodeX", "nodeY", "rx", "ry", "sx", "sy", "nx", "ny", "fscale")}
    wd = T(w, dev)
    ext = torch.full((K * NF * 4,), 7.0, dtype=torch.float64, device=dev)
    ctx.neighbour_zero(ext)
    ctx.get_neighbour_q(geo["nodeX"], geo["nodeY"], wd, ext)
    ctx.get_bedge_q(geo["nx"], geo["ny"], wd, ext)
    rhs_ref, ext_ref = o.rhs_all(case, w)
    ext_got = ext.cpu().numpy().reshape(K, NF, 4)
    # neighbour traces are exact copies; boundary states agree to rounding
    interior = np.zeros((K, NF), dtype=bool)
    for e in range(mesh.nedges):
        for side in range(2):
            c, f = mesh.edge2cells[e, side], mesh.edgeNum[e, side]
            interior[c, f * ctx.Nfp:(f + 1) * ctx.Nfp] = True
    assert np.array_equal(ext_got[interior], ext_ref[interior])
    if (~interior).any():
        assert np.abs(ext_got[~interior] - ext_ref[~interior]).max() <= 1e-13 * np.abs(ext_ref).max()
    rhs = torch.zeros(K * Np * 4, dtype=torch.float64, device=dev)
    ctx.euler_rhs(wd, ext, geo["rx"], geo["ry"], geo["sx"], geo["sy"], geo["fscale"], geo["nx"], geo["ny"], rhs)
    got = rhs.cpu().numpy()
    assert max(rel_per_var(got, rhs_ref)) <= 1e-12
    assert torch.count_nonzero(ext).item() == 0  # euler_rhs zeroes exteriorQ (OP_RW)
    fused = torch.zeros_like(rhs)
    ctx.rhs_fused(wd, geo["rx"], geo["ry"], geo["sx"], geo["sy"], geo["fscale"], geo["nx"], geo["ny"], fused)
    # the same arithmetic with the traces gathered instead of staged; the two
    # kernels are separate compilations, whose multiply-add contractions the
    # compiler picks per kernel, so they agree to rounding, not bit for bit
    assert max(rel_per_var(fused.cpu().numpy(), got)) <= 1e-13


def test_stage_updates_and_calc_dt(dg, orc, dev):
    mesh = dg.periodic_grid(5, 5, seed=1)
    ctx, o, case = setup_case(dg, orc, mesh)
    K, Np = mesh.ncells, ctx.Np
    rng = np.random.default_rng(3)
    q, k1, k2, k3 = [rng.normal(size=(K, Np, 4)) for _ in range(4)]
    q[..., 0] = 1 + 0.1 * np.abs(q[..., 0])
    q[..., 3] = 3 + np.abs(q[..., 3])
    dt = 0.013
    Q, K1, K2, K3 = [T(a, dev) for a in (q, k1, k2, k3)]
    W = torch.zeros_like(Q)
    for stage in (0, 1):
        ctx.set_workingQ(dt, stage, Q, K1, K2, W)
        ref = np.stack([o.set_workingQ(dt, stage, q[k], k1[k], k2[k]) for k in range(K)])
        assert rel(W.cpu().numpy().ravel(), ref.ravel()) <= 1e-15
    ctx.update_Q(dt, Q, K1, K2, K3, W)
    ref = np.stack([o.update_Q(dt, q[k], k1[k], k2[k], k3[k])[0] for k in range(K)])
    assert rel(Q.cpu().numpy().ravel(), ref.ravel()) <= 1e-15
    assert torch.equal(Q, W)
    qn = Q.cpu().numpy()
    got = ctx.calc_dt(Q, T(case["geo"]["fscale"], dev), 0.0)
    expect = o.calc_dt_all(case, qn)
    assert abs(got - expect) <= 1e-14 * expect
    # OP_MAX: a larger incoming value wins
    assert ctx.calc_dt(Q, T(case["geo"]["fscale"], dev), 10 * expect) == 10 * expect


def _vortex_solver(dg, n, order=3):
    mesh = dg.periodic_grid(n, n, 0.0, -5.0, 10.0, 10.0, seed=0, jitter=0.1)
    ctx = dg.Context(order)
    ctx.set_freestream(GAM, 0.5, 0.0, 1.0, 1.0)
    s = dg.Solver(ctx, mesh)
    s.set_vortex(UINF, 5.0, 5.0, 0.0, 10.0)
    return mesh, ctx, s


@pytest.fixture(scope="module")
def vortex64(dg, orc):
    """Config 0: 64x64 jittered periodic vortex, 100 steps at CFL 0.5 (oracle)."""
    mesh, ctx, s = _vortex_solver(dg, 64)
    q0 = s.get_q()
    o = orc.Oracle(dg.generate_operators(3), orc.freestream(GAM, 0.5, 0.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    st = o.new_state(case, q0)
    import os
    t, dtn = o.run(case, st, 100, cfl=0.5, nthreads=oracle_threads())
    return dict(mesh=mesh, q0=q0, q=st["q"].copy(), t=t, dt=dtn, case=case, o=o)


@pytest.fixture(scope="module")
def vortex64_cell(dg, orc, vortex64):
    """The same 100 steps in the oracle with the per-cell geometry of
    GEOM_CELL (Oracle.use_cell_geometry)."""
    import os
    o = vortex64["o"]
    mesh = vortex64["mesh"]
    case = o.use_cell_geometry(o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum,
                                           mesh.bedge2cells, mesh.bedgeNum, mesh.bedge_type))
    st = o.new_state(case, vortex64["q0"])
    t, dtn = o.run(case, st, 100, cfl=0.5, nthreads=oracle_threads())
    return dict(q=st["q"].copy(), t=t, dt=dtn)


@pytest.mark.parametrize("mode,geom", [(0, "node"), (0, "cell"), (1, "node")])
def test_vortex_100_steps_match_oracle(dg, vortex64, vortex64_cell, mode, geom):
    """Fused stages on the per-node (the reference's) or the per-cell
    geometry, and the nine-loop reference sequence (mode 1), each against the
    oracle on the same geometry."""
    mesh, ctx, s = _vortex_solver(dg, 64)
    s.set_geometry(dg.GEOM_CELL if geom == "cell" else dg.GEOM_NODE)
    ref = vortex64_cell if geom == "cell" else vortex64
    assert np.array_equal(s.get_q(), vortex64["q0"]) or rel(s.get_q(), vortex64["q0"]) < 1e-15
    s.set_q(vortex64["q0"])  # identical start
    s.run(100, cfl=0.5, mode=mode)
    assert s.check() == 0
    q = s.get_q()
    errs = rel_per_var(q, ref["q"])
    assert max(errs) <= 1e-12, errs
    t, dt = s.time()
    assert abs(t - ref["t"]) <= 1e-12 * ref["t"]
    assert abs(dt - ref["dt"]) <= 1e-11 * ref["dt"]
    if geom == "cell":
        # the per-cell geometry leaves the reference's per-node result by
        # init_grid's rounding scatter only
        assert max(rel_per_var(q, vortex64["q"])) <= 1e-11


def test_fused_run_is_deterministic_and_resumable(dg, vortex64):
    _, _, a = _vortex_solver(dg, 32)
    _, _, b = _vortex_solver(dg, 32)
    a.run(30, cfl=0.5)
    b.run(10, cfl=0.5)
    b.run(20, cfl=0.5)
    assert np.array_equal(a.get_q(), b.get_q())
    assert a.time() == b.time()


def test_vortex_error_matches_oracle_and_converges(dg, orc, dev):
    """L2 error against the analytic vortex, GPU vs oracle, at 16x16, 32x32
    and the 64x64 mesh of the CPU-sized benchmark case: the oracle's norm of
    both states, and the library's own device norm of the GPU state."""
    errs = []
    for n in (16, 32, 64):
        mesh, ctx, s = _vortex_solver(dg, n)
        q0 = s.get_q()
        steps = n  # same final time scale: dt ~ h
        s.run(steps, cfl=0.5)
        t, _ = s.time()
        o = orc.Oracle(dg.generate_operators(3), orc.freestream(GAM, 0.5, 0.0))
        case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                           mesh.bedgeNum, mesh.bedge_type)
        st = o.new_state(case, q0)
        to, _ = o.run(case, st, steps, cfl=0.5, nthreads=oracle_threads())
        g = case["geo"]
        area = 0.5 * ((g["nodeX"][:, 1] - g["nodeX"][:, 0]) * (g["nodeY"][:, 2] - g["nodeY"][:, 0])
                      - (g["nodeX"][:, 2] - g["nodeX"][:, 0]) * (g["nodeY"][:, 1] - g["nodeY"][:, 0]))
        M = dg.generate_operators(3)["mass"]
        qg = s.get_q()
        assert abs(t - to) <= 1e-12 * to
        e_gpu = o.vortex_l2(g["x"], g["y"], area / 2, M, qg, UINF, 5.0, 5.0, 0.0, t, 10.0)
        e_orc = o.vortex_l2(g["x"], g["y"], area / 2, M, st["q"], UINF, 5.0, 5.0, 0.0, to, 10.0)
        e_dev = device_l2(dg, mesh, g, qg, t, dev)
        print(f"vortex {n}x{n}, {steps} steps: L2 error vs analytic GPU {e_gpu} (device norm {e_dev}), "
              f"oracle {e_orc}", flush=True)
        assert_l2_errors_agree(area / 2, M, qg, st["q"], e_gpu, e_orc)
        assert_device_l2_agrees(area / 2, M, qg, e_dev, e_gpu)
        errs.append(e_gpu)
    ratio = np.array(errs[0]) / np.array(errs[1])
    assert np.all(ratio > 8.0), ratio  # observed order > 3 at N=3
    assert np.all(np.array(errs[2]) < np.array(errs[1])), errs


def test_full_size_vortex_matches_oracle_and_conserves(dg, orc, dev):
    """The bench workload at its full size (724x724 quads, 1,048,352 cells,
    N=3, CFL 0.5): 100 fused SSPRK3 steps against the oracle's 100 steps on
    the host threads (<= 1e-12 relative L2 per conserved variable); the
    mass-weighted totals of the conserved variables kept over the run
    (periodic domain, conservative fluxes: within 1e-12 of the largest
    total); a second run the same bits."""
    import os
    mesh, ctx, s = _vortex_solver(dg, 724)
    assert mesh.ncells == 1048352
    q0 = s.get_q()
    s.run(100, cfl=0.5)
    assert s.check() == 0
    q = s.get_q()
    o = orc.Oracle(dg.generate_operators(3), orc.freestream(GAM, 0.5, 0.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    st = o.new_state(case, q0)
    t, dt = o.run(case, st, 100, cfl=0.5, nthreads=oracle_threads())
    errs = rel_per_var(q, st["q"])
    print(f"724x724 vortex, 100 steps: rel L2 per variable vs oracle {errs}", flush=True)
    assert max(errs) <= 1e-12, errs
    assert abs(s.time()[0] - t) <= 1e-12 * t
    g = case["geo"]
    X, Y = g["nodeX"], g["nodeY"]
    area = 0.5 * ((X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (X[:, 2] - X[:, 0]) * (Y[:, 1] - Y[:, 0]))
    M = dg.generate_operators(3)["mass"]
    # L2 error against the analytic vortex, GPU state vs oracle state
    e_gpu = o.vortex_l2(g["x"], g["y"], area / 2, M, q, UINF, 5.0, 5.0, 0.0, t, 10.0)
    e_orc = o.vortex_l2(g["x"], g["y"], area / 2, M, st["q"], UINF, 5.0, 5.0, 0.0, t, 10.0)
    e_dev = device_l2(dg, mesh, g, q, t, dev)
    print(f"724x724 vortex L2 error vs analytic: GPU {e_gpu} (device norm {e_dev}), oracle {e_orc}", flush=True)
    assert_l2_errors_agree(area / 2, M, q, st["q"], e_gpu, e_orc)
    assert_device_l2_agrees(area / 2, M, q, e_dev, e_gpu)
    w = M.sum(axis=1)

    def totals(x):
        return np.einsum("k,i,kiv->v", area / 2, w, np.asarray(x).reshape(-1, 10, 4))
    t0, t1 = totals(q0), totals(q)
    assert np.max(np.abs(t1 - t0)) <= 1e-12 * np.max(np.abs(t0)), (t0, t1)
    _, _, s2 = _vortex_solver(dg, 724)
    s2.run(100, cfl=0.5)
    assert np.array_equal(s2.get_q(), q)


@pytest.mark.parametrize("geom", ["node", "cell"])
def test_airfoil_ogrid_short_run_matches_oracle(dg, orc, geom):
    """O-grid with wall and characteristic far field, alpha = 2 deg, M = 0.4,
    per-node (the reference's) and per-cell geometry in the stage kernels,
    each against the oracle on the same geometry; forces and C_p on the
    per-node geometry in both."""
    import os
    mesh = dg.naca0012_ogrid(48, 16, stretch=1.15, rfar=20.0)
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.4, 2.0, 1.0, 1.0)
    s = dg.Solver(ctx, mesh)
    s.set_geometry(dg.GEOM_CELL if geom == "cell" else dg.GEOM_NODE)
    q0 = s.get_q()
    s.run(50, cfl=1.0)
    assert s.check() == 0
    o = orc.Oracle(dg.generate_operators(3), orc.freestream(GAM, 0.4, 2.0))

    def mk():
        return o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                           mesh.bedgeNum, mesh.bedge_type)
    case = mk()
    run_case = o.use_cell_geometry(mk()) if geom == "cell" else case
    st = o.new_state(run_case, q0)
    o.run(run_case, st, 50, cfl=1.0, nthreads=oracle_threads())
    qg = s.get_q()
    assert max(rel_per_var(qg, st["q"])) <= 1e-12
    # surface force: GPU quadrature vs the oracle's on the same state, and
    # on the oracle's own trajectory
    facew = dg.generate_operators(3)["face_mass"].sum(axis=1)
    cl, cd, fxy = s.forces(1.0)
    f_same = o.wall_forces(case, qg, facew)
    f_orc = o.wall_forces(case, st["q"], facew)
    # the net force is a small difference of O(1) pressure integrals over a
    # perimeter of ~2 chords: tolerances are relative to that integral
    scale = 2.0 * np.abs(qg[..., 3]).max()
    assert np.abs(fxy - f_same).max() <= 1e-13 * scale
    assert np.abs(fxy - f_orc).max() <= 1e-11 * scale
    x, cp = s.wall_cp()
    assert x.shape == cp.shape == (48, 4)
    assert np.all((x >= -1e-12) & (x <= 1 + 1e-12))


@pytest.mark.parametrize("order", [1, 2, 4])
@pytest.mark.parametrize("geom", ["node", "cell"])
def test_fused_solver_other_orders_match_oracle(dg, orc, order, geom):
    """The fused stage kernels at N = 1, 2, 4 (other tile fillings, face-flux
    padding columns, Np not a multiple of 4) on an O-grid with wall and far
    field, 20 steps, each geometry mode against the oracle on the same
    geometry."""
    import os
    mesh = dg.naca0012_ogrid(32, 8, stretch=1.2, rfar=10.0)
    ctx = dg.Context(order)
    ctx.set_freestream(GAM, 0.4, 3.0, 1.0, 1.0)
    s = dg.Solver(ctx, mesh)
    s.set_geometry(dg.GEOM_CELL if geom == "cell" else dg.GEOM_NODE)
    q0 = s.get_q()
    s.run(20, cfl=0.5)
    assert s.check() == 0
    o = orc.Oracle(dg.generate_operators(order), orc.freestream(GAM, 0.4, 3.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    if geom == "cell":
        o.use_cell_geometry(case)
    st = o.new_state(case, q0)
    o.run(case, st, 20, cfl=0.5, nthreads=oracle_threads())
    errs = rel_per_var(s.get_q(), st["q"])
    assert max(errs) <= 1e-12, errs


@pytest.mark.parametrize("mode", [0, 1])
def test_reference_airfoil_scenario(dg, orc, naca_vtk, mode):
    """The reference's own airfoil run end to end: its mesh (naca0012.vtk,
    2450 triangles, through the VTK reader), its order and literal N=4
    tables, M 0.4, alpha 0, dt = 1/dt1 every step (src/airfoil.cpp:67-99,
    252-334); fused stages (mode 0) and the nine-loop sequence (mode 1)
    against the oracle, 20 steps."""
    import json
    import os
    tab = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "n4_tables.json")))
    mesh = dg.read_vtk(naca_vtk)
    ctx = dg.Context(4)
    for k in ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK"):
        ctx.set_const(k, tab[k])
    ctx.set_freestream(GAM, 0.4, 0.0, 1.0, 1.0)
    s = dg.Solver(ctx, mesh)
    q0 = s.get_q()
    s.run(20, cfl=1.0, mode=mode)
    assert s.check() == 0
    ops = {k: np.array(tab[k]) for k in ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK")}
    ops["N"] = 4
    o = orc.Oracle(ops, orc.freestream(GAM, 0.4, 0.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    st = o.new_state(case, q0)
    t, _ = o.run(case, st, 20, cfl=1.0, nthreads=oracle_threads())
    qg = s.get_q()
    errs = rel_per_var(qg, st["q"])
    print(f"naca0012.vtk N=4 mode {mode}, 20 steps: rel L2 per variable vs oracle {errs}", flush=True)
    assert max(errs) <= 1e-12, errs
    assert abs(s.time()[0] - t) <= 1e-12 * t


def test_reference_tables_order4(dg, orc, dev):
    """N=4 with the reference's own literal tables installed via set_const."""
    import json
    import os
    tab = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "n4_tables.json")))
    mesh = dg.periodic_grid(5, 5, seed=4)
    ctx = dg.Context(4)
    ctx.set_freestream(GAM, 0.4, 0.0, 1.0, 1.0)
    for k in ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK"):
        ctx.set_const(k, tab[k])
    assert np.array_equal(ctx.get_const("LIFT"), np.array(tab["LIFT"]))
    with pytest.raises(dg.DGError):
        ctx.set_const("not_a_constant", [1.0])
    ctx.decl_mesh(mesh)
    ops = {k: np.array(tab[k]) for k in ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK")}
    ops["N"] = 4
    o = orc.Oracle(ops, orc.freestream(GAM, 0.4, 0.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    g = case["geo"]
    w = smooth_state(g["x"], g["y"])
    geo = {k: T(g[k], dev) for k in ("rx", "ry", "sx", "sy", "nx", "ny", "fscale")}
    out = torch.zeros(w.size, dtype=torch.float64, device=dev)
    ctx.rhs_fused(T(w, dev), geo["rx"], geo["ry"], geo["sx"], geo["sy"], geo["fscale"], geo["nx"],
                  geo["ny"], out)
    ref, _ = o.rhs_all(case, w)
    assert max(rel_per_var(out.cpu().numpy(), ref)) <= 1e-12


def test_vertex_export_matches_oracle(dg, orc, tmp_path):
    """Vertex values (save_solution semantics) of a solver state and the
    legacy-VTK file carrying them."""
    mesh = dg.naca0012_ogrid(24, 6, 1.15, 8.0)
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.45, 2.0, 1.0, 1.0)
    s = dg.Solver(ctx, mesh)
    s.run(4, cfl=1.0)
    v = s.vertex_fields()
    rho, vx, vy, p = orc.vertex_fields(mesh.cells, s.get_q(), GAM, mesh.nnodes, 3)
    for got, ref in ((v["Density"], rho), (v["VelocityX"], vx), (v["VelocityY"], vy), (v["Pressure"], p)):
        assert np.array_equal(got, ref)
    path = tmp_path / "sol.vtk"
    s.write_vtk(path)
    tok = open(path).read().split()
    i = tok.index("POINTS")
    assert int(tok[i + 1]) == mesh.nnodes
    pts = np.array(tok[i + 3:i + 3 + 3 * mesh.nnodes], dtype=float).reshape(-1, 3)
    assert np.array_equal(pts[:, :2], mesh.coords)
    j = tok.index("CELLS")
    assert int(tok[j + 1]) == mesh.ncells
    k = tok.index("Density")
    dens = np.array(tok[k + 5:k + 5 + mesh.nnodes], dtype=float)
    assert np.array_equal(dens, rho)


def _exact_box(dg, orc, n):
    """Jittered box [0,10] x [-5,5] whose boundary edges all carry the exact
    vortex state; the vortex moves at u_inf from (5, -4), one unit above the
    bottom edge, so the boundary cuts through its flow."""
    mesh = dg.rect_grid(n, n, 0.0, -5.0, 10.0, 10.0, seed=4, jitter=0.1)
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.5, 0.0, 1.0, 1.0)
    o = orc.Oracle(dg.generate_operators(3), orc.freestream(GAM, 0.5, 0.0))
    o.set_exact_state(UINF, 5.0, 5.0, -4.0, 0.0)
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    return mesh, ctx, o, case


def test_exact_boundary_loop_matches_oracle(dg, orc, dev):
    """dg_get_bedge_q_exact (vortex/src/get_bedge_q.h over every boundary
    edge, OP_INC) against the oracle's loop at a stage time."""
    mesh, ctx, o, case = _exact_box(dg, orc, 12)
    ctx.decl_mesh(mesh)
    ctx.set_exact_state(UINF, 5.0, 5.0, -4.0, 0.0)
    K, NF = mesh.ncells, ctx.NF
    g = case["geo"]
    rng = np.random.default_rng(5)
    ext0 = rng.normal(size=(K, NF, 4))
    ext = T(ext0, dev)
    ctx.get_bedge_q_exact(T(g["x"], dev), T(g["y"], dev), 0.37, ext)
    ref = ext0.copy()
    for b in range(mesh.nbedges):
        c = case["bedge2cells"][b]
        ref[c] = o.get_bedge_q_exact(case["bedgeNum"][b], g["x"][c], g["y"][c], 0.37, ref[c]).reshape(NF, 4)
    got = ext.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("order,n", [(3, 12), (3, 64), (1, 9), (2, 7), (4, 8)])
def test_l2_error_loop_matches_oracle(dg, orc, dev, order, n):
    """dg_l2_error (vortex/src/l2_error.h over cells, vortex/src/vortex.cpp:
    314-328): the squared nodal errors against the oracle's exact vortex and
    their sums per variable added to the caller's accumulators (OP_INC);
    dg_l2_error_mass against the oracle's mass-weighted norm.  The 64 x 64
    box has more passes than workgroups, the other sizes one pass or fewer
    per workgroup; repeated calls give the same bits."""
    mesh = dg.rect_grid(n, n, 0.0, -5.0, 10.0, 10.0, seed=4, jitter=0.1)
    ctx = dg.Context(order)
    ctx.set_freestream(GAM, 0.5, 0.0, 1.0, 1.0)
    ops = dg.generate_operators(order)
    o = orc.Oracle(ops, orc.freestream(GAM, 0.5, 0.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    ctx.decl_mesh(mesh)
    ctx.set_exact_state(UINF, 5.0, 3.0, -1.0, 10.0)
    g = case["geo"]
    K, Np = mesh.ncells, ctx.Np
    t = 0.83
    ex = o.vortex_state(UINF, 5.0, 3.0, -1.0, t, g["x"], g["y"], 10.0)
    rng = np.random.default_rng(order * 100 + n)
    q = ex + 1e-3 * rng.normal(size=ex.shape)
    sq = (q - ex) ** 2
    x, y, qd = T(g["x"], dev), T(g["y"], dev), T(q, dev)
    error = torch.full((K, Np, 4), -1.0, dtype=torch.float64, device=dev)
    acc0 = np.array([1.0, 2.0, 3.0, 4.0])
    got = ctx.l2_error(x, y, qd, t, error, acc0)
    got_err = error.cpu().numpy()
    # the two exact states differ by the device's and the host's exp / pow
    # and by the multiply-adds the device contracts in x - x0 - uinf t and r^2
    # (relative error up to about r^2 ulps in exp(1 - r^2)): bounded by 1e-13
    # of the state's unit scale.  A difference d moves e^2 by at most
    # 2|e|d + d^2.
    tol = 1e-13 * (np.abs(ex) + 1.0)
    dsq = 2 * np.sqrt(sq) * tol + tol ** 2 + 1e-15 * sq
    worst = np.abs(got_err - sq) / dsq
    assert worst.max() <= 1.0, (worst.max(), np.unravel_index(worst.argmax(), worst.shape))
    ref = acc0 + sq.reshape(-1, 4).sum(axis=0)
    bound = dsq.reshape(-1, 4).sum(axis=0) + 1e-13 * ref
    assert np.all(np.abs(got - ref) <= bound), (got, ref, bound)
    again = ctx.l2_error(x, y, qd, t, None, acc0)
    assert np.array_equal(again, got)
    geo = [T(g[k], dev) for k in ("xr", "yr", "xs", "ys")]
    wl2 = ctx.l2_error_mass(x, y, *geo, qd, t)
    M = ops["mass"]
    e_orc = o.vortex_l2(g["x"], g["y"], g["J"], M, q, UINF, 5.0, 3.0, -1.0, t, 10.0)
    # triangle inequality in the mass norm: |d|^T |M| |d| bounds d^T M d
    bound = mass_l2(g["J"], np.abs(M), tol) + 1e-12 * e_orc
    assert np.all(np.abs(wl2 - e_orc) <= bound), (wl2, e_orc, bound)
    # the weighted norm of the errors' own closed form: at q = exact it is 0
    zero = ctx.l2_error_mass(x, y, *geo, T(ex, dev), t)
    assert np.all(zero <= mass_l2(g["J"], np.abs(M), tol)), zero
    assert np.array_equal(ctx.l2_error_mass(x, y, *geo, qd, t), wl2)


@pytest.mark.parametrize("mode", [0, 1])
def test_exact_boundary_vortex_run_matches_oracle(dg, orc, mode):
    """The vortex case with exact-state boundaries (vortex/src/vortex.cpp:
    228-317): 100 SSPRK3 steps at CFL 0.5 in the fused solver (hipGraph
    replay, exact states written per stage at t, t + dt, t + dt/2) and in the
    nine-loop sequence, against the oracle: <= 1e-12 relative L2 per
    variable, and the same L2 error against the analytic vortex."""
    mesh, ctx, o, case = _exact_box(dg, orc, 32)
    s = dg.Solver(ctx, mesh)
    s.set_vortex(UINF, 5.0, 5.0, -4.0, 0.0)
    q0 = s.get_q()
    s.run(100, cfl=0.5, mode=mode, graph=(mode == 0))
    assert s.check() == 0
    q = s.get_q()
    st = o.new_state(case, q0)
    t, _ = o.run(case, st, 100, cfl=0.5, nthreads=oracle_threads())
    errs = rel_per_var(q, st["q"])
    print(f"exact-boundary vortex 32x32, mode {mode}, 100 steps: rel L2 per variable {errs}", flush=True)
    assert max(errs) <= 1e-12, errs
    assert abs(s.time()[0] - t) <= 1e-12 * t
    g = case["geo"]
    X, Y = g["nodeX"], g["nodeY"]
    area = 0.5 * ((X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (X[:, 2] - X[:, 0]) * (Y[:, 1] - Y[:, 0]))
    M = dg.generate_operators(3)["mass"]
    e_gpu = o.vortex_l2(g["x"], g["y"], area / 2, M, q, UINF, 5.0, 5.0, -4.0, t)
    e_orc = o.vortex_l2(g["x"], g["y"], area / 2, M, st["q"], UINF, 5.0, 5.0, -4.0, t)
    assert_l2_errors_agree(area / 2, M, q, st["q"], e_gpu, e_orc)


def test_reference_vortex_scenario(dg, orc, dev, vortex_vtk):
    """The reference's vortex run end to end (vortex/src/vortex.cpp:195-328)
    on the library's per-loop entry points, driven in the reference's order:
    its mesh (vortex/grid.vtk, 618 triangles, through the VTK reader, every
    boundary edge the exact state), the literal N=4 tables, the vortex
    (beta 5 at (5, 0), u_inf 1, gamma 1.4) as the initial state, dt = 0.1 /
    dt1 from the initial state and fixed from then on (:213-222), and per
    stage neighbour_zero, get_neighbour_q, the exact-state boundary at
    t + {0, 1, 1/2} dt, euler_rhs and (stages 0, 1) set_workingQ, update_Q
    per step (:230-307); 2000 steps.  Against the oracle's fixed-dt run:
    <= 1e-12 relative L2 per variable, and the L2 errors against the
    analytic vortex at the end (l2_error's role, :314-328), the library's
    device norm and the oracle's, agree."""
    import json
    import os
    tab = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "n4_tables.json")))
    names = ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT", "FMASK")
    mesh = dg.read_vtk(vortex_vtk)
    mesh.set_bedge_types(dg.BC_EXACT)
    uinf, beta, x0, y0 = 1.0, 5.0, 5.0, 0.0
    mach = uinf / math.sqrt(GAM)  # p = rho = 1
    ctx = dg.Context(4)
    for k in names:
        ctx.set_const(k, tab[k])
    ctx.set_freestream(GAM, mach, 0.0, 1.0, 1.0)
    ctx.decl_mesh(mesh)
    ctx.set_exact_state(uinf, beta, x0, y0, 0.0)
    K, Np, NF = mesh.ncells, ctx.Np, ctx.NF

    def z(n):
        return torch.zeros(n, dtype=torch.float64, device=dev)
    geo = {k: z(K * 3) for k in ("nodeX", "nodeY")}
    geo.update({k: z(K * Np) for k in ("x", "y", "xr", "yr", "xs", "ys", "rx", "ry", "sx", "sy")})
    geo.update({k: z(K * NF) for k in ("nx", "ny", "fscale")})
    ctx.init_grid(T(mesh.cell_xy, dev), geo)

    ops = {k: np.array(tab[k]) for k in names}
    ops["N"] = 4
    o = orc.Oracle(ops, orc.freestream(GAM, mach, 0.0))
    o.set_exact_state(uinf, beta, x0, y0, 0.0)
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    g = case["geo"]
    for k in ("x", "y", "rx", "sy", "nx", "fscale"):
        assert np.array_equal(geo[k].cpu().numpy(), g[k].ravel()), k  # init_grid: the oracle's bits

    q0 = o.vortex_state(uinf, beta, x0, y0, 0.0, g["x"], g["y"])  # set_ic at t = 0
    q, w = T(q0, dev).reshape(-1), T(q0, dev).reshape(-1)
    ext = z(K * NF * 4)
    rk = [z(K * Np * 4) for _ in range(3)]
    dt1 = ctx.calc_dt(q, geo["fscale"])
    dt1_o = o.calc_dt_all(case, q0)
    assert abs(dt1 - dt1_o) <= 1e-14 * dt1_o, (dt1, dt1_o)
    dt = 0.1 / dt1
    nsteps, t = 2000, 0.0
    frac = (0.0, 1.0, 0.5)
    for _ in range(nsteps):
        for j in range(3):
            ctx.neighbour_zero(ext)
            ctx.get_neighbour_q(geo["nodeX"], geo["nodeY"], w, ext)
            ctx.get_bedge_q_exact(geo["x"], geo["y"], t + frac[j] * dt, ext)
            ctx.euler_rhs(w, ext, geo["rx"], geo["ry"], geo["sx"], geo["sy"], geo["fscale"], geo["nx"],
                          geo["ny"], rk[j])
            if j != 2:
                ctx.set_workingQ(dt, j, q, rk[0], rk[1], w)
        ctx.update_Q(dt, q, rk[0], rk[1], rk[2], w)
        t += dt
    torch.cuda.synchronize()
    qg = q.cpu().numpy().reshape(K, Np, 4)
    assert np.all(np.isfinite(qg))

    st = o.new_state(case, q0)
    t_o = o.run_dt(case, st, nsteps, dt, nthreads=oracle_threads())
    assert abs(t_o - t) <= 1e-13 * t
    errs = rel_per_var(qg, st["q"])
    print(f"vortex grid.vtk N=4, {nsteps} steps of dt {dt:.4e}: rel L2 per variable vs oracle {errs}",
          flush=True)
    assert max(errs) <= 1e-12, errs

    X, Y = g["nodeX"], g["nodeY"]
    area = 0.5 * ((X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (X[:, 2] - X[:, 0]) * (Y[:, 1] - Y[:, 0]))
    M = dg.generate_operators(4)["mass"]
    e_gpu = o.vortex_l2(g["x"], g["y"], area / 2, M, qg, uinf, beta, x0, y0, t)
    e_orc = o.vortex_l2(g["x"], g["y"], area / 2, M, st["q"], uinf, beta, x0, y0, t)
    assert_l2_errors_agree(area / 2, M, qg, st["q"], e_gpu, e_orc)
    e_dev = ctx.l2_error_mass(geo["x"], geo["y"], geo["xr"], geo["yr"], geo["xs"], geo["ys"], q, t)
    assert_device_l2_agrees(area / 2, M, qg, e_dev, e_gpu)
    print(f"vortex grid.vtk: L2 error vs the analytic vortex at t = {t:.4f}: GPU {e_gpu}, oracle {e_orc}, "
          f"device norm {e_dev}", flush=True)
