"""The benchmark configurations at their full sizes on the HIP path (FP64,
N=3), each against the oracle or its own defining property:

  * NACA0012 O-grid 256 x 128 (65,536 cells), Mach 0.4, alpha 0, slip wall
    and characteristic far field, 1000 SSPRK3 steps at CFL 0.5 (the
    reference's stepping and boundary sequence, src/airfoil.cpp:252-334):
    state <= 1e-12 relative L2 per conserved variable against the oracle,
    wall force vector <= 1e-12 relative, C_L / C_D and every wall node's C_p
    against the oracle's;
  * 64-case uneven O-grid sweep (8.5k..128k cells per case), 500 steps at
    CFL 0.25 in one persistent launch per stage with hipGraph-captured
    steps: every case bit-identical to its own single-case run;
  * lift sensitivities on the 256 x 128 mesh with 16 Hicks-Henne bumps,
    alpha 2 deg, 400 steps, checkpoints every 20: the adjoint gradient
    against the tangent-linear run (1e-7 per component) and against finite
    differences (1e-6 per component, no absolute slack).
"""
import math
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GAM = 1.4


def oracle_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, int(os.environ.get("OMP_NUM_THREADS", n) or n)))


def rel_per_var(a, b):
    a = np.asarray(a).reshape(-1, 4)
    b = np.asarray(b).reshape(-1, 4)
    return [float(np.linalg.norm(a[:, v] - b[:, v]) / np.linalg.norm(b[:, v])) for v in range(4)]


def test_airfoil_ogrid_256x128_1000_steps(dg, orc):
    """After 100 and after the full 1000 steps: state (relative L2 per
    conserved variable), wall force vector, C_L, C_D (relative to the force
    magnitude) and every wall node's C_p within 1e-12 of the oracle's."""
    mesh = dg.naca0012_ogrid(256, 128, 1.08, 20.0)
    assert mesh.ncells == 65536
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.4, 0.0, 1.0, 1.0)
    s = dg.Solver(ctx, mesh)
    q0 = s.get_q()
    o = orc.Oracle(dg.generate_operators(3), orc.freestream(GAM, 0.4, 0.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    st = o.new_state(case, q0)
    facew = dg.generate_operators(3)["face_mass"].sum(axis=1)
    qdyn = 0.5 * 0.4 ** 2 * GAM
    t = 0.0
    for steps, done in ((100, 100), (900, 1000)):
        s.run(steps, cfl=0.5, graph=True)
        assert s.check() == 0
        q = s.get_q()
        cl, cd, fxy = s.forces()
        _, cp = s.wall_cp()
        cp = np.asarray(cp).ravel()
        t += o.run(case, st, steps, cfl=0.5, nthreads=oracle_threads())[0]
        errs = rel_per_var(q, st["q"])
        f_o = np.array(o.wall_forces(case, st["q"], facew))
        cl_o, cd_o = o.lift_drag(case, st["q"], facew)
        cp_o = o.wall_cp(case, st["q"])
        df = np.linalg.norm(np.asarray(fxy) - f_o) / np.linalg.norm(f_o)
        dcp = np.abs(cp - cp_o).max() / np.abs(cp_o).max()
        print(f"O-grid 256x128, {done} steps: rel L2 per variable {errs}; force rel {df:.3e}; "
              f"C_L {cl!r} vs {cl_o!r}, C_D {cd!r} vs {cd_o!r}; C_p max rel {dcp:.3e}", flush=True)
        assert max(errs) <= 1e-12, errs
        assert abs(s.time()[0] - t) <= 1e-12 * t
        assert df <= 1e-12
        # C_L and C_D: components of the force over q_inf * chord
        fnorm = np.linalg.norm(f_o) / qdyn
        assert abs(cl - cl_o) <= 1e-12 * fnorm and abs(cd - cd_o) <= 1e-12 * fnorm
        assert cp.shape == cp_o.shape == (256 * 4,)
        assert dcp <= 1e-12


def test_sweep_64_cases_500_steps_bit_identical(dg):
    specs = dg.sweep_cases(64)
    meshes = [dg.naca0012_ogrid(nc, nr, 1.08, 20.0) for nc, nr, _, _ in specs]
    cells = [m.ncells for m in meshes]
    assert min(cells) >= 8000 and max(cells) <= 130000
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.4, 0.0, 1.0, 1.0)
    batch = dg.BatchSolver(ctx, meshes, [sp[2] for sp in specs], [sp[3] for sp in specs])
    batch.run(500, cfl=0.25, graph=True)
    assert batch.check() == 0
    print(f"sweep: {len(specs)} cases, {sum(cells)} cells; batched run done", flush=True)
    for k, (m, (_, _, mach, alpha)) in enumerate(zip(meshes, specs)):
        c1 = dg.Context(3)
        c1.set_freestream(GAM, mach, alpha, 1.0, 1.0)
        s = dg.Solver(c1, m)
        s.run(500, cfl=0.25, graph=True)
        assert np.array_equal(batch.case_q(k), s.get_q()), k
        assert batch.case_time(k) == s.time(), k
        del s


@pytest.fixture(scope="module")
def sens_full(dg):
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.4, 2.0, 1.0, 1.0)
    s = dg.Sensitivity(ctx, 256, 128, 1.08, 20.0, 8, 8, 0.4, 2.0)
    bumps = np.random.default_rng(2024).uniform(-1e-3, 1e-3, 16)
    cl = s.record(bumps, nsteps=400, cfl=1.0)
    cl_g, da, db = s.gradient(ckpt_every=20)
    return dict(s=s, ctx=ctx, bumps=bumps, cl=cl, cl_g=cl_g, g=np.concatenate([[da], db]))


def _cl_at(sf, i, x):
    s, b = sf["s"], sf["bumps"]
    if i == 0:
        return s.eval(2.0 + x, b)
    bb = b.copy()
    bb[i - 1] += x
    return s.eval(2.0, bb)


def test_adjoint_256x128_matches_tangent_linear(sens_full):
    sf = sens_full
    s, g = sf["s"], sf["g"]
    assert sf["cl_g"] == sf["cl"]  # the gradient's forward is the recorded run
    tl = np.zeros(17)
    _, tl[0] = s.tangent(1.0, None)
    for i in range(16):
        e = np.zeros(16)
        e[i] = 1.0
        _, tl[i + 1] = s.tangent(0.0, e)
    rel = np.abs(tl - g) / np.abs(g)
    print(f"adjoint vs tangent-linear per component: max {rel.max():.3e} {rel.tolist()}", flush=True)
    assert np.all(rel <= 1e-7), rel


def test_adjoint_256x128_matches_finite_differences(sens_full):
    """Finite differences with the recorded dt sequence.  C_L carries
    ~2e-14 of rounding noise from one perturbed forward run to the next, so
    the plain central difference at h = 1e-6 resolves a derivative to
    ~1e-8 absolute (printed, not asserted per component); the fourth-order
    central difference at h = 1e-5 (truncation O(h^4), noise ~3e-9
    absolute) is held to 1e-6 relative on every component."""
    sf = sens_full
    g = sf["g"]
    h = 1e-6
    fd2 = np.array([(_cl_at(sf, i, h) - _cl_at(sf, i, -h)) / (2 * h) for i in range(17)])
    h4 = 1e-5
    fd4 = np.array([(-_cl_at(sf, i, 2 * h4) + 8 * _cl_at(sf, i, h4) - 8 * _cl_at(sf, i, -h4)
                     + _cl_at(sf, i, -2 * h4)) / (12 * h4) for i in range(17)])
    r2 = np.abs(fd2 - g) / np.abs(g)
    r4 = np.abs(fd4 - g) / np.abs(g)
    print(f"central h=1e-6: max rel {r2.max():.3e} (abs {np.abs(fd2 - g).max():.3e}); "
          f"4th order h=1e-5: max rel {r4.max():.3e}; per component {r4.tolist()}", flush=True)
    assert np.all(r4 <= 1e-6), r4
    assert np.abs(fd2 - g).max() <= 5e-8
    assert np.linalg.norm(fd2 - g) <= 1e-6 * np.linalg.norm(g)
