"""Batched sweep and hipGraph replay on the GPU.

Bar: bit-identical.  A case inside a batch runs the same per-cell arithmetic
with the same dt as the case alone, and a graph replays the same launches."""
import math

import numpy as np
import pytest

pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GAM = 1.4


def single(dg, mesh, mach, alpha):
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, mach, alpha, 1.0, 1.0)
    return dg.Solver(ctx, mesh)


def test_batch_cases_bit_identical_to_single_runs(dg):
    specs = [(24, 6, 0.35, -2.0), (40, 10, 0.5, 4.0), (32, 8, 0.45, 7.5)]
    meshes = [dg.naca0012_ogrid(nc, nr, 1.15, 15.0) for nc, nr, _, _ in specs]
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.4, 0.0, 1.0, 1.0)
    batch = dg.BatchSolver(ctx, meshes, [s[2] for s in specs], [s[3] for s in specs])
    batch.run(21, cfl=1.0, graph=True)
    assert batch.check() == 0
    for k, (m, (_, _, mach, alpha)) in enumerate(zip(meshes, specs)):
        s = single(dg, m, mach, alpha)
        s.run(21, cfl=1.0)
        assert np.array_equal(batch.case_q(k), s.get_q()), k
        assert batch.case_time(k) == s.time(), k
        assert batch.case_forces(k)[2].tolist() == s.forces()[2].tolist(), k


def test_failed_case_does_not_reach_other_cases(dg):
    """A case whose state turns non-finite stays confined to itself: the
    other cases of the batch still match their single runs bit for bit
    (the stage kernel re-zeroes its padded operand columns on every tile)."""
    specs = [(24, 6, 0.4, 1.0), (16, 5, 0.45, 2.0), (40, 10, 0.5, -3.0), (20, 4, 0.3, 0.5)]
    meshes = [dg.naca0012_ogrid(nc, nr, 1.15, 15.0) for nc, nr, _, _ in specs]
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.4, 0.0, 1.0, 1.0)
    batch = dg.BatchSolver(ctx, meshes, [s[2] for s in specs], [s[3] for s in specs])
    q1 = batch.case_q(1)
    q1[len(q1) // 2, 3, 0] = np.nan
    batch.set_case_q(1, q1)
    batch.run(6, cfl=1.0)
    assert batch.check() != 0
    assert np.isnan(batch.case_q(1)).any()
    for k in (0, 2, 3):
        s = single(dg, meshes[k], specs[k][2], specs[k][3])
        s.run(6, cfl=1.0)
        assert np.array_equal(batch.case_q(k), s.get_q()), k


def _vortex(dg, n=24):
    mesh = dg.periodic_grid(n, n, 0.0, -5.0, 10.0, 10.0, seed=0, jitter=0.1)
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.5, 0.0, 1.0, 1.0)
    s = dg.Solver(ctx, mesh)
    s.set_vortex(0.5 * math.sqrt(GAM), 5.0, 5.0, 0.0, 10.0)
    return s


def test_launch_shapes_and_autotune_keep_results(dg):
    """Every launch shape of the matrix-core stage kernel (waves per
    workgroup, walk order, grid cap) gives the same bits, and the autotuner
    restores the state it tried its shapes from."""
    ref = _vortex(dg)
    ref.run(5, cfl=0.5)
    q_ref, t_ref = ref.get_q(), ref.time()
    for wpb, walk, cap in ((1, 1, 0), (2, 0, 0), (2, 1, 40), (4, 0, 0), (4, 1, 8), (1, 0, 3)):
        s = _vortex(dg)
        s.set_launch(wpb, walk, cap)
        s.run(5, cfl=0.5, graph=(wpb == 4))
        assert np.array_equal(s.get_q(), q_ref), (wpb, walk, cap)
        assert s.time() == t_ref
    s = _vortex(dg)
    q0 = s.get_q()
    tab = s.autotune(trial_steps=2, cfl=0.5)
    assert len(tab) == 18 and all(r["ms_per_step"] > 0 for r in tab)
    assert np.array_equal(s.get_q(), q0) and s.time()[0] == 0.0
    best = min(tab, key=lambda r: r["ms_per_step"])
    cfg = s.launch_config()
    assert (cfg["wpb"], cfg["walk"]) == (best["wpb"], best["walk"])
    s.run(5, cfl=0.5)
    assert np.array_equal(s.get_q(), q_ref)


def test_batch_launch_shapes_bit_identical(dg):
    specs = [(24, 6, 0.35, -2.0), (40, 10, 0.5, 4.0), (32, 8, 0.45, 7.5)]
    meshes = [dg.naca0012_ogrid(nc, nr, 1.15, 15.0) for nc, nr, _, _ in specs]
    out = []
    for wpb, walk in ((1, 0), (4, 1), (2, 0)):
        ctx = dg.Context(3)
        ctx.set_freestream(GAM, 0.4, 0.0, 1.0, 1.0)
        b = dg.BatchSolver(ctx, meshes, [s[2] for s in specs], [s[3] for s in specs])
        b.set_launch(wpb, walk, 0)
        b.run(7, cfl=1.0)
        out.append([b.case_q(k) for k in range(3)])
    for k in range(3):
        assert np.array_equal(out[0][k], out[1][k]) and np.array_equal(out[0][k], out[2][k]), k


def test_graph_replay_matches_direct_launches(dg):
    def vortex():
        mesh = dg.periodic_grid(24, 24, 0.0, -5.0, 10.0, 10.0, seed=0, jitter=0.1)
        ctx = dg.Context(3)
        ctx.set_freestream(GAM, 0.5, 0.0, 1.0, 1.0)
        s = dg.Solver(ctx, mesh)
        s.set_vortex(0.5 * math.sqrt(GAM), 5.0, 5.0, 0.0, 10.0)
        return s
    a, b = vortex(), vortex()
    a.run(15, cfl=0.5)
    b.run(1, cfl=0.5)
    b.run(14, cfl=0.5, graph=True)  # starts on an odd step
    assert np.array_equal(a.get_q(), b.get_q())
    assert a.time() == b.time()
    b.run(6, cfl=0.5, graph=True)   # replay of the cached graph
    a.run(6, cfl=0.5)
    assert np.array_equal(a.get_q(), b.get_q())
