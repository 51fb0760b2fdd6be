"""dg_rhs_adjoint_noalloc (static per-GPU scratch, no allocation) against
dg_rhs_adjoint (caller's scratch) through the C ABI: the same kernels, so the
state and geometry cotangents must be the same bits.  Two calls in a row on
one GPU, the second on a smaller mesh (it reuses the scratch the first call
left behind), and a mesh above the scratch's capacity rejected with
DG_ERR_ARG before any device work."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GAM = 1.4
GKEYS = ("rx", "ry", "sx", "sy", "fscale", "nx", "ny", "bcu", "bcv")


def T(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)


def _case(dg, orc, dev, mesh, seed):
    ctx = dg.Context(3)
    ctx.set_freestream(GAM, 0.45, 3.0, 1.0, 1.0)
    ctx.decl_mesh(mesh)
    o = orc.Oracle(dg.generate_operators(3), orc.freestream(GAM, 0.45, 3.0))
    case = o.make_case(mesh.cells, mesh.cell_xy, mesh.edge2cells, mesh.edgeNum, mesh.bedge2cells,
                       mesh.bedgeNum, mesh.bedge_type)
    geo = {k: T(case["geo"][k], dev) for k in ("rx", "ry", "sx", "sy", "fscale", "nx", "ny")}
    ph = o.phys
    qinf = np.array([ph["bc_r"], ph["bc_r"] * ph["bc_u"], ph["bc_r"] * ph["bc_v"], ph["bc_e"]])
    rng = np.random.default_rng(seed)
    w = qinf * (1 + 0.03 * rng.normal(size=(mesh.ncells, ctx.Np, 4)))
    mu = rng.normal(size=w.shape)
    return ctx, geo, T(w, dev), T(mu, dev)


def _zeros(ctx, dev):
    K, Np, NF = ctx.ncells, ctx.Np, ctx.NF
    acc = torch.zeros(K * Np * 4, dtype=torch.float64, device=dev)
    g = {k: torch.zeros(K * (Np if k in ("rx", "ry", "sx", "sy") else NF), dtype=torch.float64, device=dev)
         for k in GKEYS}
    return acc, g


def test_noalloc_adjoint_matches_the_scratch_entry_point(dg, orc):
    dev = torch.device("cuda:0")
    meshes = [dg.naca0012_ogrid(48, 16, 1.15, 10.0), dg.periodic_grid(6, 5, seed=3)]
    assert meshes[1].ncells < meshes[0].ncells
    for i, mesh in enumerate(meshes):
        ctx, geo, w, mu = _case(dg, orc, dev, mesh, 11 + i)
        acc_a, g_a = _zeros(ctx, dev)
        ctx.rhs_adjoint(w, mu, 0.75, geo, acc_a, g_a)
        acc_b, g_b = _zeros(ctx, dev)
        ctx.rhs_adjoint_noalloc(w, mu, 0.75, geo, acc_b, g_b)
        torch.cuda.synchronize()
        print(f"call {i}: {mesh.ncells} cells, max |acc| {acc_a.abs().max().item():.3e}", flush=True)
        assert acc_a.abs().max().item() > 0
        assert torch.equal(acc_a, acc_b), i
        for k in GKEYS:
            assert torch.equal(g_a[k], g_b[k]), (i, k)


def test_noalloc_adjoint_rejects_meshes_beyond_its_scratch(dg):
    from importlib import import_module
    lib = import_module("airfoil-op2-fem_amd._lib").lib
    mesh = dg.periodic_grid(210, 210, seed=1)  # 88,200 cells: above 87,381 at N=3
    ctx = dg.Context(3)
    ctx.decl_mesh(mesh)
    rc = lib.dg_rhs_adjoint_noalloc(ctx._h, None, None, 1.0, *([None] * 17))
    assert rc == -1, rc  # DG_ERR_ARG
    msg = lib.dg_last_error(ctx._h).decode()
    assert "scratch" in msg and "87381" in msg, msg
