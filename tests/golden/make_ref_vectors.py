"""Golden vectors from the reference's own kernels at its order N=4.

Runs the reference's BLAS-free per-element kernel headers, compiled from the
reference tree into oracle/_ref/libref.so by ``make -C oracle ref``
(get_neighbour_q, get_bedge_q, set_workingQ, update_Q, calc_dt, set_ic,
neighbour_zero), on seeded random valid inputs and stores inputs and
outputs in ``ref_n4_vectors.npz`` next to this script.

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle  # noqa: E402

NP, NFP, NF = 15, 5, 15


def P(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def random_state(rng, n):
    """Valid conserved states: rho in [0.5, 2], |u| < 1, p in [0.5, 2]."""
    rho = rng.uniform(0.5, 2.0, n)
    u = rng.uniform(-0.8, 0.8, n)
    v = rng.uniform(-0.8, 0.8, n)
    p = rng.uniform(0.5, 2.0, n)
    E = p / 0.4 + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E], axis=-1)


def main():
    ref = oracle.ref_lib()
    if ref is None:
        sys.exit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    rng = np.random.default_rng(20241203)
    phys = oracle.freestream(gam=1.4, mach=0.4, alpha=3.0)
    D = ctypes.c_double
    ref.ref_set_physics(*[D(phys[k]) for k in ("gam", "bc_mach", "bc_alpha", "bc_p", "bc_r", "bc_u",
                                               "bc_v", "bc_e")])
    out = {"phys": np.array([phys[k] for k in ("gam", "bc_mach", "bc_alpha", "bc_p", "bc_r", "bc_u",
                                               "bc_v", "bc_e")])}
    fm = np.zeros(15, np.int32)
    ref.ref_get_fmask(P(fm))
    out["FMASK"] = fm

    # get_neighbour_q: random edge numbers and vertex coordinates; orientation
    # by the coordinate test, both outcomes represented
    n = 64
    en = rng.integers(0, 3, (n, 2)).astype(np.int32)
    xL, yL = rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3))
    xR, yR = rng.uniform(-1, 1, (n, 3)), rng.uniform(-1, 1, (n, 3))
    same = rng.random(n) < 0.5
    for k in range(n):
        if same[k]:
            xR[k, en[k, 1]] = xL[k, en[k, 0]]
            yR[k, en[k, 1]] = yL[k, en[k, 0]]
    qL, qR = random_state(rng, (n, NP)), random_state(rng, (n, NP))
    exL0, exR0 = rng.normal(size=(n, NF * 4)), rng.normal(size=(n, NF * 4))
    exL, exR = exL0.copy(), exR0.copy()
    for k in range(n):
        a, b = exL[k].copy(), exR[k].copy()
        ref.ref_get_neighbour_q(P(en[k].copy()), P(xL[k].copy()), P(yL[k].copy()), P(xR[k].copy()),
                                P(yR[k].copy()), P(qL[k].copy()), P(qR[k].copy()), P(a), P(b))
        exL[k], exR[k] = a, b
    out.update(gn_en=en, gn_xL=xL, gn_yL=yL, gn_xR=xR, gn_yR=yR, gn_qL=qL, gn_qR=qR, gn_exL0=exL0,
               gn_exR0=exR0, gn_exL=exL, gn_exR=exR)

    # get_bedge_q: every type 0..2 and every local edge
    m = 36
    bt = np.repeat(np.arange(3, dtype=np.int32), 12)
    bn = np.tile(np.arange(3, dtype=np.int32), 12)
    ang = rng.uniform(0, 2 * np.pi, (m, NF))
    nx, ny = np.cos(ang), np.sin(ang)
    q = random_state(rng, (m, NP))
    ex0 = rng.normal(size=(m, NF * 4))
    ex = ex0.copy()
    for k in range(m):
        a = ex[k].copy()
        ref.ref_get_bedge_q(P(bt[k:k + 1].copy()), P(bn[k:k + 1].copy()), P(nx[k].copy()), P(ny[k].copy()),
                            P(q[k].copy()), P(a))
        ex[k] = a
    out.update(gb_type=bt, gb_num=bn, gb_nx=nx, gb_ny=ny, gb_q=q, gb_ex0=ex0, gb_ex=ex)

    # set_workingQ (both stages) and update_Q
    s = 8
    qs, k1, k2, k3 = [rng.normal(size=(s, NP * 4)) for _ in range(4)]
    dts = rng.uniform(1e-4, 1e-1, s)
    w0, w1 = np.zeros((s, NP * 4)), np.zeros((s, NP * 4))
    qn, wn = qs.copy(), np.zeros((s, NP * 4))
    for k in range(s):
        dt = np.array([dts[k]])
        for stage, w in ((0, w0), (1, w1)):
            a = np.zeros(NP * 4)
            ref.ref_set_workingQ(P(dt), P(np.array([stage], np.int32)), P(qs[k].copy()), P(k1[k].copy()),
                                 P(k2[k].copy()), P(a))
            w[k] = a
        a, b = qn[k].copy(), np.zeros(NP * 4)
        ref.ref_update_Q(P(dt), P(a), P(k1[k].copy()), P(k2[k].copy()), P(k3[k].copy()), P(b))
        qn[k], wn[k] = a, b
    out.update(sw_q=qs, sw_k1=k1, sw_k2=k2, sw_k3=k3, sw_dt=dts, sw_w0=w0, sw_w1=w1, uq_q=qn, uq_w=wn)

    # calc_dt with running max
    c = 32
    qc = random_state(rng, (c, NP))
    fs = rng.uniform(1.0, 50.0, (c, NF))
    dt_in = rng.uniform(0.0, 40.0, c)
    dt_out = np.zeros(c)
    for k in range(c):
        v = np.array([dt_in[k]])
        ref.ref_calc_dt(P(qc[k].copy()), P(fs[k].copy()), P(v))
        dt_out[k] = v[0]
    out.update(cd_q=qc, cd_fs=fs, cd_in=dt_in, cd_out=dt_out)

    # set_ic and neighbour_zero
    q1, w1_ = np.zeros(NP * 4), np.zeros(NP * 4)
    ref.ref_set_ic(P(q1), P(w1_))
    z = rng.normal(size=NF * 4)
    ref.ref_neighbour_zero(P(z))
    out.update(ic_q=q1, ic_w=w1_, nz=z)
    np.savez_compressed(os.path.join(HERE, "ref_n4_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_n4_vectors.npz"))


if __name__ == "__main__":
    main()
