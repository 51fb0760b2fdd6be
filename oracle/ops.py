"""Oracle: nodal DG operators on the reference triangle (numpy, N-generic).

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this file;
only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use it, as the checker.

Restates the operator tables the reference declares as constants
(``src/constants/constant_{r,s,Dr,Ds,Drw,Dsw,LIFT}.h``, ``FMASK`` at
``src/constants/constant_util.h:15``) from the textbook construction of
Hesthaven & Warburton, *Nodal Discontinuous Galerkin Methods* (2008):

* ``nodes2d``       warp & blend nodes with the alpha-optimised table (sec. 6.1)
* ``vandermonde2d`` orthonormal Proriol-Koornwinder-Dubiner basis (sec. 6.1)
* ``Dr = Vr V^-1``, ``Ds = Vs V^-1`` (sec. 6.2)
* ``Drw = (V Vr^T)(V V^T)^-1``, ``Dsw`` likewise: weak derivative M^-1 S^T
* ``LIFT = V (V^T E)`` with E the face mass matrices (sec. 6.2)
* ``FMASK``: face 0 is s=-1 (ascending r), face 1 is r+s=0 (ascending node
  index), face 2 is r=-1 listed in reverse so that every face is traversed
  counter-clockwise, which is the order the reference's table uses.

Parity: pinned at N=4 against the reference's literal tables
(``tests/golden/n4_tables.json``) and at every N by the structural identities
``Dr 1 = 0``, ``Dr r = 1``, ``Ds s = 1``, exactness on polynomials of degree N,
and the LIFT/face-mass identity.
"""
import math

import numpy as np

ALPHA_OPT = [0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
             1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258]


def jacobi_p(x, alpha, beta, n):
    """Orthonormal Jacobi polynomial P_n^{(alpha,beta)} evaluated at x."""
    x = np.asarray(x, dtype=np.float64)
    pl = np.zeros((n + 1,) + x.shape)
    gamma0 = (2 ** (alpha + beta + 1) / (alpha + beta + 1)
              * math.gamma(alpha + 1) * math.gamma(beta + 1) / math.gamma(alpha + beta + 1))
    pl[0] = 1.0 / math.sqrt(gamma0)
    if n == 0:
        return pl[0]
    gamma1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * gamma0
    pl[1] = ((alpha + beta + 2) * x / 2 + (alpha - beta) / 2) / math.sqrt(gamma1)
    if n == 1:
        return pl[1]
    aold = 2 / (2 + alpha + beta) * math.sqrt((alpha + 1) * (beta + 1) / (alpha + beta + 3))
    for i in range(1, n):
        h1 = 2 * i + alpha + beta
        anew = 2 / (h1 + 2) * math.sqrt((i + 1) * (i + 1 + alpha + beta) * (i + 1 + alpha)
                                        * (i + 1 + beta) / (h1 + 1) / (h1 + 3))
        bnew = -(alpha ** 2 - beta ** 2) / h1 / (h1 + 2)
        pl[i + 1] = 1 / anew * (-aold * pl[i - 1] + (x - bnew) * pl[i])
        aold = anew
    return pl[n]


def grad_jacobi_p(x, alpha, beta, n):
    if n == 0:
        return np.zeros_like(np.asarray(x, dtype=np.float64))
    return math.sqrt(n * (n + alpha + beta + 1)) * jacobi_p(x, alpha + 1, beta + 1, n - 1)


def jacobi_gq(alpha, beta, n):
    """Gauss quadrature nodes/weights for the Jacobi weight (Golub-Welsch)."""
    if n == 0:
        return np.array([-(alpha - beta) / (alpha + beta + 2)]), np.array([2.0])
    h1 = 2 * np.arange(n + 1) + alpha + beta
    main = -0.5 * (alpha ** 2 - beta ** 2) / (h1 + 2) / h1
    k = np.arange(1, n + 1)
    off = 2 / (h1[:n] + 2) * np.sqrt(k * (k + alpha + beta) * (k + alpha) * (k + beta)
                                     / (h1[:n] + 1) / (h1[:n] + 3))
    if alpha + beta < 10 * np.finfo(float).eps:
        main[0] = 0.0
    J = np.diag(main) + np.diag(off, 1) + np.diag(off, -1)
    x, v = np.linalg.eigh(J)
    w = v[0, :] ** 2 * 2 ** (alpha + beta + 1) / (alpha + beta + 1) \
        * math.gamma(alpha + 1) * math.gamma(beta + 1) / math.gamma(alpha + beta + 1)
    return x, w


def jacobi_gl(alpha, beta, n):
    """Gauss-Lobatto nodes for the Jacobi weight."""
    if n == 1:
        return np.array([-1.0, 1.0])
    xint, _ = jacobi_gq(alpha + 1, beta + 1, n - 2)
    return np.concatenate([[-1.0], xint, [1.0]])


def vandermonde1d(n, r):
    r = np.asarray(r, dtype=np.float64)
    return np.stack([jacobi_p(r, 0, 0, j) for j in range(n + 1)], axis=1)


def warp_factor(n, rout):
    lglr = jacobi_gl(0, 0, n)
    req = np.linspace(-1, 1, n + 1)
    veq = vandermonde1d(n, req)
    pmat = np.stack([jacobi_p(rout, 0, 0, i) for i in range(n + 1)], axis=0)
    lmat = np.linalg.solve(veq.T, pmat)
    warp = lmat.T @ (lglr - req)
    zerof = (np.abs(rout) < 1.0 - 1.0e-10).astype(np.float64)
    sf = 1.0 - (zerof * rout) ** 2
    return warp / sf + warp * (zerof - 1)


def nodes2d(n):
    """Warp & blend nodes on the equilateral triangle, then mapped to (r, s)."""
    alpha = ALPHA_OPT[n - 1] if n < 16 else 5.0 / 3.0
    L1, L3 = [], []
    for i in range(1, n + 2):
        for j in range(1, n + 3 - i):
            L1.append((i - 1) / n)
            L3.append((j - 1) / n)
    L1 = np.array(L1)
    L3 = np.array(L3)
    L2 = 1.0 - L1 - L3
    x = -L2 + L3
    y = (-L2 - L3 + 2 * L1) / math.sqrt(3.0)
    blend1 = 4 * L2 * L3
    blend2 = 4 * L1 * L3
    blend3 = 4 * L1 * L2
    warpf1 = warp_factor(n, L3 - L2)
    warpf2 = warp_factor(n, L1 - L3)
    warpf3 = warp_factor(n, L2 - L1)
    warp1 = blend1 * warpf1 * (1 + (alpha * L1) ** 2)
    warp2 = blend2 * warpf2 * (1 + (alpha * L2) ** 2)
    warp3 = blend3 * warpf3 * (1 + (alpha * L3) ** 2)
    x = x + warp1 + math.cos(2 * math.pi / 3) * warp2 + math.cos(4 * math.pi / 3) * warp3
    y = y + math.sin(2 * math.pi / 3) * warp2 + math.sin(4 * math.pi / 3) * warp3
    # equilateral (x, y) -> reference (r, s)
    l1 = (math.sqrt(3.0) * y + 1.0) / 3.0
    l2 = (-3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    l3 = (3.0 * x - math.sqrt(3.0) * y + 2.0) / 6.0
    r = -l2 + l3 - l1
    s = -l2 - l3 + l1
    return r, s


def rs_to_ab(r, s):
    a = np.where(np.abs(s - 1.0) > 1e-14, 2 * (1 + r) / np.where(s != 1.0, 1 - s, 1.0) - 1, -1.0)
    return a, s.copy()


def simplex2dp(a, b, i, j):
    h1 = jacobi_p(a, 0, 0, i)
    h2 = jacobi_p(b, 2 * i + 1, 0, j)
    return math.sqrt(2.0) * h1 * h2 * (1 - b) ** i


def grad_simplex2dp(a, b, idx, jdx):
    fa = jacobi_p(a, 0, 0, idx)
    dfa = grad_jacobi_p(a, 0, 0, idx)
    gb = jacobi_p(b, 2 * idx + 1, 0, jdx)
    dgb = grad_jacobi_p(b, 2 * idx + 1, 0, jdx)
    dmodedr = dfa * gb
    if idx > 0:
        dmodedr = dmodedr * (0.5 * (1 - b)) ** (idx - 1)
    dmodeds = dfa * (gb * (0.5 * (1 + a)))
    if idx > 0:
        dmodeds = dmodeds * (0.5 * (1 - b)) ** (idx - 1)
    tmp = dgb * (0.5 * (1 - b)) ** idx
    if idx > 0:
        tmp = tmp - 0.5 * idx * gb * (0.5 * (1 - b)) ** (idx - 1)
    dmodeds = dmodeds + fa * tmp
    dmodedr = dmodedr * 2 ** (idx + 0.5)
    dmodeds = dmodeds * 2 ** (idx + 0.5)
    return dmodedr, dmodeds


def vandermonde2d(n, r, s):
    a, b = rs_to_ab(r, s)
    cols = []
    for i in range(n + 1):
        for j in range(n - i + 1):
            cols.append(simplex2dp(a, b, i, j))
    return np.stack(cols, axis=1)


def grad_vandermonde2d(n, r, s):
    a, b = rs_to_ab(r, s)
    vr, vs = [], []
    for i in range(n + 1):
        for j in range(n - i + 1):
            dr, ds = grad_simplex2dp(a, b, i, j)
            vr.append(dr)
            vs.append(ds)
    return np.stack(vr, axis=1), np.stack(vs, axis=1)


def fmask(n, r, s, tol=1e-12):
    f0 = np.where(np.abs(s + 1) < tol)[0]
    f1 = np.where(np.abs(r + s) < tol)[0]
    f2 = np.where(np.abs(r + 1) < tol)[0][::-1]
    return np.concatenate([f0, f1, f2]).astype(np.int32)


def build_operators(n):
    """Return a dict of the reference-named constants at order n (row-major)."""
    if n < 1:
        raise ValueError("order must be >= 1")
    np_ = (n + 1) * (n + 2) // 2
    nfp = n + 1
    r, s = nodes2d(n)
    V = vandermonde2d(n, r, s)
    Vr, Vs = grad_vandermonde2d(n, r, s)
    Vinv = np.linalg.inv(V)
    Dr = Vr @ Vinv
    Ds = Vs @ Vinv
    VVt_inv = np.linalg.inv(V @ V.T)
    Drw = (V @ Vr.T) @ VVt_inv
    Dsw = (V @ Vs.T) @ VVt_inv
    fm = fmask(n, r, s)
    assert fm.size == 3 * nfp, fm
    E = np.zeros((np_, 3 * nfp))
    for f in range(3):
        idx = fm[f * nfp:(f + 1) * nfp]
        coord = s[idx] if f == 2 else r[idx]
        v1 = vandermonde1d(n, coord)
        E[idx, f * nfp:(f + 1) * nfp] = np.linalg.inv(v1 @ v1.T)
    LIFT = V @ (V.T @ E)
    return {"N": n, "Np": np_, "Nfp": nfp, "r": r, "s": s, "Dr": Dr, "Ds": Ds,
            "Drw": Drw, "Dsw": Dsw, "LIFT": LIFT, "FMASK": fm, "V": V,
            "MassMatrix": np.linalg.inv(V @ V.T)}


def face_mass(n, ops=None):
    """1D face mass matrix for one face (edge of length 2 in reference coords)."""
    ops = ops or build_operators(n)
    r = ops["r"][ops["FMASK"][:ops["Nfp"]]]
    v1 = vandermonde1d(n, r)
    return np.linalg.inv(v1 @ v1.T)
