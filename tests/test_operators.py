"""DG operator tables: both generators pinned against the reference's literal
order-4 tables (src/constants/*.h, extracted to tests/golden/n4_tables.json)
and, at every order, against identities the mathematics fixes."""
import json
import os

import numpy as np
import pytest

from oracle.ops import build_operators

GOLD = os.path.join(os.path.dirname(__file__), "golden", "n4_tables.json")
NAMES = ("r", "s", "Dr", "Ds", "Drw", "Dsw", "LIFT")


@pytest.fixture(scope="module")
def n4():
    return json.load(open(GOLD))


def test_oracle_generator_matches_reference_tables(n4):
    o = build_operators(4)
    for k in NAMES:
        assert np.abs(np.array(n4[k]) - o[k].ravel()).max() < 3e-14, k
    assert list(o["FMASK"]) == n4["FMASK"]


def test_library_generator_matches_reference_tables(dg, n4):
    o = dg.generate_operators(4)
    for k in NAMES:
        assert np.abs(np.array(n4[k]) - o[k].ravel()).max() < 3e-14, k
    assert list(o["FMASK"]) == n4["FMASK"]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_library_and_oracle_generators_agree(dg, n):
    a, b = dg.generate_operators(n), build_operators(n)
    for k in NAMES:
        assert np.abs(a[k] - b[k]).max() < 5e-13, (n, k)
    assert np.array_equal(a["FMASK"], b["FMASK"])


def test_order3_face_mask(dg):
    assert list(dg.generate_operators(3)["FMASK"]) == [0, 1, 2, 3, 3, 6, 8, 9, 9, 7, 4, 0]


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_operator_identities(dg, n):
    o = dg.generate_operators(n)
    r, s, Dr, Ds = o["r"], o["s"], o["Dr"], o["Ds"]
    M, Drw, Dsw, LIFT = o["mass"], o["Drw"], o["Dsw"], o["LIFT"]
    # exact differentiation of every monomial of degree <= N
    for a in range(n + 1):
        for b in range(n + 1 - a):
            f = r ** a * s ** b
            dfr = a * r ** max(a - 1, 0) * s ** b if a else 0 * r
            dfs = b * r ** a * s ** max(b - 1, 0) if b else 0 * s
            assert np.abs(Dr @ f - dfr).max() < 1e-11
            assert np.abs(Ds @ f - dfs).max() < 1e-11
    # mass matrix integrates 1 over the reference triangle (area 2)
    one = np.ones_like(r)
    assert abs(one @ M @ one - 2.0) < 1e-13
    # weak derivative: M Drw = Dr^T M
    assert np.abs(M @ Drw - Dr.T @ M).max() < 1e-11
    assert np.abs(M @ Dsw - Ds.T @ M).max() < 1e-11
    # lift: M LIFT = face mass matrices scattered to the face nodes
    nfp = n + 1
    E = np.zeros_like(LIFT)
    fm = o["FMASK"]
    for f in range(3):
        for i in range(nfp):
            for j in range(nfp):
                E[fm[f * nfp + i], f * nfp + j] = o["face_mass"][i, j]
    assert np.abs(M @ LIFT - E).max() < 1e-11
    # face mass integrates 1 over a face of reference length 2
    assert abs(np.ones(nfp) @ o["face_mass"] @ np.ones(nfp) - 2.0) < 1e-13
