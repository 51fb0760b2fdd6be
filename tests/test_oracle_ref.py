"""The oracle's restatement against the reference's own kernels at N=4.

tests/golden/ref_n4_vectors.npz holds outputs of the reference's BLAS-free
kernel headers (src/get_neighbour_q.h, get_bedge_q.h, set_workingQ.h,
update_Q.h, calc_dt.h, set_ic.h, neighbour_zero.h) compiled from the
reference tree (tests/golden/make_ref_vectors.py).  The oracle follows the
same operation order, so agreement is bit-for-bit.  When oracle/_ref is
built, the live reference library is compared as well."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle
from oracle.ops import build_operators

VEC = os.path.join(os.path.dirname(__file__), "golden", "ref_n4_vectors.npz")


@pytest.fixture(scope="module")
def vec():
    return dict(np.load(VEC))


@pytest.fixture(scope="module")
def orc4(vec):
    ph = dict(zip(("gam", "bc_mach", "bc_alpha", "bc_p", "bc_r", "bc_u", "bc_v", "bc_e"), vec["phys"]))
    return oracle.Oracle(build_operators(4), ph)


def test_fmask_matches(vec, orc4):
    assert np.array_equal(vec["FMASK"], orc4.ops["FMASK"])


def test_get_neighbour_q(vec, orc4):
    for k in range(vec["gn_en"].shape[0]):
        a, b = orc4.get_neighbour_q(vec["gn_en"][k], vec["gn_xL"][k], vec["gn_yL"][k], vec["gn_xR"][k],
                                    vec["gn_yR"][k], vec["gn_qL"][k], vec["gn_qR"][k], vec["gn_exL0"][k],
                                    vec["gn_exR0"][k])
        assert np.array_equal(a, vec["gn_exL"][k]) and np.array_equal(b, vec["gn_exR"][k]), k


def test_get_bedge_q(vec, orc4):
    for k in range(vec["gb_type"].shape[0]):
        ex = orc4.get_bedge_q(vec["gb_type"][k], vec["gb_num"][k], vec["gb_nx"][k], vec["gb_ny"][k],
                              vec["gb_q"][k], vec["gb_ex0"][k])
        assert np.array_equal(ex, vec["gb_ex"][k]), k


def test_set_workingQ_and_update_Q(vec, orc4):
    for k in range(vec["sw_dt"].size):
        dt = vec["sw_dt"][k]
        q, k1, k2, k3 = vec["sw_q"][k], vec["sw_k1"][k], vec["sw_k2"][k], vec["sw_k3"][k]
        assert np.array_equal(orc4.set_workingQ(dt, 0, q, k1, k2), vec["sw_w0"][k])
        assert np.array_equal(orc4.set_workingQ(dt, 1, q, k1, k2), vec["sw_w1"][k])
        qn, wn = orc4.update_Q(dt, q, k1, k2, k3)
        assert np.array_equal(qn, vec["uq_q"][k]) and np.array_equal(wn, vec["uq_w"][k])


def test_calc_dt(vec, orc4):
    for k in range(vec["cd_in"].size):
        assert orc4.calc_dt(vec["cd_q"][k], vec["cd_fs"][k], vec["cd_in"][k]) == vec["cd_out"][k], k


def test_set_ic(vec, orc4):
    q, w = orc4.set_ic()
    assert np.array_equal(q, vec["ic_q"]) and np.array_equal(w, vec["ic_w"])
    assert np.all(vec["nz"] == 0.0)


def test_live_reference_library_matches_vectors(vec):
    ref = oracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref is not built (needs the reference tree)")
    D = ctypes.c_double
    ref.ref_set_physics(*[D(v) for v in vec["phys"]])
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    for k in range(vec["cd_in"].size):
        v = np.array([vec["cd_in"][k]])
        ref.ref_calc_dt(P(vec["cd_q"][k].copy()), P(vec["cd_fs"][k].copy()), P(v))
        assert v[0] == vec["cd_out"][k]
    for k in range(vec["gb_type"].shape[0]):
        a = vec["gb_ex0"][k].copy()
        ref.ref_get_bedge_q(P(vec["gb_type"][k:k + 1].copy()), P(vec["gb_num"][k:k + 1].copy()),
                            P(vec["gb_nx"][k].copy()), P(vec["gb_ny"][k].copy()), P(vec["gb_q"][k].copy()), P(a))
        assert np.array_equal(a, vec["gb_ex"][k])
