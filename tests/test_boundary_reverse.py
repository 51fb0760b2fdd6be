"""The hand-written reverse of the boundary states (csrc/dg_bc.h
boundary_vjp, used by the adjoint's face-pair pass) against the same states
in dual arithmetic, on the host: bp . (J dx) == (J^T bp) . dx to 1e-12
relative in every branch (types 0-3, the far field's four branches)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_boundary_reverse_matches_dual_derivatives(tmp_path):
    cxx = os.environ.get("CXX", "g++")
    if shutil.which(cxx) is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "bc_reverse_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-o", exe, os.path.join(HERE, "bc", "bc_reverse_check.cpp")],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bc reverse ok" in r.stdout
