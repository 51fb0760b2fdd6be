"""The library's host C++ (mesh generators, legacy-VTK reader, edge and
face-map builders, operator generator) under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/asan/host_checks.cpp built by
tests/asan/Makefile with g++ -fsanitize=address,undefined, run on the
reference airfoil mesh fixture and on faulty meshes it writes itself."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_host_code_under_address_and_ub_sanitizers(naca_vtk, tmp_path):
    if shutil.which(os.environ.get("CXX", "g++")) is None:
        pytest.skip("no host C++ compiler")
    d = os.path.join(HERE, "asan")
    subprocess.run(["make", "-s", "-C", d], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(d, "host_checks"), str(tmp_path), naca_vtk], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host checks ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
