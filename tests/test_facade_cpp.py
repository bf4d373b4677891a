"""The header-only C++ facade (include/conicxx_amd.hpp) driven from C++.

tests/cpp/facade_qp.cpp mirrors how the reference's own tests call
conicxx::Solver (test/solver/test_solve_qp.cpp:9-66,
test/solver/test_update_reuse.cpp:32-73); conicxx_amd/build.py compiles it.
"""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_qp")


def run():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run the build first")
    return subprocess.run([BIN], capture_output=True, text=True, timeout=120)


def test_facade_links_and_reports_defaults():
    r = run()
    assert "PASS Settings defaults equal cx_default_settings" in r.stdout, r.stdout + r.stderr
    assert "FAIL" not in r.stdout, r.stdout
    assert r.returncode in (0, 3), r.stdout + r.stderr


@pytest.mark.gpu
def test_facade_solves_reference_known_answers():
    r = run()
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("PASS") >= 11
