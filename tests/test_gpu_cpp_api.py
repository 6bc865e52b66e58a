"""The C++ mirror of the reference API (include/cpolyg.hpp) drives the GPU path from a
plain g++ program; the program checks itself against the oracle and prints PASS."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_cpp_api_program():
    exe = os.path.join(HERE, "cpp", "test_api")
    assert os.path.exists(exe), "tests/cpp/test_api not built (run __graft_entry__.build())"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout
