"""Runs the C++ mirror's test program (tests/cpp/api_test.cpp): the
reference's own unit tests written against include/goldilocks.hpp."""
import os
import subprocess

import pytest

from gpu_util import note

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_api_program():
    exe = os.path.join(ROOT, "tests", "cpp", "api_test")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    note("cpp api_test")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout
