"""Host build of the device field arithmetic and codelets.

tests/cpp/field_host_test.cpp includes the same headers the HIP kernels use
(gl_field.hpp, gl_dft.hpp), compiles them for the CPU and checks add, sub,
mul, mont_mul and every shift form against 128-bit closed forms, and the
register codelets against the oracle's naive DFT (ntt/src/ntt/naive.rs:3-25
restated in oracle/ntt.c)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and not shutil.which("hipcc"), reason="hipcc not installed")
def test_field_and_codelets_on_host(oracle, tmp_path):
    exe = str(tmp_path / "field_host_test")
    src = os.path.join(ROOT, "tests", "cpp", "field_host_test.cpp")
    odir = os.path.join(ROOT, "oracle")
    cmd = [HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-host-only", src,
           "-L", odir, "-loracle", f"-Wl,-rpath,{odir}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:]
