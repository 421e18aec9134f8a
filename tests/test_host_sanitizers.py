"""The host side of the C ABI under ThreadSanitizer and AddressSanitizer.

The runtime's host sources (plan cache, tune-file schedules, table caches,
pool, server, batch planner, entry points) are compiled with a host
sanitizer -- the device code is not instrumented -- and linked with the
library's kernel objects into tests/cpp/host_threads_test, which hammers the
plan cache and entry points from eight threads.  Runs on the CPU (without a
device the device calls must fail cleanly); the program exits non-zero on a
sanitizer report or a failed check.
"""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "goldilocks_amd", "csrc")
BUILD = os.path.join(ROOT, "goldilocks_amd", "_build")
OUT = os.path.join(ROOT, "tests", "cpp", "_san")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANGXX = "/opt/rocm/llvm/bin/clang++"
INSTRUMENTED = ["gl_runtime.cpp", "gl_api.cpp", "gl_hostio.cpp", "gl_server.cpp", "gl_batch.hip"]
PLAIN = ["gl_ntt_kernels.hip"] + [f"gl_variants_{k}.hip" for k in range(8)] + [
    "gl_rader.hip", "gl_k12.hip", "gl_permute.hip", "gl_noalloc.hip"]


def _build(kind):
    """Objects for sanitizer `kind` ("thread" / "address"); the host flag is
    placed behind -Xarch_host so the device code is built as usual."""
    os.makedirs(OUT, exist_ok=True)
    plain = [os.path.join(BUILD, s + ".o") for s in PLAIN]
    if not all(os.path.exists(p) for p in plain):
        pytest.skip("library objects not built (python -m goldilocks_amd.build)")

    def compile_one(src):
        obj = os.path.join(OUT, f"{kind}_{src}.o")
        cmd = [HIPCC, "-O1", "-g", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-Xarch_host", f"-fsanitize={kind}", "-fno-omit-frame-pointer",
               "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-x", "hip",
               "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    with ThreadPoolExecutor(4) as ex:
        objs = list(ex.map(compile_one, INSTRUMENTED))
    exe = os.path.join(OUT, f"host_threads_{kind}")
    # host-only link line (no device sources): the host compiler driver
    cmd = [CLANGXX, "-g", f"-fsanitize={kind}", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "host_threads_test.cpp")] + objs + plain + \
          ["-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-lpthread", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.mark.parametrize("kind", ["thread", "address"])
def test_host_threads_under_sanitizer(kind):
    exe = _build(kind)
    env = dict(os.environ)
    env["GOLDILOCKS_TUNE_FILE"] = os.path.join(ROOT, "goldilocks_amd", "tune_gfx950.txt")
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66"
    env["ASAN_OPTIONS"] = "detect_leaks=0 exitcode=66"
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "host_threads_test: ok" in r.stdout
