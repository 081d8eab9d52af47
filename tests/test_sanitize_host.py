"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

Builds tests/cpp/sanitize_host.cpp together with the C oracle and the CpuFFT
header with -fsanitize=address,undefined (no recovery) and runs it: any
out-of-bounds access, leak or undefined operation aborts with a report, wrong
values exit non-zero.  No GPU code is built here.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="gcc/g++ not found")
def test_host_code_under_asan_ubsan(tmp_path):
    san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
    obj = tmp_path / "spx_oracle.o"
    exe = tmp_path / "sanitize_host"
    subprocess.run(["gcc", "-std=gnu99", *san, "-c", os.path.join(ROOT, "oracle", "spx_oracle.c"), "-o", str(obj)],
                   check=True)
    subprocess.run(["g++", "-std=c++17", *san, "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp"), str(obj), "-o", str(exe), "-lm"],
                   check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert "sanitized host checks passed" in r.stdout
