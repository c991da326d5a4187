"""Sanitizer builds on the CPU: the library's host-only algorithms
(fortranfvcell_amd/csrc/fvc_host.h: workspace carving, the transposed-gather
index, batch block allocation) and the C oracle (every SpMV kind, the dots,
getres and every branch of the solver dispatch), each compiled with
AddressSanitizer + UndefinedBehaviorSanitizer (leak checking included) and
run; any report fails the test.  Device code is not built here: these
sanitizers instrument host code only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
ENV.pop("LD_PRELOAD", None)


def _build_and_run(tmp_path, compiler, args, name):
    if shutil.which(compiler) is None:
        pytest.skip(f"{compiler} not found")
    exe = str(tmp_path / name)
    b = subprocess.run([compiler, *SAN, "-Wall", "-o", exe, *args], capture_output=True, text=True)
    if b.returncode != 0 and "asan" in b.stderr.lower() and "cannot find" in b.stderr.lower():
        pytest.skip("sanitizer runtime not installed")
    assert b.returncode == 0, b.stderr
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "runtime error" not in out and "Sanitizer" not in out, out
    return out


def test_host_algorithms_sanitized(tmp_path):
    out = _build_and_run(tmp_path, "g++", ["-std=c++17", os.path.join(HERE, "sanitize", "host_checks.cpp")],
                         "host_checks")
    assert "0 failures" in out


def test_oracle_sanitized(tmp_path):
    out = _build_and_run(tmp_path, "gcc", ["-std=c11", "-ffp-contract=off",
                                           os.path.join(HERE, "sanitize", "oracle_checks.c"),
                                           os.path.join(ROOT, "oracle", "fvc_oracle.c"), "-lm"], "oracle_checks")
    assert "0 failures" in out
