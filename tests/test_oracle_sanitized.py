"""Runs the oracle's own tests against a build of the oracle with
AddressSanitizer and UndefinedBehaviorSanitizer (``make -C oracle asan``), so an
out-of-bounds access or undefined arithmetic in the checker's C++ fails the
suite instead of silently corrupting the vectors the GPU results are compared
with.  CPU only: the sanitizers instrument host code.

The sanitized library is loaded in a child interpreter with the sanitizer
runtimes preloaded (Python itself is not instrumented; leak checking is off
because the interpreter keeps allocations alive at exit)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "oracle", "_asan", "liboracle_acoustic2d.so")


def _runtime(name):
    out = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_suite_under_address_and_ub_sanitizers():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("g++ sanitizer runtimes not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    env = dict(os.environ, LD_PRELOAD=f"{asan} {ubsan}", ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1", GBW_ORACLE_LIB=ASAN_LIB,
               PYTHONDONTWRITEBYTECODE="1")
    probe = ("import oracle, sys; oracle._load(); "
             "sys.exit(0 if oracle._lib._name == sys.argv[1] else 3)")
    subprocess.run([sys.executable, "-c", probe, ASAN_LIB], cwd=ROOT, env=env, check=True)
    # the two config-1 finite-difference tests are left out for time (each runs
    # 32 full forwards); every other oracle test runs instrumented
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                          "tests/test_oracle.py", "tests/test_oracle3d.py", "-k", "not config1"],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (res.stdout + res.stderr)[-3000:]
    assert res.returncode == 0, tail
    assert "passed" in res.stdout and "ERROR: AddressSanitizer" not in tail, tail
