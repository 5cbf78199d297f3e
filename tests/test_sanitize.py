"""The library's host code under AddressSanitizer and UndefinedBehavior-
Sanitizer (tests/sanitize): random hierarchies committed in host-only
contexts (every host planner), host ghost programs, the regrid clustering
and the frame formatter. Built with the sanitizers on the host side only."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "sanitize", "_build", "host_check")


def test_host_code_under_sanitizers():
    if shutil.which("/opt/rocm/bin/hipcc") is None:
        pytest.skip("hipcc not available")
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize"), "-j8"], check=True,
                   capture_output=True, timeout=900)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0 and "host_check ok" in out.stdout, out.stdout + out.stderr[-4000:]
