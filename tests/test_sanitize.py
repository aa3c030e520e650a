"""The library's host code (seeded generators, format converters, CSR check,
Matrix Market reader) and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (tests/sanitize/host_checks.cpp, built by
tests/sanitize/Makefile with g++ and gcc; no GPU involved), over the
reference's Matrix Market files and malformed ones."""
import glob
import os
import shutil
import subprocess

import pytest

from conftest import GOLDEN, ROOT

SAN = os.path.join(ROOT, "tests", "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "host_checks"
    subprocess.run(["make", "-s", "-C", SAN, f"OUT={exe}", str(exe)], check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), str(tmp_path), *sorted(glob.glob(os.path.join(GOLDEN, "*.mtx")))],
                       capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout, r.stderr[-4000:])
    assert r.returncode == 0
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
