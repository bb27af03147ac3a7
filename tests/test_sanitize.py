"""Host-only AddressSanitizer / UndefinedBehaviorSanitizer build of the code
that reads untrusted input or manages lifetimes on the host (edge-list
parser, prepared-graph readers, generators, the launch cache's LRU map, the
oracle): tests/sanitize/Makefile builds tests/sanitize/host_checks.cpp with
g++/gcc (no GPU code) and the test runs it.  Any sanitizer report fails the
run."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HERE = os.path.join(ROOT, "tests", "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None,
                    reason="needs g++ and make")
def test_host_code_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:exitcode=86"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1:exitcode=87"
    env["OMP_NUM_THREADS"] = "4"
    r = subprocess.run([os.path.join(HERE, "_build", "host_checks")], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.returncode, r.stdout[-4000:], r.stderr[-8000:])
    assert "host checks passed" in r.stdout
