"""Host logic of libftdx.so under AddressSanitizer + UndefinedBehaviorSanitizer (CPU).

tests/cpp/test_host_sanitized.cpp is built with g++ -fsanitize=address,undefined from
the library's host-only sources (ftdx_host.cpp: grouped-launch schedule and threaded
host copy; ftdx_tune.cpp: tuning-cache parser and writer) and run against malformed
cache files, concurrent writers of one cache path and extreme problem lists, as the
reference's sanitizer build type runs its host code (CMakeLists.txt:273-274).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_logic_under_asan_ubsan(tmp_path):
    b = subprocess.run(["make", "-s", "-C", CPP, "test_host_sanitized"], capture_output=True, text=True,
                       timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(CPP, "test_host_sanitized"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
