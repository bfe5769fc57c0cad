"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer
on the CPU: `make -C fidibench_amd/csrc sanitize` builds
fidibench_amd/csrc/san_host.cpp with every source's host side instrumented,
and the test program exercises the checkpoint schedule (T up to 400), the
reference checksum/std folds, both VTK writers, the raw-file header checks
(good, truncated, foreign, malformed, overflowing axes), the CLI parser,
constructor argument validation and the tuning-variable parser.  Any
sanitizer report aborts the program; leak detection is on."""
from __future__ import annotations

import os
import subprocess

import pytest

from conftest import ROOT

SAN = os.path.join(ROOT, "fidibench_amd", "_build", "san_host")


@pytest.fixture(scope="module")
def san_host():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "fidibench_amd", "csrc"), "sanitize"],
                   check=True, capture_output=True, timeout=1200)
    return SAN


@pytest.mark.parametrize("mode", [[], ["tuning"]])
def test_host_code_is_clean_under_asan_and_ubsan(san_host, mode, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", TMPDIR=str(tmp_path))
    p = subprocess.run([san_host] + mode, capture_output=True, text=True, timeout=600, env=env)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "san_host ok" in p.stdout
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr
