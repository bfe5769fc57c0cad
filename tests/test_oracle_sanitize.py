"""The oracle's C restatement under the host address and undefined-
behaviour sanitizers (make -C oracle sanitize): every entry point on small,
odd shapes -- 1-D, 2-D, a size-1 axis, every velocity sign, filter offsets
wider than the axes -- without an out-of-bounds access, leak or UB."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_oracle_clean_under_asan_and_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    p = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "sanitized oracle run ok" in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr
