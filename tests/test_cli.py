"""The drop-in upwindCxx executable against the reference executable's stdout
(tests/golden/manifest.json "cli", captured from the reference's own main).

The reference prints two OpenMP banner lines that a GPU run has no use for;
they are removed from the golden text before comparing.  Help output is
compared option by option (upwindCxx here adds -precision and -timeBlock and
its own footer).  The reference's CTest pass rule (upwind/cxx/
CMakeLists.txt:54-91) is applied too.
"""
from __future__ import annotations

import os
import re
import subprocess

import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "fidibench_amd", "bin", "upwindCxx")
CTEST_RE = re.compile(r"[Cc]heck sum:[ ]*[1|0\\.999]")


def run(args):
    return subprocess.run(["upwindCxx"] + args, executable=EXE, capture_output=True, text=True,
                          timeout=300)


def strip_openmp(text):
    return "".join(l for l in text.splitlines(True)
                   if not l.startswith("Running with OpenMP") and not l.startswith("number of threads"))


def help_lines(text):
    keep = []
    for l in text.splitlines():
        if l.startswith("\t-precision") or l.startswith("\t-timeBlock"):
            continue
        if l.startswith("Report bugs") or l.startswith("fidibench-amd"):
            continue
        keep.append(l.rstrip())
    while keep and keep[-1] == "":
        keep.pop()
    return keep


def cases(manifest, help_only):
    out = []
    for c in manifest["cli"]:
        is_help = "-h" in c["args"] or "-bogus" in c["args"]
        if is_help == help_only:
            out.append(c)
    return out


def test_help_and_invalid_option(manifest):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built")
    for c in cases(manifest, help_only=True):
        p = run(c["args"])
        assert help_lines(p.stdout) == help_lines(c["stdout"]), c["args"]
        assert p.stderr == c["stderr"], c["args"]
        assert p.returncode == 0


@pytest.mark.gpu
def test_runs_match_reference_stdout(manifest):
    for c in cases(manifest, help_only=False):
        p = run(c["args"])
        assert p.returncode == 0, p.stderr
        assert p.stdout == strip_openmp(c["stdout"]), c["args"]
        assert CTEST_RE.search(p.stdout), "reference CTest regex must match"


@pytest.mark.gpu
@pytest.mark.parametrize("extra", [["-precision", "f32"], ["-precision", "bf16"],
                                   ["-precision", "bf16sr"], ["-timeBlock", "4"],
                                   ["-timeBlock", "8"], ["-timeBlock", "0"]])
def test_extension_options(extra):
    p = run(["-numCells", "32", "-numSteps", "10"] + extra)
    assert p.returncode == 0, p.stderr
    assert CTEST_RE.search(p.stdout)


@pytest.mark.gpu
def test_vtk_option_writes_reference_format(tmp_path, manifest):
    p = subprocess.run(["upwindCxx", "-numCells", "6", "-numSteps", "2", "-vtk"], executable=EXE,
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert p.returncode == 0, p.stderr
    for f in ("up0.vtk", "up1.vtk"):
        text = (tmp_path / f).read_text()
        assert text.startswith("# vtk DataFile Version 2.0\nupwind.cxx\nASCII\n"
                               "DATASET RECTILINEAR_GRID\nDIMENSIONS 7 7 7\n")
        assert "CELL_DATA 216\nSCALARS f double 1\nLOOKUP_TABLE default\n" in text
