"""Build recipe: compiles libfdb_upwind.so (gfx950) and the upwindCxx executable.

    python -m fidibench_amd.build

Runs ``make -B`` in fidibench_amd/csrc (see its Makefile for the flags:
-ffp-contract=off, --offload-arch=gfx950, packed FP32 off).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")


def build(force: bool = True, jobs: int = 4) -> None:
    cmd = ["make", "-C", CSRC, f"-j{jobs}"] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build(force="--no-force" not in sys.argv)
