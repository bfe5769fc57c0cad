"""The GPU stencil filter (cxx/Filter.cpp, one process) against the oracle,
bit for bit: Laplacians in 1-3 dimensions, random stencils with offsets up
to +-3 (also at sizes where the reference's unsigned wrap of negative
indices is not periodic), repeated applyFilter/copyOutToIn, checksum and
VTK output, and the drop-in laplacian executable."""
from __future__ import annotations

import itertools
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, bits

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "fidibench_amd", "bin", "laplacian")


def _lap(nd):
    st = {tuple([0] * nd): -2.0 * nd}
    for a in range(nd):
        for s in (1, -1):
            o = [0] * nd
            o[a] = s
            st[tuple(o)] = 1.0
    return st


def _rand_stencil(rng, nd, nterms):
    nterms = min(nterms, 7 ** nd)  # offsets in [-3, 3]: 7^nd distinct vectors
    st = {}
    while len(st) < nterms:
        st[tuple(int(x) for x in rng.integers(-3, 4, size=nd))] = float(rng.standard_normal())
    return st


def _lap4(nd):
    """Fourth-order Laplacian: radius 2 along every axis."""
    st = {tuple([0] * nd): -30.0 / 12.0 * nd}
    for a in range(nd):
        for s, w in ((1, 16.0 / 12.0), (2, -1.0 / 12.0)):
            for sg in (1, -1):
                o = [0] * nd
                o[a] = sg * s
                st[tuple(o)] = w
    return st


def _box27(rng):
    return {o: float(rng.standard_normal()) for o in itertools.product((-1, 0, 1), repeat=3)}


# 2-D grids take the row-marching kernel (512-wide tiles; 1030 x 21 has a
# partial third tile), 3-D grids the plane-marching one: 128 x 8 tiles for
# radius 1 (40 x 30 x 1 with axis-2 offsets that all wrap onto plane 0),
# 128 x 4 tiles for radius 2-3 (rings of 45-79 KB of LDS; 300 x 5 x 3 has
# axis 1 shorter than a tile); 1-D grids the gather kernel.
CASES = [((100,), "lap"), ((64, 48), "lap"), ((6, 10), "lap"), ((33, 8, 20), "lap"),
         ((16, 16, 16), "lap"), ((37,), "rand"), ((24, 9), "rand"), ((10, 12, 7), "rand"),
         ((300, 5, 3), "rand"), ((40, 24, 18), "box"), ((130, 17, 5), "box"),
         ((1030, 21), "rand"), ((40, 30, 1), "lap"), ((40, 24, 30), "lap4"),
         ((70, 9, 11), "lap4"), ((300, 40), "lap4")]


def _star(rng, R):
    """A full 3-D star of radius R with random weights (the term order in
    std::map order pins which weight goes with which neighbour)."""
    st = {(0, 0, 0): float(rng.standard_normal())}
    for a in range(3):
        for k in range(1, R + 1):
            for sg in (1, -1):
                o = [0, 0, 0]
                o[a] = sg * k
                st[tuple(o)] = float(rng.standard_normal())
    return st


# Full 3-D stars of radius 2 take the star kernel (128 x 8 tiles, own
# cells queued in registers, one halo plane in LDS): partial tiles on both
# axes, grids narrower than a tile (12 x 8), sizes whose negative-index wrap
# is not periodic (non-powers of two), and n2 = 5, the smallest it takes;
# n2 = 4 goes to the ring kernel, as do radius-1 stars (random weights).
STAR_CASES = [((257, 9, 5), 2, "star"), ((128, 16, 64), 2, "star"), ((12, 8, 9), 2, "star"),
              ((1000, 24, 7), 2, "star"), ((130, 17, 6), 2, "star"),
              ((64, 16, 4), 2, "planes_mid"), ((130, 17, 6), 1, "planes"),
              ((40, 8, 3), 1, "planes"), ((30, 7, 9), 1, "planes_mid")]


@pytest.mark.parametrize("dims,R,kernel", STAR_CASES)
def test_star_stencils_bit_exact_vs_oracle(oracle, dims, R, kernel):
    from fidibench_amd import Filter
    rng = np.random.default_rng(sum(dims) + R)
    st = _star(rng, R)
    f0 = rng.standard_normal(int(np.prod(dims)))
    flt = Filter(dims, [0.0] * 3, [1.0] * 3, st)
    assert flt.kernel() == kernel
    flt.setInData(f0)
    ref = f0
    for it in range(3):
        flt.applyFilter()
        ref = oracle.filter_apply(ref, dims, list(st.items()))
        got = flt.get_data("output").ravel(order="F")
        assert np.array_equal(bits(got), bits(ref)), (dims, R, it)
        flt.copyOutToIn()
    flt.close()


@pytest.mark.parametrize("dims,kind", CASES)
def test_filter_bit_exact_vs_oracle(oracle, dims, kind):
    from fidibench_amd import Filter
    rng = np.random.default_rng(sum(dims))
    nd = len(dims)
    st = {"lap": lambda: _lap(nd), "rand": lambda: _rand_stencil(rng, nd, 9),
          "box": lambda: _box27(rng), "lap4": lambda: _lap4(nd)}[kind]()
    f0 = rng.standard_normal(int(np.prod(dims)))
    flt = Filter(dims, [0.0] * nd, [1.0] * nd, st)
    flt.setInData(f0)
    ref = f0
    for it in range(3):
        flt.applyFilter()
        ref = oracle.filter_apply(ref, dims, list(st.items()))
        got = flt.get_data("output").ravel(order="F")
        assert np.array_equal(bits(got), bits(ref)), (dims, kind, it)
        flt.copyOutToIn()
    flt.close()


def test_copy_out_to_in_and_checksums(oracle):
    from fidibench_amd import Filter
    dims = (20, 12)
    st = _lap(2)
    rng = np.random.default_rng(0)
    f0 = rng.standard_normal(240)
    flt = Filter(dims, [0, 0], [1, 1], st)
    flt.setInData(f0)
    assert flt.computeCheckSum("output") == 0.0          # output starts at zero
    assert flt.computeCheckSum("input") == _fold(f0)
    flt.applyFilter()
    out1 = oracle.filter_apply(f0, dims, list(st.items()))
    assert flt.computeCheckSum("output") == _fold(out1)
    flt.copyOutToIn()                                     # input := output
    assert flt.computeCheckSum("input") == _fold(out1)
    assert flt.computeCheckSum("output") == _fold(out1)
    flt.setInData(f0)                                     # output keeps its values
    assert flt.computeCheckSum("output") == _fold(out1)
    assert flt.computeCheckSum("input") == _fold(f0)
    ms = flt.iterate(4)
    ref = f0
    for _ in range(4):
        ref = oracle.filter_apply(ref, dims, list(st.items()))
    assert ms > 0 and np.array_equal(flt.get_data("input").ravel(order="F"), ref)
    flt.close()


def _fold(x):
    s = 0.0
    for v in np.asarray(x, dtype=np.float64).ravel():
        s = s + float(v)
    return s


def test_save_vtk_is_host_writer_on_output(tmp_path):
    import ctypes as C
    from fidibench_amd import Filter, _lib as L
    dims, st = (5, 4, 3), _lap(3)
    flt = Filter(dims, [0, 0, 0], [1, 2, 3], st)
    flt.setInData(np.arange(60, dtype=np.float64))
    flt.applyFilter()
    flt.saveVTK(str(tmp_path / "a.vtk"))
    out = flt.get_data("output").ravel(order="F").copy()
    L.check(L.lib().fdb_host_write_vtk(3, L.u64s(dims), L.dbls([0, 0, 0]), L.dbls([1, 2, 3]),
                                       out.ctypes.data_as(C.POINTER(C.c_double)), b"outData",
                                       str(tmp_path / "b.vtk").encode()))
    assert (tmp_path / "a.vtk").read_bytes() == (tmp_path / "b.vtk").read_bytes()
    flt.close()


def _sin_field(dims):
    """laplacian.cxx's func at the cell centres, column-major, with the
    host's sin (numpy's sin is the C library's for float64 scalars)."""
    n = int(np.prod(dims))
    vals = np.empty(n)
    for flat in range(n):
        r, res = flat, 1.0
        for d in dims:
            i = r % d
            r //= d
            pos = 0.0 + (i + 0.5) * ((1.0 - 0.0) / float(d))
            res *= math.sin(2.0 * math.pi * pos)
        vals[flat] = res
    return vals


@pytest.mark.parametrize("nd,n", [(2, 48), (3, 12), (1, 64)])
def test_laplacian_executable(oracle, nd, n):
    p = subprocess.run([EXE, "-numCells", str(n), "-numDims", str(nd)], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    assert lines[0] == "Number of procs: 1"
    assert lines[1] == "global dimensions: " + " ".join([str(n)] * nd) + " "
    assert lines[2] == "Domain decomp " + "1 " * nd
    assert lines[3].startswith("Laplace times min/max/avg: ") and lines[3].endswith(" [seconds]")
    dims = (n,) * nd
    ref = _sin_field(dims)
    st = list(_lap(nd).items())
    for _ in range(10):
        ref = oracle.filter_apply(ref, dims, st)
    s = _fold(ref)
    assert lines[4] == f"Check sums: input = {s:g} output = {s:g}"


def test_laplacian_default_size_vs_oracle(oracle):
    """The reference's default run, 8000^2 x 10 Laplacian iterations,
    against the oracle (input from the device sin, read back)."""
    from fidibench_amd import Filter
    dims, st = (8000, 8000), _lap(2)
    flt = Filter(dims, [0, 0], [1, 1], st)
    flt.set_in_sin()
    f0 = flt.get_data("input").ravel(order="F").copy()
    flt.iterate(10)
    got = flt.get_data("input").ravel(order="F")
    ref = f0
    for _ in range(10):
        ref = oracle.filter_apply(ref, dims, list(st.items()))
    assert np.array_equal(bits(got), bits(ref))
    flt.close()
