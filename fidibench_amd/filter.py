"""Python host side of the GPU stencil filter: the one-process form of
pletzer/fidibench's Filter class (cxx/Filter.h, cxx/Filter.cpp), over the
C ABI in include/fidibench_amd/fdb_filter.h.

Method names and argument meaning follow the reference:
  Filter(globalDims, xmins, xmaxs, stencil)      Filter.cpp:11-78
  setInData(f | array), setInDataByIndices(f)    Filter.cpp:131-190
  applyFilter(), copyOutToIn()                   Filter.cpp:192-263, 441-463
  computeCheckSum("input" | "output")            Filter.cpp:466-485
  saveVTK(filename)                              Filter.cpp:488-538
  getPosition(inds), getRank(), getNumProcs(), isDecompValid()
Arrays are indexed in the reference's axis order; the flat layout is the
reference's column-major one (axis 0 fastest).  `stencil` maps offset
tuples to weights; the terms apply in sorted (std::map) order.
"""
from __future__ import annotations

import ctypes as C
from typing import Callable, Mapping, Sequence

import numpy as np

from . import _lib as L


class Filter:
    def __init__(self, globalDims: Sequence[int], xmins: Sequence[float],
                 xmaxs: Sequence[float], stencil: Mapping[tuple, float]):
        self.dims = tuple(int(d) for d in globalDims)
        self.xmins = tuple(float(x) for x in xmins)
        self.xmaxs = tuple(float(x) for x in xmaxs)
        nd = len(self.dims)
        if len(self.xmins) != nd or len(self.xmaxs) != nd:
            raise ValueError("globalDims, xmins and xmaxs need one entry per axis")
        offs, ws = [], []
        for off, w in stencil.items():
            if len(off) != nd:
                raise ValueError("stencil offsets need one entry per axis")
            offs.extend(int(o) for o in off)
            ws.append(float(w))
        h = C.c_void_p()
        L.check(L.lib().fdb_filter_create(nd, L.u64s(self.dims), L.dbls(self.xmins),
                                          L.dbls(self.xmaxs), len(ws),
                                          (C.c_int * len(offs))(*offs), L.dbls(ws), C.byref(h)))
        self._h = h

    # -- reference members -------------------------------------------------
    def getPosition(self, globalInds: Sequence[int]) -> list:
        return [self.xmins[i] + (globalInds[i] + 0.5) * ((self.xmaxs[i] - self.xmins[i])
                                                         / float(self.dims[i]))
                for i in range(len(self.dims))]

    def setInData(self, f) -> None:
        """f: a function of the cell-centre position (list of floats), or an
        array of all cells (reference axis order, or flat column-major)."""
        if callable(f):
            data = np.array([f(self.getPosition(ix)) for ix in self._indices()])
        else:
            a = np.asarray(f, dtype=np.float64)
            data = a.ravel(order="F") if a.ndim > 1 else a
        self._set(data)

    def setInDataByIndices(self, f: Callable[[list], float]) -> None:
        self._set(np.array([f(list(ix)) for ix in self._indices()], dtype=np.float64))

    def applyFilter(self) -> None:
        L.check(L.lib().fdb_filter_apply(self._h))

    def copyOutToIn(self) -> None:
        L.check(L.lib().fdb_filter_copy_out_to_in(self._h))

    def computeCheckSum(self, inOrOut: str) -> float:
        out = C.c_double()
        L.check(L.lib().fdb_filter_checksum(self._h, 0 if inOrOut == "input" else 1,
                                            C.byref(out)))
        return out.value

    def saveVTK(self, filename: str) -> None:
        L.check(L.lib().fdb_filter_save_vtk(self._h, str(filename).encode()))

    def getRank(self) -> int:
        return 0

    def getNumProcs(self) -> int:
        return 1

    def isDecompValid(self) -> bool:
        return True

    # -- extensions --------------------------------------------------------
    def kernel(self) -> str:
        """The kernel applyFilter launches ("star", "planes", "planes_mid",
        "planes_wide", "rows" or "gather"; fixed at creation)."""
        name = C.c_char_p()
        L.check(L.lib().fdb_filter_kernel(self._h, C.byref(name)))
        return name.value.decode()

    def iterate(self, niters: int) -> float:
        """niters x (applyFilter(); copyOutToIn()); returns kernel ms."""
        ms = C.c_float()
        L.check(L.lib().fdb_filter_iterate_timed(self._h, int(niters), C.byref(ms)))
        return ms.value

    def set_in_sin(self) -> None:
        """The laplacian.cxx input function on the GPU (device sin)."""
        L.check(L.lib().fdb_filter_set_in_sin(self._h))

    def get_data(self, inOrOut: str = "output") -> np.ndarray:
        """All cells as an array in the reference's axis order."""
        flat = np.empty(int(np.prod(self.dims)), dtype=np.float64)
        L.check(L.lib().fdb_filter_get_data(self._h, 0 if inOrOut == "input" else 1,
                                            flat.ctypes.data_as(C.POINTER(C.c_double))))
        return flat.reshape(self.dims, order="F")

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            L.lib().fdb_filter_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers -----------------------------------------------------------
    def _indices(self):
        # column-major order: axis 0 fastest, as MultiArrayIter iterates
        for flat in range(int(np.prod(self.dims))):
            ix, r = [], flat
            for d in self.dims:
                ix.append(r % d)
                r //= d
            yield ix

    def _set(self, data: np.ndarray) -> None:
        data = np.ascontiguousarray(data, dtype=np.float64)
        if data.size != int(np.prod(self.dims)):
            raise ValueError("setInData: wrong number of cells")
        L.check(L.lib().fdb_filter_set_in_data(self._h, data.ctypes.data_as(C.POINTER(C.c_double))))
