"""Python mirror of the reference's Upwind<NDIMS> class over the C ABI.

pletzer/fidibench upwind/cxx/upwind.cxx:19-135:
    Upwind(velocity, lengths, numCells)   -> Upwind(...)            :23-49
    advect(numTimeSteps, deltaTime)       -> Upwind.advect(...)     :51-86
    checksum() / std() / print()          -> same names             :91-109
    saveVTK(filename)                     -> Upwind.saveVTK(...)    saveVTK.h:1-60
Extensions: dtype ('f64' bit-exact, 'f32', 'bf16'), set_field/get_field,
init_field, mass (compensated FP64 device sum), time_block, advect_timed.
Errors raise FdbError (the reference has no error reporting).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib as L

_NP = {L.F64: np.float64, L.F32: np.float32, L.BF16: np.uint16, L.BF16_SR: np.uint16}


class Upwind:
    def __init__(self, velocity: Sequence[float], lengths: Sequence[float],
                 numCells: Sequence[int], dtype: str = "f64"):
        nd = len(numCells)
        if not (len(velocity) >= nd and len(lengths) >= nd):
            raise ValueError("velocity, lengths and numCells need NDIMS entries")
        self.ndims = nd
        self.velocity = tuple(float(x) for x in velocity[:nd])
        self.lengths = tuple(float(x) for x in lengths[:nd])
        self.numCells = tuple(int(x) for x in numCells)
        self.dtype = L.DTYPES[dtype] if isinstance(dtype, str) else int(dtype)
        h = C.c_void_p()
        L.check(L.lib().fdb_upwind_create(nd, L.dbls(self.velocity), L.dbls(self.lengths),
                                          L.u64s(self.numCells), self.dtype, C.byref(h)))
        self._h = h

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            L.lib().fdb_upwind_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- the reference surface
    def advect(self, numTimeSteps: int, deltaTime: float | None = None) -> None:
        if deltaTime is None:  # advect(double): one step (upwindAcc.cxx:49)
            numTimeSteps, deltaTime = 1, float(numTimeSteps)
        L.check(L.lib().fdb_upwind_advect(self._h, int(numTimeSteps), float(deltaTime)))

    def checksum(self) -> float:
        r = C.c_double()
        L.check(L.lib().fdb_upwind_checksum(self._h, C.byref(r)))
        return r.value

    def std(self) -> float:
        r = C.c_double()
        L.check(L.lib().fdb_upwind_std(self._h, C.byref(r)))
        return r.value

    def print(self) -> None:
        L.check(L.lib().fdb_upwind_print(self._h))

    def saveVTK(self, filename: str) -> None:
        L.check(L.lib().fdb_upwind_save_vtk(self._h, str(filename).encode()))

    # -- extensions
    def save_raw(self, filename: str) -> None:
        """Binary dump (64-byte header, then the storage-type values)."""
        L.check(L.lib().fdb_upwind_save_raw(self._h, str(filename).encode()))

    def load_raw(self, filename: str) -> None:
        """Resume from save_raw's file: field, step count and seed."""
        L.check(L.lib().fdb_upwind_load_raw(self._h, str(filename).encode()))

    @property
    def ntot(self) -> int:
        n = C.c_uint64()
        L.check(L.lib().fdb_upwind_ntot(self._h, C.byref(n)))
        return n.value

    def set_field(self, f: np.ndarray, dtype: str | None = None) -> None:
        """Replace the field.  f64/f32 arrays convert; uint16 arrays are bf16 bits."""
        f = np.ascontiguousarray(f)
        if f.size != self.ntot:
            raise ValueError(f"field has {f.size} values, expected {self.ntot}")
        code = {np.dtype(np.float64): L.F64, np.dtype(np.float32): L.F32,
                np.dtype(np.uint16): L.BF16}[f.dtype] if dtype is None else L.DTYPES[dtype]
        L.check(L.lib().fdb_upwind_set_field(self._h, f.ctypes.data_as(C.c_void_p), code))

    def get_field(self, dtype: str | None = None) -> np.ndarray:
        """The field as numpy (storage dtype by default; bf16 as uint16 bits)."""
        code = self.dtype if dtype is None else L.DTYPES[dtype]
        out = np.empty(self.numCells, dtype=_NP[code])
        L.check(L.lib().fdb_upwind_get_field(self._h, out.ctypes.data_as(C.c_void_p), code))
        return out

    def init_field(self, kind: int, params: Sequence[float] = (), seed: int = 0) -> None:
        L.check(L.lib().fdb_upwind_init_field(self._h, int(kind), L.dbls(params), len(params),
                                              int(seed)))

    def mass(self) -> float:
        r = C.c_double()
        L.check(L.lib().fdb_upwind_mass(self._h, C.byref(r)))
        return r.value

    def set_seed(self, seed: int) -> None:
        """Seed of the stochastic rounding (dtype 'bf16sr'); restarts the step count."""
        L.check(L.lib().fdb_upwind_set_seed(self._h, int(seed)))

    def set_time_block(self, k: int) -> None:
        """Steps fused per pass (1..8), or 0 (TIME_BLOCK_AUTO) for the
        library's least-time split of each advect; same results for all."""
        L.check(L.lib().fdb_upwind_set_time_block(self._h, int(k)))

    def plan_passes(self, numTimeSteps: int, deltaTime: float, time_block: int) -> list[int]:
        """Steps per launch of advect(numTimeSteps, deltaTime) with this
        time_block (launches nothing)."""
        return L.plan_passes(self.numCells, self.velocity, self.lengths, deltaTime, self.dtype,
                             numTimeSteps, time_block)

    def advect_timed(self, numTimeSteps: int, deltaTime: float) -> tuple[float, int]:
        """advect(); returns (device milliseconds of its kernels, kernel launches)."""
        ms, n = C.c_float(), C.c_int()
        L.check(L.lib().fdb_upwind_advect_timed(self._h, int(numTimeSteps), float(deltaTime),
                                                C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def device_ptr(self) -> int:
        p = C.c_void_p()
        L.check(L.lib().fdb_upwind_field_device(self._h, C.byref(p)))
        return p.value or 0
