"""ctypes binding of libfdb_upwind.so (the C ABI in include/fidibench_amd/fdb_upwind.h).

The library is built in-tree (fidibench_amd/lib) by ``fidibench_amd.build``.
There is no fallback: if the shared library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfdb_upwind.so")

F64, F32, BF16, BF16_SR = 0, 1, 2, 3
DTYPES = {"f64": F64, "f32": F32, "bf16": BF16, "bf16sr": BF16_SR}
DTYPE_NAMES = {v: k for k, v in DTYPES.items()}
ITEMSIZE = {F64: 8, F32: 4, BF16: 2, BF16_SR: 2}

INIT_DELTA, INIT_SPHERE, INIT_GAUSS, INIT_GAUSS_SUM, INIT_UNIFORM = range(5)
CK_ADVANCE, CK_STORE, CK_RESTORE, CK_POP, CK_ADJOINT = 1, 2, 3, 4, 5
TIME_BLOCK_AUTO = 0  # time_block: the library's least-time split of each advect
GRAD_CHECKPOINT, GRAD_SHIFT = 0, 1
GRAD_METHODS = {"checkpoint": GRAD_CHECKPOINT, "shift": GRAD_SHIFT}

OK = 0
ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_IO, ERR_NODEV = -1, -2, -3, -4, -5

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_i = C.c_int

# name -> (restype, argtypes); every symbol the header declares.
SIGNATURES = {
    "fdb_last_error": (C.c_char_p, []),
    "fdb_version": (_i, []),
    "fdb_upwind_create": (_i, [_i, _dp, _dp, _u64p, _i, C.POINTER(_vp)]),
    "fdb_upwind_destroy": (_i, [_vp]),
    "fdb_upwind_advect": (_i, [_vp, _i, C.c_double]),
    "fdb_upwind_checksum": (_i, [_vp, _dp]),
    "fdb_upwind_std": (_i, [_vp, _dp]),
    "fdb_upwind_print": (_i, [_vp]),
    "fdb_upwind_save_vtk": (_i, [_vp, C.c_char_p]),
    "fdb_upwind_save_raw": (_i, [_vp, C.c_char_p]),
    "fdb_upwind_load_raw": (_i, [_vp, C.c_char_p]),
    "fdb_host_raw_info": (_i, [C.c_char_p, C.POINTER(_i), C.POINTER(_i), _u64p, _u64p, _u64p]),
    "fdb_upwind_ntot": (_i, [_vp, _u64p]),
    "fdb_upwind_dtype": (_i, [_vp, C.POINTER(_i)]),
    "fdb_upwind_set_field": (_i, [_vp, _vp, _i]),
    "fdb_upwind_get_field": (_i, [_vp, _vp, _i]),
    "fdb_upwind_field_device": (_i, [_vp, C.POINTER(_vp)]),
    "fdb_upwind_init_field": (_i, [_vp, _i, _dp, _i, C.c_uint64]),
    "fdb_upwind_mass": (_i, [_vp, _dp]),
    "fdb_upwind_set_time_block": (_i, [_vp, _i]),
    "fdb_upwind_set_seed": (_i, [_vp, C.c_uint64]),
    "fdb_upwind_advect_timed": (_i, [_vp, _i, C.c_double, C.POINTER(C.c_float), C.POINTER(_i)]),
    "fdb_advect_dev": (_i, [_i, _u64p, _dp, _dp, C.c_double, _i, _vp, _vp, _i, _i, _vp,
                            C.POINTER(_i)]),
    "fdb_advect_dev_seeded": (_i, [_i, _u64p, _dp, _dp, C.c_double, _i, _vp, _vp, _i, _i,
                                   C.c_uint64, C.c_uint64, _vp, C.POINTER(_i)]),
    "fdb_plan_passes": (_i, [_i, _u64p, _dp, _dp, C.c_double, _i, _i, _i, C.POINTER(_i), _i,
                             C.POINTER(_i)]),
    "fdb_adjoint_step_dev": (_i, [_i, _u64p, _dp, _dp, C.c_double, _i, _vp, _vp, _i, _vp, _vp,
                                  _vp]),
    "fdb_adjoint_step_dev_noalloc": (_i, [_i, _u64p, _dp, _dp, C.c_double, _i, _vp, _vp, _i,
                                          _vp, _vp, _vp]),
    "fdb_sum_dev": (_i, [_i, _vp, C.c_uint64, _dp, _vp]),
    "fdb_misfit_dev": (_i, [_i, _vp, _vp, C.c_uint64, _i, _vp, _dp, _vp]),
    "fdb_tape_create": (_i, [_i, _u64p, _dp, _dp, C.c_double, _i, _i, _i, C.c_uint64,
                             C.POINTER(_vp)]),
    "fdb_tape_forward": (_i, [_vp, _vp, _vp, _vp]),
    "fdb_tape_backward": (_i, [_vp, _vp, _vp, _dp, _vp]),
    "fdb_tape_stats": (_i, [_vp, C.POINTER(_i), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "fdb_tape_set_time_block": (_i, [_vp, _i]),
    "fdb_tape_set_method": (_i, [_vp, _i]),
    "fdb_tape_destroy": (_i, [_vp]),
    "fdb_init_field_dev": (_i, [_i, _u64p, _dp, _i, _i, _dp, _i, C.c_uint64, _vp, _vp]),
    "fdb_compare_dev": (_i, [_i, _vp, _i, _vp, C.c_uint64, _dp, _vp]),
    "fdb_host_checksum": (C.c_double, [_dp, C.c_uint64]),
    "fdb_host_std": (C.c_double, [_dp, C.c_uint64]),
    "fdb_host_save_vtk": (_i, [_i, _u64p, _dp, _dp, C.c_char_p]),
    "fdb_ckpt_plan": (_i, [_i, _i, C.POINTER(_i), C.POINTER(_i), _i, C.POINTER(C.c_int64)]),
    # fdb_filter.h
    "fdb_filter_create": (_i, [_i, _u64p, _dp, _dp, _i, C.POINTER(_i), _dp, C.POINTER(_vp)]),
    "fdb_filter_destroy": (_i, [_vp]),
    "fdb_filter_set_in_data": (_i, [_vp, _dp]),
    "fdb_filter_set_in_sin": (_i, [_vp]),
    "fdb_filter_apply": (_i, [_vp]),
    "fdb_filter_copy_out_to_in": (_i, [_vp]),
    "fdb_filter_iterate_timed": (_i, [_vp, _i, C.POINTER(C.c_float)]),
    "fdb_filter_kernel": (_i, [_vp, C.POINTER(C.c_char_p)]),
    "fdb_filter_checksum": (_i, [_vp, _i, _dp]),
    "fdb_filter_get_data": (_i, [_vp, _i, _dp]),
    "fdb_filter_save_vtk": (_i, [_vp, C.c_char_p]),
    "fdb_host_write_vtk": (_i, [_i, _u64p, _dp, _dp, _dp, C.c_char_p, C.c_char_p]),
}

_lib = None


class FdbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib() -> C.CDLL:
    """Load the in-tree libfdb_upwind.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -m fidibench_amd.build`")
        # One HIP runtime per process: torch bundles its own libamdhip64
        # (same soname).  Loading torch first makes this library bind to
        # that copy, so device pointers and streams are shared with torch;
        # loading it first would hand torch a runtime it was not built for.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise FdbError(rc, lib().fdb_last_error().decode(errors="replace"))


def u64s(a):
    return (C.c_uint64 * len(a))(*[int(x) for x in a])


def dbls(a):
    return (C.c_double * max(1, len(a)))(*[float(x) for x in a])


def plan_passes(num_cells, velocity, lengths, dt, dtype, nsteps, time_block):
    """Steps per launch, in launch order, of an advect of nsteps steps with
    this time_block (0 = TIME_BLOCK_AUTO): fdb_plan_passes."""
    nd = len(num_cells)
    n = C.c_int()
    check(lib().fdb_plan_passes(nd, u64s(num_cells), dbls(velocity[:nd]), dbls(lengths[:nd]),
                                float(dt), int(dtype), int(nsteps), int(time_block), None, 0,
                                C.byref(n)))
    ks = (C.c_int * max(1, n.value))()
    check(lib().fdb_plan_passes(nd, u64s(num_cells), dbls(velocity[:nd]), dbls(lengths[:nd]),
                                float(dt), int(dtype), int(nsteps), int(time_block), ks, n.value,
                                C.byref(n)))
    return list(ks[:n.value])
