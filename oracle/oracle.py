"""ctypes bindings for the CPU oracle (and the reference include-harness).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package
(fidibench_amd) never imports this module.

* ``liboracle.so`` (oracle/_build) is the C restatement in
  oracle/upwind_oracle.c of pletzer/fidibench upwind/cxx/upwind.cxx:23-134.
* ``libupwind_ref.so`` (oracle/_ref) is the reference's own upwind.cxx,
  compiled where it lies by oracle/Makefile.  It exists only where the
  reference checkout was present at build time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libupwind_ref.so")

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u16p = C.POINTER(C.c_uint16)

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile the oracle (and _ref when the reference is present)."""
    cmd = ["make", "-s", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(cmd, check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        sig = {
            "orc_advect_f64": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int, _dp]),
            "orc_advect_f32": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int, _fp]),
            "orc_advect_bf16": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int, _u16p]),
            "orc_advect_bf16_sr": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int,
                                             C.c_uint64, C.c_uint64, _u16p]),
            "orc_bf16_sr": (C.c_uint16, [C.c_float, C.c_uint64, C.c_int, C.c_uint32]),
            "orc_sr_bits": (C.c_uint32, [C.c_uint64, C.c_int, C.c_uint32]),
            "orc_sr_step_key": (C.c_uint32, [C.c_uint64, C.c_uint64]),
            "orc_adjoint_f64": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int, _dp]),
            "orc_adjoint_step_lowp": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int,
                                                C.c_void_p, C.c_void_p]),
            "orc_vgrad_terms": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int,
                                          C.c_void_p, C.c_int, C.c_void_p, _dp]),
            "orc_filter_apply": (C.c_int, [C.c_int, _u64p, C.c_int, C.POINTER(C.c_int), _dp,
                                           _dp, _dp]),
            "orc_misfit_grad_f64": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int,
                                              _dp, _dp, _dp, _dp, _dp]),
            "orc_tlm_f64": (C.c_int, [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int,
                                      _dp, _dp, _dp, _dp, _dp]),
            "orc_checksum": (C.c_double, [_dp, C.c_uint64]),
            "orc_std": (C.c_double, [_dp, C.c_uint64]),
            "orc_f32_to_bf16": (C.c_uint16, [C.c_float]),
            "orc_bf16_to_f32": (C.c_float, [C.c_uint16]),
            "orc_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "orc_uniform": (C.c_double, [C.c_uint64, C.c_uint64]),
            "orc_init_sphere_noise": (None, [_u64p, _dp, _dp, C.c_double, C.c_double,
                                             C.c_uint64, _dp]),
            "orc_init_uniform": (None, [C.c_uint64, C.c_double, C.c_double, C.c_uint64, _dp]),
            "orc_num_threads": (C.c_int, []),
            "orc_wtime": (C.c_double, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        if not have_ref():
            raise FileNotFoundError(f"{REF_SO} not built (needs the reference checkout)")
        L = C.CDLL(REF_SO)
        L.ref_upwind_run.restype = C.c_int
        L.ref_upwind_run.argtypes = [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int, _dp, _dp,
                                     _dp, _dp, C.c_char_p]
        L.ref_upwind_print.restype = C.c_int
        L.ref_upwind_print.argtypes = [C.c_int, _u64p, _dp, _dp, C.c_double, C.c_int, _dp]
        L.ref_upwind_time_advect.restype = C.c_double
        L.ref_upwind_time_advect.argtypes = [_u64p, _dp, _dp, C.c_double, C.c_int, _dp, _dp]
        L.ref_upwind_main.restype = C.c_int
        L.ref_upwind_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.ref_num_threads.restype = C.c_int
        _ref = L
    return _ref


# ------------------------------------------------------------------ helpers

def _u64(a):
    return (C.c_uint64 * len(a))(*[int(x) for x in a])


def _d(a):
    return (C.c_double * len(a))(*[float(x) for x in a])


def _ptr(arr: np.ndarray, ct):
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(C.POINTER(ct))


def advect(f0: np.ndarray, v, lengths, dt: float, nsteps: int, dtype: str | None = None) -> np.ndarray:
    """Oracle advect of a field of shape numCells.  dtype f64/f32/bf16 is
    taken from the array (float64, float32, uint16 bf16 bit patterns)
    unless given."""
    n = f0.shape
    nd = len(n)
    if dtype is None:
        dtype = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32",
                 np.dtype(np.uint16): "bf16"}.get(np.asarray(f0).dtype, "f64")
    if dtype == "f64":
        f = np.ascontiguousarray(f0, dtype=np.float64).copy()
        rc = lib().orc_advect_f64(nd, _u64(n), _d(v), _d(lengths), dt, nsteps, _ptr(f, C.c_double))
    elif dtype == "f32":
        f = np.ascontiguousarray(f0, dtype=np.float32).copy()
        rc = lib().orc_advect_f32(nd, _u64(n), _d(v), _d(lengths), dt, nsteps, _ptr(f, C.c_float))
    elif dtype == "bf16":
        assert f0.dtype == np.uint16, "bf16 fields are passed as uint16 bit patterns"
        f = np.ascontiguousarray(f0).copy()
        rc = lib().orc_advect_bf16(nd, _u64(n), _d(v), _d(lengths), dt, nsteps, _ptr(f, C.c_uint16))
    else:
        raise ValueError(dtype)
    if rc:
        raise RuntimeError(f"oracle advect failed: {rc}")
    return f


def advect_bf16_sr(f0: np.ndarray, v, lengths, dt: float, nsteps: int, seed: int = 0,
                   step0: int = 0) -> np.ndarray:
    """bf16 storage (uint16 bit patterns) with stochastically rounded steps."""
    assert f0.dtype == np.uint16
    n = f0.shape
    f = np.ascontiguousarray(f0).copy()
    rc = lib().orc_advect_bf16_sr(len(n), _u64(n), _d(v), _d(lengths), dt, nsteps, seed, step0,
                                  _ptr(f, C.c_uint16))
    if rc:
        raise RuntimeError(rc)
    return f


def adjoint(lam: np.ndarray, v, lengths, dt: float, nsteps: int) -> np.ndarray:
    n = lam.shape
    out = np.ascontiguousarray(lam, dtype=np.float64).copy()
    rc = lib().orc_adjoint_f64(len(n), _u64(n), _d(v), _d(lengths), dt, nsteps, _ptr(out, C.c_double))
    if rc:
        raise RuntimeError(rc)
    return out


_KIND = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.uint16): 2}


def adjoint_step_lowp(lam_next: np.ndarray, v, lengths, dt: float) -> np.ndarray:
    """One adjoint step in float arithmetic on float32 or bf16 (uint16)
    storage, the operation order of the FP64 adjoint."""
    a = np.ascontiguousarray(lam_next)
    kind = _KIND[a.dtype]
    assert kind in (1, 2)
    out = np.empty_like(a)
    rc = lib().orc_adjoint_step_lowp(a.ndim, _u64(a.shape), _d(v), _d(lengths), dt, kind,
                                     a.ctypes.data, out.ctypes.data)
    if rc:
        raise RuntimeError(rc)
    return out


def vgrad_terms(lam_next: np.ndarray, f_t: np.ndarray, v, lengths, dt: float) -> np.ndarray:
    """D_j = sum_i f_t[i] (lam_next[i - u_j e_j] - lam_next[i]) per axis
    (float64, bf16 as uint16 bit patterns), long-double sums."""
    a, f = np.ascontiguousarray(lam_next), np.ascontiguousarray(f_t)
    out = np.zeros(a.ndim)
    rc = lib().orc_vgrad_terms(a.ndim, _u64(a.shape), _d(v), _d(lengths), dt, _KIND[a.dtype],
                               a.ctypes.data, _KIND[f.dtype], f.ctypes.data,
                               out.ctypes.data_as(C.POINTER(C.c_double)))
    if rc:
        raise RuntimeError(rc)
    return out


def filter_apply(data: np.ndarray, dims, stencil) -> np.ndarray:
    """One applyFilter() (cxx/Filter.cpp, one process) on a column-major
    flat array `data` of prod(dims) cells; stencil = list of (offset
    tuple, weight), in insertion order (std::map semantics applied)."""
    d = np.ascontiguousarray(data, dtype=np.float64).ravel()
    nd = len(dims)
    offs = np.array([o for o, _ in stencil], dtype=np.int32).reshape(-1, nd)
    w = np.array([x for _, x in stencil], dtype=np.float64)
    out = np.empty_like(d)
    rc = lib().orc_filter_apply(nd, _u64(dims), len(w),
                                offs.ctypes.data_as(C.POINTER(C.c_int)), _ptr(w, C.c_double),
                                _ptr(d, C.c_double), _ptr(out, C.c_double))
    if rc:
        raise RuntimeError(rc)
    return out


def misfit_grad(f0, target, v, lengths, dt, nsteps):
    """(loss, dL/df0, dL/dv) for L = 0.5*sum((f_T - target)^2), FP64."""
    n = f0.shape
    f0 = np.ascontiguousarray(f0, dtype=np.float64)
    target = np.ascontiguousarray(target, dtype=np.float64)
    gf = np.empty_like(f0)
    gv = np.zeros(3)
    loss = C.c_double()
    rc = lib().orc_misfit_grad_f64(len(n), _u64(n), _d(v), _d(lengths), dt, nsteps,
                                   _ptr(f0, C.c_double), _ptr(target, C.c_double),
                                   C.byref(loss), _ptr(gf, C.c_double), _ptr(gv, C.c_double))
    if rc:
        raise RuntimeError(rc)
    return loss.value, gf, gv[: len(n)].copy()


def tlm(f0, df0, v, dv, lengths, dt, nsteps):
    """(f_T, d f_T) along direction (df0, dv), FP64."""
    n = f0.shape
    f0 = np.ascontiguousarray(f0, dtype=np.float64)
    df0 = np.ascontiguousarray(df0, dtype=np.float64)
    fT = np.empty_like(f0)
    dfT = np.empty_like(f0)
    dvp = np.zeros(3)
    dvp[: len(n)] = dv
    rc = lib().orc_tlm_f64(len(n), _u64(n), _d(v), _d(lengths), dt, nsteps, _ptr(f0, C.c_double),
                           _ptr(df0, C.c_double), _ptr(dvp, C.c_double), _ptr(fT, C.c_double),
                           _ptr(dfT, C.c_double))
    if rc:
        raise RuntimeError(rc)
    return fT, dfT


def checksum(f: np.ndarray) -> float:
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_checksum(_ptr(f, C.c_double), f.size)


def std(f: np.ndarray) -> float:
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_std(_ptr(f, C.c_double), f.size)


def f32_to_bf16(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    return np.array([lib().orc_f32_to_bf16(float(v)) for v in x], dtype=np.uint16)


def bf16_to_f32(h: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16)
    return (h.astype(np.uint32) << 16).view(np.float32)


def sphere_noise(n, lengths=(1.0, 1.0, 1.0), centre=(0.5, 0.5, 0.5), radius=0.15,
                 amp=0.0, seed=0) -> np.ndarray:
    f = np.empty(tuple(n), dtype=np.float64)
    lib().orc_init_sphere_noise(_u64(n), _d(lengths), _d(centre), radius, amp, seed,
                                _ptr(f, C.c_double))
    return f


def uniform(n, lo=0.0, hi=1.0, seed=0) -> np.ndarray:
    f = np.empty(tuple(n), dtype=np.float64)
    lib().orc_init_uniform(f.size, lo, hi, seed, _ptr(f, C.c_double))
    return f


# ------------------------------------------------------- reference harness

def ref_run(f0, v, lengths, dt, nsteps, vtk_path=None, use_ctor_field=False):
    """Run the reference Upwind<nd> on f0 (or its own f[0]=1 field)."""
    n = f0.shape if f0 is not None else None
    f_in = None if use_ctor_field else np.ascontiguousarray(f0, dtype=np.float64)
    out = np.empty(n, dtype=np.float64)
    cs, sd = C.c_double(), C.c_double()
    rc = ref().ref_upwind_run(len(n), _u64(n), _d(v), _d(lengths), dt, nsteps,
                              None if f_in is None else _ptr(f_in, C.c_double),
                              _ptr(out, C.c_double), C.byref(cs), C.byref(sd),
                              None if vtk_path is None else vtk_path.encode())
    if rc:
        raise RuntimeError(rc)
    return out, cs.value, sd.value


def ref_time_advect(f0, v, lengths, dt, nsteps):
    n = f0.shape
    f_in = np.ascontiguousarray(f0, dtype=np.float64)
    cs = C.c_double()
    t = ref().ref_upwind_time_advect(_u64(n), _d(v), _d(lengths), dt, nsteps,
                                     _ptr(f_in, C.c_double), C.byref(cs))
    return t, cs.value
