"""Gradient workloads of the upwind path on the GPU (torch-owned buffers).

misfit_gradient(): loss 0.5*sum((f_T - f_target)^2) after T steps, with the
target made by a forward run at the target velocity and the gradient
(dL/df0, dL/dv) taken by the checkpointed adjoint at the evaluation
velocity, dt fixed by the target's sum|v|.  Used by bench.py (the
1024x1024x768 x 400-step BF16-forward / FP32-adjoint run against the same
run in FP64) and by the GPU tests (256^3 x 50 against finite differences).
"""
from __future__ import annotations

import ctypes as C
import time

import torch

from . import _lib as L
from .autograd import Tape, compensated_sum  # noqa: F401
from .configs import Workload

TORCH = {L.F64: torch.float64, L.F32: torch.float32, L.BF16: torch.bfloat16,
         L.BF16_SR: torch.bfloat16}
# Storage code of a tensor for the device-pointer entry points (bf16
# tensors are plain bf16: no stochastic rounding outside a tape).
_CODE = {torch.float64: L.F64, torch.float32: L.F32, torch.bfloat16: L.BF16}


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def init_field(w: Workload, dtype: int, params=None, seed=None) -> torch.Tensor:
    f = torch.empty(w.num_cells, dtype=TORCH[dtype], device="cuda")
    p = w.init_params if params is None else params
    L.check(L.lib().fdb_init_field_dev(3, L.u64s(w.num_cells), L.dbls(w.lengths), dtype,
                                       w.init_kind, L.dbls(p), len(p),
                                       w.init_seed if seed is None else seed, f.data_ptr(),
                                       _stream()))
    return f


def advect(f0: torch.Tensor, velocity, w: Workload, nsteps: int, time_block: int = 8):
    """nsteps plain forward steps (no tape); returns a new tensor."""
    a = f0.clone()
    b = torch.empty_like(a)
    res = C.c_int()
    L.check(L.lib().fdb_advect_dev(3, L.u64s(w.num_cells), L.dbls(velocity), L.dbls(w.lengths),
                                   w.dt, _CODE[a.dtype],
                                   a.data_ptr(), b.data_ptr(), int(nsteps), int(time_block),
                                   _stream(), C.byref(res)))
    return b if res.value else a


def misfit(fT: torch.Tensor, target: torch.Tensor, lam_dtype: int):
    """(loss, lam_T) with lam_T = fT - target in lam_dtype; FP64 compensated loss."""
    lam = torch.empty(fT.shape, dtype=TORCH[lam_dtype], device="cuda")
    loss = C.c_double()
    fd = _CODE[fT.dtype]
    L.check(L.lib().fdb_misfit_dev(fd, fT.data_ptr(), target.to(fT.dtype).data_ptr(),
                                   fT.numel(), lam_dtype, lam.data_ptr(), C.byref(loss),
                                   _stream()))
    return loss.value, lam


def compare(a: torch.Tensor, b: torch.Tensor) -> dict:
    """max-abs, RMS and max-rel (max|a-b|/max|b|) of a against b."""
    out = (C.c_double * 4)()
    fd = _CODE
    L.check(L.lib().fdb_compare_dev(fd[a.dtype], a.data_ptr(), fd[b.dtype], b.data_ptr(),
                                    a.numel(), out, _stream()))
    return {"max_abs": out[0], "rms": out[1], "max_rel": out[3]}


def misfit_gradient(w_eval: Workload, target: torch.Tensor, f_dtype: int, lam_dtype: int,
                    max_bytes: int, f0: torch.Tensor | None = None,
                    method: str = "checkpoint", reps: int = 1) -> dict:
    """Gradient of the misfit at w_eval's velocity and field, by the tape's
    checkpointed or shift-invariant method.  With reps > 1 the forward,
    misfit and backward run reps times on one tape (same results) and
    `seconds` is the median; `seconds_all` lists every run."""
    if f0 is None:
        f0 = init_field(w_eval, f_dtype)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tape = Tape(w_eval.num_cells, w_eval.velocity, w_eval.lengths, w_eval.dt, w_eval.steps,
                f_dtype, lam_dtype, max_bytes, method=method)
    t1 = time.perf_counter()
    runs = []
    for _ in range(max(1, reps)):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        fT = torch.empty_like(f0)
        tape.forward(f0, fT)
        loss, lamT = misfit(fT, target, lam_dtype)
        lam0 = torch.empty_like(lamT)
        grad_v = tape.backward(lamT, lam0)
        torch.cuda.synchronize()
        runs.append(time.perf_counter() - ta)
        del fT, lamT
    st = tape.stats()
    tape.close()
    return {"loss": loss, "grad_f0": lam0, "grad_v": grad_v,
            "seconds": sorted(runs)[len(runs) // 2], "seconds_all": runs,
            "alloc_seconds": t1 - t0, **st}
