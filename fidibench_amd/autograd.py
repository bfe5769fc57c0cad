"""torch.autograd.Function over the upwind advection and its hand-written adjoint.

    fT = advect(f0, velocity, dt, nsteps, lengths=(1, 1, 1))
    loss = misfit(fT, target)            # 0.5*sum((fT-target)^2), compensated FP64
    loss.backward()                      # d loss / d f0 and d loss / d velocity

The forward runs nsteps of the upwind step (libfdb_upwind.so, the same
kernels as Upwind.advect) with dt held fixed, on a tape (fdb_tape_*).
backward() runs the adjoint kernel -- the exact transpose of the step --
and the block-reduced velocity gradient
dL/dv_j = -(dt*u_j/dx_j) * sum_t <lam_{t+1}, S_j f_t - f_t>, by one of two
methods (fdb_tape_set_method):
  "shift" (default)  the step and the axis shifts are circulant and
      commute, so every term of the sum equals the t = 0 one: T fused
      transpose steps and one product step against f0; no stored states;
  "checkpoint"  every step's products against its own forward state, the
      states recomputed from binomial checkpoints within `max_bytes`.
Both give the same dL/df0 bit for bit and dL/dv equal up to rounding.
f0 may be float64 (the reference's type), float32 or bfloat16; the adjoint
runs in float64 for float64 fields and in float32 otherwise.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import torch

from . import _lib as L

_TORCH_TO_FDB = {torch.float64: L.F64, torch.float32: L.F32, torch.bfloat16: L.BF16}
_FDB_TO_TORCH = {L.F64: torch.float64, L.F32: torch.float32, L.BF16: torch.bfloat16}


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class Tape:
    """Owner of one fdb_tape_t (checkpoint storage for one forward/backward)."""

    def __init__(self, shape: Sequence[int], velocity: Sequence[float],
                 lengths: Sequence[float], dt: float, nsteps: int, f_dtype: int,
                 lam_dtype: int, max_bytes: int, method: str = "checkpoint"):
        self.shape = tuple(int(x) for x in shape)
        self.ndims = len(self.shape)
        self.f_dtype, self.lam_dtype = f_dtype, lam_dtype
        h = C.c_void_p()
        if method == "shift":
            # No snapshots in this method: ask for the fixed working set only
            # (initial state, two working buffers, one adjoint buffer).
            ntot = 1
            for n in self.shape:
                ntot *= n
            fixed = ntot * (3 * L.ITEMSIZE[f_dtype] + L.ITEMSIZE[lam_dtype])
            max_bytes = min(int(max_bytes), fixed)
        L.check(L.lib().fdb_tape_create(self.ndims, L.u64s(self.shape), L.dbls(velocity),
                                        L.dbls(lengths), float(dt), int(nsteps), f_dtype,
                                        lam_dtype, int(max_bytes), C.byref(h)))
        self._h = h
        self.method = method
        L.check(L.lib().fdb_tape_set_method(h, L.GRAD_METHODS[method]))
        self._f0 = None

    def forward(self, f0: torch.Tensor, fT: torch.Tensor) -> None:
        L.check(L.lib().fdb_tape_forward(self._h, f0.data_ptr(), fT.data_ptr(), _stream()))
        # "shift" reads f0 again in backward: keep the tensor alive until then.
        self._f0 = f0 if self.method == "shift" else None

    def backward(self, lamT: torch.Tensor, lam0: torch.Tensor) -> list:
        gv = (C.c_double * 3)()
        L.check(L.lib().fdb_tape_backward(self._h, lamT.data_ptr(), lam0.data_ptr(), gv,
                                          _stream()))
        return [gv[i] for i in range(self.ndims)]

    def set_time_block(self, k: int) -> None:
        """Steps per launch (1..8): fused forward runs and, where every state
        is stored, fused 4-step adjoint runs; 1 = one step per launch."""
        L.check(L.lib().fdb_tape_set_time_block(self._h, int(k)))

    def stats(self) -> dict:
        s, f, a = C.c_int(), C.c_int64(), C.c_int64()
        L.check(L.lib().fdb_tape_stats(self._h, C.byref(s), C.byref(f), C.byref(a)))
        return {"snapshots": s.value, "forward_steps": f.value, "adjoint_steps": a.value}

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            L.lib().fdb_tape_destroy(self._h)
            self._h = C.c_void_p()
        self._f0 = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_budget(fraction: float = 0.85) -> int:
    free, _ = torch.cuda.mem_get_info()
    return int(free * fraction)


class UpwindAdvect(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, velocity, dt, nsteps, lengths, max_bytes, lam_dtype, method):
        if not f0.is_cuda:
            raise ValueError("f0 must be a GPU tensor")
        f_dtype = _TORCH_TO_FDB[f0.dtype]
        lam = lam_dtype if lam_dtype is not None else (L.F64 if f_dtype == L.F64 else L.F32)
        v = [float(x) for x in velocity.detach().double().cpu().reshape(-1).tolist()]
        f0c = f0.detach().contiguous()
        if not any(ctx.needs_input_grad[:2]):
            # No gradient wanted: plain ping-pong advect, no checkpoints.
            a = f0c.clone()
            b = torch.empty_like(a)
            res = C.c_int()
            L.check(L.lib().fdb_advect_dev(a.dim(), L.u64s(a.shape), L.dbls(v), L.dbls(lengths),
                                           float(dt), f_dtype, a.data_ptr(), b.data_ptr(),
                                           int(nsteps), 1, _stream(), C.byref(res)))
            ctx.tape = None
            return b if res.value else a
        budget = max_bytes if max_bytes is not None else default_budget()
        tape = Tape(f0c.shape, v, lengths, dt, nsteps, f_dtype, lam, budget, method=method)
        fT = torch.empty_like(f0c)
        tape.forward(f0c, fT)
        ctx.tape = tape
        ctx.lam = lam
        ctx.f_dtype = f0.dtype
        ctx.v_meta = (velocity.dtype, velocity.device)
        return fT

    @staticmethod
    def backward(ctx, grad_fT):
        tape = ctx.tape
        lam_t = _FDB_TO_TORCH[ctx.lam]
        lamT = grad_fT.detach().to(lam_t).contiguous()
        lam0 = torch.empty_like(lamT)
        gv = tape.backward(lamT, lam0)
        tape.close()
        vdtype, vdev = ctx.v_meta
        g_f0 = lam0.to(ctx.f_dtype) if ctx.needs_input_grad[0] else None
        g_v = (torch.tensor(gv, dtype=torch.float64).to(vdtype).to(vdev)
               if ctx.needs_input_grad[1] else None)
        return g_f0, g_v, None, None, None, None, None, None


def advect(f0: torch.Tensor, velocity: torch.Tensor, dt: float, nsteps: int,
           lengths: Sequence[float] = (1.0, 1.0, 1.0), max_bytes: int | None = None,
           lam_dtype: int | None = None, method: str = "shift") -> torch.Tensor:
    """f_T after nsteps upwind steps from f0 (differentiable in f0 and
    velocity); method "shift" or "checkpoint" (module docstring)."""
    nd = f0.dim()
    if method not in L.GRAD_METHODS:
        raise ValueError(f"method must be one of {sorted(L.GRAD_METHODS)}")
    return UpwindAdvect.apply(f0, velocity, float(dt), int(nsteps), tuple(lengths)[:nd],
                              max_bytes, lam_dtype, method)


class _Misfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fT, target):
        lam_dtype = L.F64 if fT.dtype == torch.float64 else L.F32
        lam = torch.empty(fT.shape, dtype=_FDB_TO_TORCH[lam_dtype], device=fT.device)
        loss = C.c_double()
        L.check(L.lib().fdb_misfit_dev(_TORCH_TO_FDB[fT.dtype], fT.contiguous().data_ptr(),
                                       target.to(fT.dtype).contiguous().data_ptr(), fT.numel(),
                                       lam_dtype, lam.data_ptr(), C.byref(loss), _stream()))
        ctx.save_for_backward(lam)
        return torch.tensor(loss.value, dtype=torch.float64, device=fT.device)

    @staticmethod
    def backward(ctx, g):
        (lam,) = ctx.saved_tensors
        return lam * g.to(lam.dtype), None


def misfit(fT: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """0.5*sum((fT - target)^2) as an FP64 compensated device reduction."""
    return _Misfit.apply(fT, target)


def compensated_sum(x: torch.Tensor) -> float:
    """FP64 compensated (Neumaier + TwoSum) sum of a device tensor."""
    out = C.c_double()
    L.check(L.lib().fdb_sum_dev(_TORCH_TO_FDB[x.dtype], x.contiguous().data_ptr(), x.numel(),
                                C.byref(out), _stream()))
    return out.value
