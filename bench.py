#!/usr/bin/env python3
"""Benchmark of the upwind advection hot path on one MI355X.

    python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--device-warmup-ms 200]
                    [--no-cpu] [--only f64,f32,...]

A step is one time step of Upwind::advect over the whole field.  The
headline workload is sphere_noise512: 512^3 cells, FP64, seeded mixed-sign
velocity, CFL 0.5, sphere + 0.01 noise (fidibench_amd/configs.py).  Fields
are initialised in HBM before the timed region; the timed region is K
advect steps between two device synchronisations, and kernel time is
measured with HIP events on the stream the kernels run on
(fdb_upwind_advect_timed).  Right before its timed region each leg keeps
the device busy, untimed, for --device-warmup-ms of GPU time (default
200 ms; recorded as device_warmup_ms) -- the forward legs with their own
advect, then their W warm-up steps -- because clocks drop within 0.2 s of
idle host work between legs, which would cost the next 100-step run
12-15 %.

One JSON line: value = FP64 cell-updates/s at the best time-blocking depth
k (steps fused per launch; k = 0: the library's automatic least-time split
of the run into passes), plus variants:
  {f64,f32,bf16}_k{k}  the workload per storage type and depth (_auto for
                   k = 0); every k's final field is compared bit for bit
                   with k = 1's
  mixed_gauss1024  1024^3 x 200 Gaussian bubble in FP64/FP32/BF16 (round to
                   nearest and stochastic): rates, max-abs / RMS / max-rel
                   error vs FP64, mass drift
  fwd_adj_*        forward + adjoint (dL/df0, dL/dv of a misfit) at 512^3
                   over the same steps, by the shift-invariant and the
                   checkpointed method; time ratio to the best forward
  adjoint_step     one adjoint step per storage type (24 / 12 / 6 B a cell)
  adjoint_large    1024x1024x768 x 400 gradient, BF16-SR/FP32 against FP64
  laplacian        the stencil filter (cxx/Filter.cpp) on laplacian.cxx's
                   2-D 8000^2 and a 3-D 512^3 Laplacian, 10 iterations
  copy             the device's plain-copy bandwidth (torch's elementwise
                   kernel over 1 GiB), the ceiling of one-pass kernels and
                   the calibration of the PMC byte counters
and parity (never null: the GPU's FP64 checksum against the reference's at
the CPU leg's step count, against the reference's 512^3 x 100 checksum,
and every k against k = 1), the roofline of the dominant kernel, and the
CPU baseline (the reference's own upwind.cxx, compiled by oracle/Makefile,
timed on this host's cores on a bounded sample of the same workload,
median of 3).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s & %HBM roofline at 512³: FP64 plain vs FP32/BF16 vs fwd+adjoint"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
BYTES = {"f64": 8, "f32": 4, "bf16": 2, "bf16sr": 2}
ALL = ("f64", "f32", "bf16", "mixed", "fwdadj", "adjlarge", "laplacian", "copy")
# VALU lane-operations per second the clock allows: 256 CUs x 4 SIMDs x 16
# lanes x 2.4 GHz (FP64 add/mul and non-packed FP32 issue one lane-op per
# lane per cycle; a probe of independent adds/multiplies reached 35.5 T).
VALU_PEAK = 256 * 4 * 16 * 2.4e9


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--device-warmup-ms", type=float, default=200.0,
                   help="untimed GPU work right before every timed region, for this much "
                        "device time (clocks drop within 0.2 s of idle host work)")
    p.add_argument("--n", type=int, default=512, help="cells per axis")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    p.add_argument("--only", default="", help="comma list of: " + ",".join(ALL))
    p.add_argument("--time-blocks", default="1,4,6,7,8,0",
                   help="comma list of time-blocking depths k (steps fused per launch; 0 = "
                        "the library's automatic split); value is the best one's")
    return p.parse_args()


def sync():
    import torch
    torch.cuda.synchronize()


# ---------------------------------------------------------------- CPU leg

def _cpu_advect_run(w, threads, budget_s, steps=None):
    """One timed advect of the reference (oracle/_ref) -- or the oracle port
    when the reference harness was not built -- in a child process with
    `threads` OpenMP threads, on a sample of the workload (same field,
    velocity and dt): `steps` steps, or as many as fit in about budget_s.
    Returns kind, steps, seconds, threads and the checksum of the result."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    code = r"""
import json, sys, time
sys.path.insert(0, %(root)r)
from oracle import oracle as O
n, v, dt, p, seed, budget, steps = %(args)r
f0 = O.sphere_noise(n, (1.0, 1.0, 1.0), p[:3], p[3], p[4], seed)
kind = "reference" if O.have_ref() else "port"
def run(k):
    if kind == "reference":
        return O.ref_time_advect(f0, v, (1.0, 1.0, 1.0), dt, k)
    t0 = time.perf_counter(); f = O.advect(f0, v, (1.0, 1.0, 1.0), dt, k)
    return time.perf_counter() - t0, O.checksum(f)
if steps is None:
    t1, _ = run(1)
    steps = max(1, min(50, int(budget / max(t1, 1e-3))))
t, cs = run(steps)
print(json.dumps(dict(kind=kind, steps=steps, seconds=t, checksum=repr(cs),
                      threads=O.lib().orc_num_threads())))
""" % dict(root=ROOT, args=(list(w.num_cells), list(w.velocity), w.dt, list(w.init_params),
                            w.init_seed, budget_s, steps))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                         text=True, timeout=600, check=True).stdout.strip().splitlines()[-1]
    return json.loads(out)


def cpu_baseline(w, budget_s=4.0, reps=3, single_budget_s=6.0):
    """The CPU figure on this host: the reference's advect on all its cores,
    the median of `reps` runs of the same sample (about budget_s each), plus
    one run on one core (single_core).  Also returns the sample's step count
    and reference checksum, which the parity check reproduces on the GPU."""
    ncores = len(os.sched_getaffinity(0))
    try:
        runs = [_cpu_advect_run(w, ncores, budget_s)]
        for _ in range(reps - 1):
            runs.append(_cpu_advect_run(w, ncores, budget_s, steps=runs[0]["steps"]))
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "cell-updates/s", "cores": ncores, "kind": None,
                "sample": f"failed: {e}"}
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except Exception:
        pass
    rates = sorted(w.ntot * r["steps"] / r["seconds"] for r in runs)
    r0 = runs[0]
    secs = ", ".join("%.2f" % r["seconds"] for r in runs)
    out = {"value": rates[len(rates) // 2], "unit": "cell-updates/s",
           "cores": r0["threads"], "kind": r0["kind"], "runs": rates,
           "sample": (f"{w.num_cells[0]}^3 x {r0['steps']} steps of the same workload, median "
                      f"of {len(runs)} runs ({secs} s, advect only, "
                      f"OMP_NUM_THREADS={r0['threads']}, {model})"),
           "steps": r0["steps"], "checksum": r0["checksum"]}
    if any(r["checksum"] != r0["checksum"] for r in runs):
        out["checksum_mismatch_between_runs"] = [r["checksum"] for r in runs]
    try:
        r1 = _cpu_advect_run(w, 1, single_budget_s)
        out["single_core"] = {"value": w.ntot * r1["steps"] / r1["seconds"],
                              "sample": (f"{w.num_cells[0]}^3 x {r1['steps']} steps "
                                         f"({r1['seconds']:.2f} s, OMP_NUM_THREADS=1)")}
    except Exception as e:
        out["single_core"] = {"value": None, "sample": f"failed: {e}"}
    return out


# ---------------------------------------------------------------- GPU legs

WARM_MS = 200.0  # set from --device-warmup-ms
_BUSY = []


def gpu_busy(ms=None):
    """Keep the GPU busy, untimed, for about `ms` of device time (default
    WARM_MS) right before a timed region, so that the region starts at the
    clocks a running simulation sees: an idle pause of 0.2 s (host work
    between legs) costs the next 100-step run 12-15 %.  Returns the GPU
    milliseconds spent."""
    import torch
    ms = WARM_MS if ms is None else ms
    if ms <= 0:
        return 0.0
    if not _BUSY:
        _BUSY.append(torch.ones(64 << 20, dtype=torch.float64, device="cuda"))  # 512 MB
    x = _BUSY[0]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    spent = 0.0
    while spent < ms:
        e0.record()
        for _ in range(16):
            x.mul_(1.0)
        e1.record()
        e1.synchronize()
        spent += e0.elapsed_time(e1)
    return spent


def run_forward(w, dtype, steps, warmup, time_block):
    """K timed steps on a fresh field; returns (ms_total, launches, handle)."""
    from fidibench_amd import Upwind
    up = Upwind(w.velocity, w.lengths, w.num_cells, dtype=dtype)
    up.set_time_block(time_block)
    up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
    spent = 0.0
    while spent < WARM_MS:  # untimed: this leg's own kernels, clocks to steady state
        ms, _ = up.advect_timed(48, w.dt)
        spent += ms
    if warmup:
        up.advect(warmup, w.dt)
    up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
    sync()
    ms, launches = up.advect_timed(steps, w.dt)
    sync()
    return ms, launches, up


class _DevArray:
    """A device buffer seen by torch through __cuda_array_interface__."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr,
                                         "data": (ptr, False), "version": 3}


def field_bits(up):
    """The handle's current field as a torch integer tensor of its bits (no copy)."""
    import torch
    ts = {8: "<i8", 4: "<i4", 2: "<i2"}[BYTES[{0: "f64", 1: "f32", 2: "bf16", 3: "bf16"}[up.dtype]]]
    return torch.as_tensor(_DevArray(up.device_ptr(), up.ntot, ts), device="cuda")


def same_bits(a_up, b_up):
    import torch
    return bool(torch.equal(field_bits(a_up), field_bits(b_up)))


def kernel_launch_ms(w, dtype, k, launches=12):
    """Average duration of one launch of the k-step kernel (k = 1: the
    one-step kernel): `launches` launches of exactly k steps each, HIP events
    on the handle's stream around them, after the clock warm-up."""
    ms, n, up = run_forward(w, dtype, k * launches, 0, k)
    up.close()
    return ms / n, n


def compare(a_up, b_up):
    """{max_abs, rms, max_rel} of handle a's field against handle b's."""
    from fidibench_amd import _lib as L
    out = (C.c_double * 4)()
    L.check(L.lib().fdb_compare_dev(a_up.dtype, a_up.device_ptr(), b_up.dtype, b_up.device_ptr(),
                                    a_up.ntot, out, None))
    return {"max_abs": out[0], "rms": out[1], "max_rel": out[3]}


def mixed_precision(n=1024, steps=200):
    """The Gaussian-bubble workload in FP64, FP32 and BF16 storage (round to
    nearest even, and stochastic rounding), each split into fused passes by
    the library's planner (time_block 0; the split changes no value)."""
    from fidibench_amd import configs as K
    w = K.gauss1024(n=n, steps=steps)
    res, ups = {}, {}
    for dt_name in ("f64", "f32", "bf16", "bf16sr"):
        from fidibench_amd import Upwind
        up = Upwind(w.velocity, w.lengths, w.num_cells, dtype=dt_name)
        up.set_time_block(0)
        up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
        m0 = up.mass()
        sync()
        gpu_busy()
        ms, launches = up.advect_timed(steps, w.dt)
        m1 = up.mass()
        res[dt_name] = {"cell_updates_per_s": w.ntot * steps / (ms * 1e-3), "ms": ms,
                        "launches": launches,
                        "hbm_frac_compulsory": w.ntot * steps / (ms * 1e-3) * 2 *
                        BYTES[dt_name] / HBM_PEAK,
                        "mass_drift_rel": abs(m1 - m0) / abs(m0)}
        ups[dt_name] = up
    for dt_name in ("f32", "bf16", "bf16sr"):
        res[dt_name]["error_vs_f64"] = compare(ups[dt_name], ups["f64"])
    for up in ups.values():
        up.close()
    res["workload"] = (f"gauss{n}: {n}^3, {steps} steps, v=(0.7,-0.4,0.2), CFL 0.5, "
                       "exp(-|x-c|^2/0.005), seeded centre")
    return res


def forward_adjoint(w, steps, f_dtype, fwd_ms, method):
    """Forward + backward of a misfit loss over `steps` steps at the
    workload's size through the tape (method "shift" or "checkpoint"); time
    ratio to the forward alone (fwd_ms: {name: ms of a plain forward})."""
    import torch
    from fidibench_amd import _lib as L
    from fidibench_amd.autograd import Tape, default_budget
    tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[f_dtype]
    fd = {"f64": L.F64, "f32": L.F32, "bf16": L.BF16}[f_dtype]
    f0 = torch.empty(w.num_cells, dtype=tdt, device="cuda")
    L.check(L.lib().fdb_init_field_dev(3, L.u64s(w.num_cells), L.dbls(w.lengths), fd, w.init_kind,
                                       L.dbls(w.init_params), len(w.init_params), w.init_seed,
                                       f0.data_ptr(), torch.cuda.current_stream().cuda_stream))
    fT = torch.empty_like(f0)
    lam0 = torch.empty_like(f0)
    lamT = torch.empty_like(f0)
    tape = Tape(w.num_cells, w.velocity, w.lengths, w.dt, steps, fd, fd, default_budget(0.8),
                method=method)
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    sync()
    gpu_busy()
    e0.record()
    tape.forward(f0, fT)
    e1.record()
    lamT.copy_(fT)  # dL/df_T of 0.5*|f_T|^2, a stand-in misfit gradient
    gv = tape.backward(lamT, lam0)
    e2.record()
    sync()
    fwd = e0.elapsed_time(e1)
    tot = e0.elapsed_time(e2)
    st = tape.stats()
    tape.close()
    out = {"method": method, "ms_forward": fwd, "ms_forward_plus_adjoint": tot,
           "cell_steps_per_s": w.ntot * steps / (tot * 1e-3), "snapshots": st["snapshots"],
           "forward_steps": st["forward_steps"], "adjoint_steps": st["adjoint_steps"],
           "grad_v": gv}
    for name, ms in fwd_ms.items():
        if ms:
            out[f"ratio_to_{name}"] = tot / ms
    return out


def adjoint_step_rates(w, reps=10):
    """One adjoint step with the velocity-gradient products (reads lam_{t+1}
    and f_t, writes lam_t) per storage type, HIP-event timed over `reps`
    steps: FP64 moves 24 B per cell, FP32 12 B, bf16 6 B."""
    import torch
    from fidibench_amd import _lib as L
    out = {}
    s = torch.cuda.current_stream().cuda_stream
    for name, tdt, code in (("f64", torch.float64, L.F64), ("f32", torch.float32, L.F32),
                            ("bf16", torch.bfloat16, L.BF16)):
        lam_next = torch.rand(w.num_cells, dtype=torch.float32, device="cuda").to(tdt)
        lam = torch.empty_like(lam_next)
        ft = torch.rand(w.num_cells, dtype=torch.float32, device="cuda").to(tdt)
        acc = torch.zeros(6, dtype=torch.float64, device="cuda")

        def step():
            L.check(L.lib().fdb_adjoint_step_dev(3, L.u64s(w.num_cells), L.dbls(w.velocity),
                                                 L.dbls(w.lengths), w.dt, code,
                                                 lam_next.data_ptr(), lam.data_ptr(), code,
                                                 ft.data_ptr(), acc.data_ptr(), s))
        step()
        sync()
        gpu_busy()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            step()
        e1.record()
        sync()
        ms = e0.elapsed_time(e1) / reps
        nbytes = w.ntot * 3 * lam.element_size()
        out[name] = {"ms_per_step": ms, "cell_steps_per_s": w.ntot / (ms * 1e-3),
                     "bytes_per_cell": 3 * lam.element_size(),
                     "hbm_frac_compulsory": nbytes / (ms * 1e-3) / HBM_PEAK}
        del lam_next, lam, ft
    for name in ("f32", "bf16"):
        out[name]["speedup_vs_f64"] = out["f64"]["ms_per_step"] / out[name]["ms_per_step"]
    return out


def laplacian(iters=10, with_cpu=True):
    """The stencil filter of cxx/Filter.cpp (one process) on laplacian.cxx's
    workloads: the 5-point 2-D Laplacian at its default 8000^2 and the
    7-point 3-D one at 512^3, and their 4th-order (radius 2) forms, `iters`
    x (applyFilter; copyOutToIn), input
    prod sin(2 pi x) made on the GPU.  16 B per cell-update compulsory
    (read input, write output); the oracle port (bit-identical result,
    checked) is timed on the host for one iteration as the CPU figure."""
    import time
    import numpy as np
    from fidibench_amd import Filter
    out = {}
    for name, nd, n, order in (("lap2d_8000", 2, 8000, 2), ("lap3d_512", 3, 512, 2),
                               ("lap2d4_8000", 2, 8000, 4), ("lap3d4_512", 3, 512, 4)):
        # 2nd order: 1, -2, 1 per axis; 4th order: -1/12, 4/3, -5/2, 4/3, -1/12
        taps = {1: 1.0} if order == 2 else {1: 16.0 / 12.0, 2: -1.0 / 12.0}
        st = {tuple([0] * nd): (-2.0 if order == 2 else -30.0 / 12.0) * nd}
        for a in range(nd):
            for k, wk in taps.items():
                for s in (k, -k):
                    o = [0] * nd
                    o[a] = s
                    st[tuple(o)] = wk
        dims = (n,) * nd
        flt = Filter(dims, [0.0] * nd, [1.0] * nd, st)
        flt.set_in_sin()
        flt.iterate(2)  # warm-up
        flt.set_in_sin()
        gpu_busy()
        ms = flt.iterate(iters)
        ntot = n ** nd
        cu = ntot * iters / (ms * 1e-3)
        launch_ms = ms / iters  # one kernel launch per iteration (copyOutToIn swaps buffers)
        achieved = 16 * ntot / (launch_ms * 1e-3) / 1e9
        entry = {"cells": ntot, "iterations": iters, "ms": ms, "cell_updates_per_s": cu,
                 "hbm_frac_compulsory": cu * 16 / HBM_PEAK,
                 "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK / 1e9,
                              "unit": "GB/s", "frac": achieved * 1e9 / HBM_PEAK,
                              "traffic": pmc_traffic(name,
                                                     "filter_march_kernel_laplacian_stats.md"),
                              "kernel": {"star": "filter_star_kernel (3-D star stencil)",
                                         "rows": "filter_march_kernel (rows, 2-D)"}.get(
                                             flt.kernel(),
                                             f"filter_march_kernel ({flt.kernel()}, 3-D)"),
                              "avg_launch_ms": launch_ms},
                 "checksum_output": flt.computeCheckSum("output")}
        if name == "lap2d_8000" and with_cpu:
            try:
                from oracle import oracle as O
                f0 = flt.get_data("input").ravel(order="F").copy()
                flt.applyFilter()
                gpu1 = flt.get_data("output").ravel(order="F")
                t0 = time.perf_counter()
                cpu1 = O.filter_apply(f0, dims, list(st.items()))
                t = time.perf_counter() - t0
                entry["cpu_port"] = {"cell_updates_per_s": ntot / t,
                                     "threads": O.lib().orc_num_threads(),
                                     "sample": "one iteration of the same grid"}
                entry["bit_exact_vs_oracle"] = bool(np.array_equal(gpu1.view(np.uint64),
                                                                   cpu1.view(np.uint64)))
            except Exception as e:  # the CPU figure is reported, never required
                entry["cpu_port"] = {"error": str(e)}
        out[name] = entry
        flt.close()
    return out


def large_adjoint(shape=(1024, 1024, 768), steps=400):
    """The gradient at 1024x1024x768 x 400 steps: BF16-SR forward with FP32
    adjoint by both methods (checkpointed: snapshots sized to the free HBM),
    against the same gradient in FP64; time ratios to the plain BF16
    forwards over the same steps (round to nearest, and the stochastic-
    rounding forward the gradient runs), each split into passes by the
    library's planner (time_block 0)."""
    import torch
    from fidibench_amd import _lib as L, configs as K
    from fidibench_amd.workloads import advect, compare, init_field, misfit_gradient
    w = K.misfit_large(shape, steps)
    wt = K.Workload("misfit_large_target", w.num_cells, steps, (0.8, 0.3, -0.5), w.init_kind,
                    w.init_params, dt_velocity=w.dt_velocity)
    target = advect(init_field(wt, L.F64), wt.velocity, wt, steps)
    g64 = misfit_gradient(w, target, L.F64, L.F64, 64 << 30, method="shift")
    lam64, gv64 = g64.pop("grad_f0"), g64.pop("grad_v")
    fwd = {}
    for name, code, tb in (("plain_bf16_forward_s", L.BF16, 0),
                           ("plain_bf16sr_forward_s", L.BF16_SR, 0)):
        f0 = init_field(w, L.BF16)
        sync()
        gpu_busy()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _advect_code(f0, w, steps, tb, code)
        e1.record()
        sync()
        fwd[name] = e0.elapsed_time(e1) * 1e-3
        del f0
    out = {"workload": (f"misfit_large: {shape[0]}x{shape[1]}x{shape[2]}, {steps} steps, "
                        "8 Gaussian bubbles a*exp(-|x-c|^2/w^2), v=(0.9,0.25,-0.45), target "
                        "at v*=(0.8,0.3,-0.5)"),
           "fp64_shift": g64, "grad_v_fp64": gv64, **fwd}
    for method in ("shift", "checkpoint"):
        free = torch.cuda.mem_get_info()[0]
        sync()
        gpu_busy()
        g = misfit_gradient(w, target, L.BF16_SR, L.F32, int(free * 0.8), method=method,
                            reps=3 if method == "shift" else 1)
        lam, gv = g.pop("grad_f0"), g.pop("grad_v")
        g["grad_f0_error_vs_fp64"] = compare(lam, lam64)
        g["grad_v_rel_error_vs_fp64"] = [abs(a - b) / abs(b) for a, b in zip(gv, gv64)]
        for name, sec in fwd.items():
            g["ratio_to_" + name[:-2]] = g["seconds"] / sec
        out[f"bf16sr_fp32_{method}"] = g
        del lam
    return out


def _advect_code(f0, w, nsteps, time_block, code):
    """nsteps forward steps of a torch field in place of storage type `code`
    (FDB_BF16_SR keyed from step 0, seed 0)."""
    import torch
    from fidibench_amd import _lib as L
    a, b = f0, torch.empty_like(f0)
    res = C.c_int()
    L.check(L.lib().fdb_advect_dev(3, L.u64s(w.num_cells), L.dbls(w.velocity), L.dbls(w.lengths),
                                   w.dt, code, a.data_ptr(), b.data_ptr(), int(nsteps),
                                   int(time_block), torch.cuda.current_stream().cuda_stream,
                                   C.byref(res)))
    return b if res.value else a


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """The device's plain-copy rate: torch's elementwise kernel b = a * 1.0
    over `nbytes` of FP64 (reads and writes nbytes once per launch), HIP
    events around `reps` launches on torch's stream.  A rocprofv3 PMC run of
    this leg also calibrates FETCH_SIZE / WRITE_SIZE against known bytes."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.mul(a, 1.0, out=b)
    sync()
    gpu_busy()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.mul(a, 1.0, out=b)
    e1.record()
    sync()
    ms = e0.elapsed_time(e1) / reps
    return {"bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
            "avg_launch_ms": ms, "GB_per_s": 2 * nbytes / (ms * 1e-3) / 1e9,
            "frac_of_peak": 2 * nbytes / (ms * 1e-3) / HBM_PEAK,
            "kernel": "torch elementwise mul (b = a * 1.0, FP64)"}


def pmc_traffic(kernel_key, summary="pmc_traffic.txt"):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", summary)
    try:
        for line in open(path):
            m = re.match(r"\s*" + re.escape(kernel_key) + r"\s*:\s*([0-9.eE+]+)", line)
            if m:
                return float(m.group(1))
    except OSError:
        pass
    return None


def fp64_checksum(w, steps, k):
    """The reference's checksum (host left fold) of the FP64 GPU field after
    `steps` steps of the workload at time-blocking depth k (untimed)."""
    from fidibench_amd import Upwind
    up = Upwind(w.velocity, w.lengths, w.num_cells)
    up.set_time_block(k)
    up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
    up.advect(steps, w.dt)
    cs = up.checksum()
    up.close()
    return cs


def klabel(dt_name, k):
    return f"{dt_name}_auto" if k == 0 else f"{dt_name}_k{k}"


def plan(w, dtype, steps, k):
    """Steps per launch of the run (fdb_plan_passes)."""
    from fidibench_amd import _lib as L
    return L.plan_passes(w.num_cells, w.velocity, w.lengths, w.dt, L.DTYPES[dtype], steps, k)


def roofline(w, dtype, k, steps):
    """Roofline of the dominant kernel: the k-step forward kernel of
    `dtype` (k = 1: the one-step kernel; k = 0, the automatic split: the
    depth that advances the most steps of the run).  Its average launch duration is
    measured here (kernel_launch_ms).  frac is a position on the HBM
    roofline: compulsory bytes per launch (read state t, write t+k: 2*s per
    cell) / duration / 8 TB/s.  effective_frac credits 2*s bytes to every
    fused step (k x compulsory), the bytes k one-step launches would move.
    traffic: HBM bytes per launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE, WRITE_SIZE, calibrated on the copy leg).  valu: vector-ALU
    lane-operations issued per second (SQ_INSTS_VALU x 64 per launch, from
    the committed summary) against the 39.3 T the clock allows."""
    s = BYTES[dtype]
    schedule = plan(w, dtype, steps, k)
    if k == 0:
        k = max(set(schedule), key=lambda d: (d * schedule.count(d), d))
    launch_ms, _ = kernel_launch_ms(w, dtype, k)
    comp = 2.0 * s * w.ntot
    achieved = comp / (launch_ms * 1e-3)
    out = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": achieved / HBM_PEAK, "traffic": pmc_traffic(f"{dtype}_k{k}"),
           "kernel": (f"tb_kernel<{ {'f64': 'double', 'f32': 'float'}.get(dtype, 'bf16')}> "
                      f"({dtype.upper()} forward, {k} fused steps per launch)" if k > 1 else
                      f"rows_kernel ({dtype.upper()} forward, 1 step per launch)"),
           "steps_per_launch": k, "avg_launch_ms": launch_ms, "schedule": schedule,
           "algorithmic_bytes_per_launch": comp,
           "note": ("algorithmic bytes = compulsory bytes per launch: read state t and write "
                    "state t+k once (2*s per cell); effective_* credit them to each of the k "
                    "fused steps")}
    out["effective_achieved"] = k * achieved / 1e9
    out["effective_frac"] = k * achieved / HBM_PEAK
    valu = pmc_traffic(f"valu_{dtype}_k{k}")
    if valu:
        rate = valu * 64 / (launch_ms * 1e-3)
        out["valu"] = {"lane_ops_per_s": rate, "peak": VALU_PEAK, "frac": rate / VALU_PEAK,
                       "instructions_per_launch": valu,
                       "instructions_per_cell_step": valu * 64 / (w.ntot * k)}
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}): the single-GPU "
              "build has no sharded path to run on more.", file=sys.stderr)
        sys.exit(2)
    import torch  # noqa: F401  (device sync around the timed region; loads the HIP runtime)
    from fidibench_amd import configs as K

    w = K.sphere_noise512(n=a.n, steps=a.steps)
    only = set(a.only.split(",")) if a.only else set(ALL)
    global WARM_MS
    WARM_MS = a.device_warmup_ms
    res = {}
    parity = {"checks": {}}

    # Every storage type at every time-blocking depth k (k fused steps per
    # launch; k = 1 is one launch per step).  Every k's final field must
    # equal k = 1's bit for bit (compared on the device).
    ks = [int(x) for x in a.time_blocks.split(",")]
    if 1 not in ks:
        ks = [1] + ks
    best_of = {}
    for dt_name in ("f64", "f32", "bf16"):
        if dt_name not in only:
            continue
        ref_up = None
        same = {}
        for k in ks:
            ms, launches, up = run_forward(w, dt_name, a.steps, a.warmup, k)
            cu = w.ntot * a.steps / (ms * 1e-3)
            res[klabel(dt_name, k)] = {"cell_updates_per_s": cu, "ms": ms, "launches": launches,
                                      "time_block": k,
                                      "hbm_frac_compulsory": cu * 2 * BYTES[dt_name] / HBM_PEAK}
            if dt_name not in best_of or cu > best_of[dt_name][0]:
                best_of[dt_name] = (cu, k, ms, launches)
            if ref_up is None:
                ref_up = up
            else:
                same[k] = same_bits(up, ref_up)
                res[klabel(dt_name, k)]["bit_identical_to_k1"] = same[k]
                up.close()
        ref_up.close()
        parity["checks"][f"{dt_name}_every_k_equals_k1"] = all(same.values())
    best = best_of.get("f64")
    for dt_name in ("f32", "bf16"):
        if dt_name in best_of and best:
            res[f"{dt_name}_best_speedup_vs_f64_best"] = best_of[dt_name][0] / best[0]
    if "mixed" in only:
        res["mixed_gauss1024"] = mixed_precision()
    if "fwdadj" in only:
        for dt_name in ("f64", "f32", "bf16"):
            fwd = {"best_forward": best_of.get(dt_name, (0, 0, None))[2],
                   "one_step_forward": res.get(f"{dt_name}_k1", {}).get("ms")}
            for method in ("shift", "checkpoint"):
                key = f"fwd_adj_{dt_name}" + ("" if method == "shift" else "_checkpoint")
                res[key] = forward_adjoint(w, a.steps, dt_name, fwd, method)
        res["adjoint_step"] = adjoint_step_rates(w)
    if "adjlarge" in only:
        res["adjoint_large"] = large_adjoint()
    if "laplacian" in only:
        res["laplacian"] = laplacian(with_cpu=not a.no_cpu)
    if "copy" in only:
        res["copy"] = copy_bandwidth()

    out = {"metric": METRIC, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "device_warmup_ms": WARM_MS}
    if best:
        cu, k, ms, launches = best
        out["value"] = cu
        out["ms_per_step"] = ms / a.steps
        out["time_block"] = k
        out["roofline"] = roofline(w, "f64", k, a.steps)
    out["config"] = {"workload": f"sphere_noise{a.n}: {a.n}^3 cells, {a.steps} steps, FP64, "
                                 "CFL 0.5, seeded mixed-sign velocity, sphere+0.01 noise",
                     "num_cells": list(w.num_cells), "velocity": list(w.velocity), "dt": w.dt}
    out["variants"] = res
    k_best = best[1] if best else 8
    if not a.no_cpu:
        cpu = cpu_baseline(w)
        out["cpu_baseline"] = cpu
        if cpu.get("value") and out.get("value"):
            out["gpu_over_cpu"] = out["value"] / cpu["value"]
        if cpu.get("checksum"):
            # The GPU from the same field for the CPU sample's step count,
            # against the checksum the reference computed in that leg.
            g = fp64_checksum(w, cpu["steps"], k_best)
            parity["cpu_leg"] = {"steps": cpu["steps"], "reference_checksum": cpu["checksum"],
                                 "gpu_checksum": repr(g), "kind": cpu["kind"]}
            parity["checks"]["cpu_leg_checksum"] = repr(g) == cpu["checksum"]
    try:
        with open(os.path.join(ROOT, "tests", "golden", "manifest.json")) as fh:
            gold = json.load(fh)["large"].get(f"sphere_noise{a.n}")
    except (OSError, ValueError, KeyError):
        gold = None
    if gold:
        # The reference's checksum of this workload at 100 steps (fixture
        # written by the reference itself), reproduced at the best k.
        g = fp64_checksum(w, 100, k_best)
        parity["golden_100_steps"] = {"reference_checksum": gold["checksum"],
                                      "gpu_checksum": repr(g), "time_block": k_best}
        parity["checks"]["golden_100_steps"] = repr(g) == gold["checksum"]
    parity["bit_exact"] = bool(parity["checks"]) and all(parity["checks"].values())
    out["parity"] = parity
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
