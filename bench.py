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
k (steps fused per launch; bit-exact with the reference for every k,
checked against its 512^3 x 100 checksum), plus variants:
  {f64,f32,bf16}_k{k}  the workload per storage type and depth
  mixed_gauss1024  1024^3 x 200 Gaussian bubble in FP64/FP32/BF16 (round to
                   nearest and stochastic): rates, max-abs / RMS / max-rel
                   error vs FP64, mass drift
  fwd_adj_*        forward + adjoint (dL/df0, dL/dv of a misfit) at 512^3
                   over the same steps, time ratio to the one-step forward
  adjoint_step     one adjoint step per storage type (24 / 12 / 6 B a cell)
  adjoint_large    1024x1024x768 x 400 checkpointed gradient, BF16-SR/FP32
                   against FP64
  laplacian        the stencil filter (cxx/Filter.cpp) on laplacian.cxx's
                   2-D 8000^2 and a 3-D 512^3 Laplacian, 10 iterations
and the roofline of the dominant kernel and the CPU baseline (the
reference's own upwind.cxx, compiled by oracle/Makefile, timed on this
host's cores on a bounded sample of the same workload).
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
ALL = ("f64", "f32", "bf16", "mixed", "fwdadj", "adjlarge", "laplacian")


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
    p.add_argument("--time-blocks", default="1,4,6,8",
                   help="comma list of time-blocking depths k (steps fused per launch)")
    return p.parse_args()


def sync():
    import torch
    torch.cuda.synchronize()


# ---------------------------------------------------------------- CPU leg

def _cpu_advect_run(w, threads, budget_s):
    """One timed advect of the reference (oracle/_ref) -- or the oracle port
    when the reference harness was not built -- in a child process with
    `threads` OpenMP threads, on a sample of the workload (same field,
    velocity and dt) of as many steps as fit in about budget_s."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close",
               OMP_PLACES="cores")
    code = r"""
import json, sys, time
sys.path.insert(0, %(root)r)
from oracle import oracle as O
n, v, dt, p, seed, budget = %(args)r
f0 = O.sphere_noise(n, (1.0, 1.0, 1.0), p[:3], p[3], p[4], seed)
kind = "reference" if O.have_ref() else "port"
if kind == "reference":
    t1, _ = O.ref_time_advect(f0, v, (1.0, 1.0, 1.0), dt, 1)
    steps = max(1, min(50, int(budget / max(t1, 1e-3))))
    t, cs = O.ref_time_advect(f0, v, (1.0, 1.0, 1.0), dt, steps)
else:
    t0 = time.perf_counter(); O.advect(f0, v, (1.0, 1.0, 1.0), dt, 1); t1 = time.perf_counter() - t0
    steps = max(1, min(50, int(budget / max(t1, 1e-3))))
    t0 = time.perf_counter(); O.advect(f0, v, (1.0, 1.0, 1.0), dt, steps); t = time.perf_counter() - t0
print(json.dumps(dict(kind=kind, steps=steps, seconds=t, threads=O.lib().orc_num_threads())))
""" % dict(root=ROOT, args=(list(w.num_cells), list(w.velocity), w.dt, list(w.init_params),
                            w.init_seed, budget_s))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                         text=True, timeout=600, check=True).stdout.strip().splitlines()[-1]
    return json.loads(out)


def cpu_baseline(w, budget_s=12.0, single_budget_s=6.0):
    """The CPU figure on this host: the reference's advect on all its cores
    (about budget_s of work), plus the same on one core (single_core)."""
    ncores = len(os.sched_getaffinity(0))
    try:
        r = _cpu_advect_run(w, ncores, budget_s)
    except Exception as e:  # the baseline is reported, never required
        return {"value": None, "unit": "cell-updates/s", "cores": ncores, "kind": None,
                "sample": f"failed: {e}"}
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
    except Exception:
        pass
    out = {"value": w.ntot * r["steps"] / r["seconds"], "unit": "cell-updates/s",
           "cores": r["threads"], "kind": r["kind"],
           "sample": (f"{w.num_cells[0]}^3 x {r['steps']} steps of the same workload "
                      f"({r['seconds']:.2f} s, advect only, OMP_NUM_THREADS={r['threads']}, "
                      f"{model})")}
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


def compare(a_up, b_up):
    """{max_abs, rms, max_rel} of handle a's field against handle b's."""
    from fidibench_amd import _lib as L
    out = (C.c_double * 4)()
    L.check(L.lib().fdb_compare_dev(a_up.dtype, a_up.device_ptr(), b_up.dtype, b_up.device_ptr(),
                                    a_up.ntot, out, None))
    return {"max_abs": out[0], "rms": out[1], "max_rel": out[3]}


def mixed_precision(n=1024, steps=200):
    """The Gaussian-bubble workload in FP64, FP32 and BF16 storage (round to
    nearest even, and stochastic rounding); FP64/FP32 fuse 8 steps per pass
    (same results)."""
    from fidibench_amd import configs as K
    w = K.gauss1024(n=n, steps=steps)
    res, ups = {}, {}
    for dt_name in ("f64", "f32", "bf16", "bf16sr"):
        from fidibench_amd import Upwind
        up = Upwind(w.velocity, w.lengths, w.num_cells, dtype=dt_name)
        up.set_time_block(8)
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


def forward_adjoint(w, steps, f_dtype, fwd_ms):
    """Forward (keeping checkpoints) + backward of a misfit loss over
    `steps` steps at the workload's size; time ratio to the forward alone."""
    import torch
    from fidibench_amd import _lib as L
    from fidibench_amd.autograd import Tape, default_budget
    from fidibench_amd import configs as K
    tdt = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}[f_dtype]
    fd = {"f64": L.F64, "f32": L.F32, "bf16": L.BF16}[f_dtype]
    f0 = torch.empty(w.num_cells, dtype=tdt, device="cuda")
    L.check(L.lib().fdb_init_field_dev(3, L.u64s(w.num_cells), L.dbls(w.lengths), fd, w.init_kind,
                                       L.dbls(w.init_params), len(w.init_params), w.init_seed,
                                       f0.data_ptr(), torch.cuda.current_stream().cuda_stream))
    fT = torch.empty_like(f0)
    lam0 = torch.empty_like(f0)
    tape = Tape(w.num_cells, w.velocity, w.lengths, w.dt, steps, fd, fd, default_budget(0.8))
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    sync()
    gpu_busy()
    e0.record()
    tape.forward(f0, fT)
    e1.record()
    lamT = fT.clone()  # dL/df_T of 0.5*|f_T|^2, a stand-in misfit gradient
    gv = tape.backward(lamT, lam0)
    e2.record()
    sync()
    fwd = e0.elapsed_time(e1)
    tot = e0.elapsed_time(e2)
    st = tape.stats()
    tape.close()
    return {"ms_forward_with_checkpoints": fwd, "ms_forward_plus_adjoint": tot,
            "ratio_to_plain_forward": tot / fwd_ms if fwd_ms else None,
            "cell_steps_per_s": w.ntot * steps / (tot * 1e-3), "snapshots": st["snapshots"],
            "forward_steps": st["forward_steps"], "adjoint_steps": st["adjoint_steps"],
            "grad_v": gv}


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
    7-point 3-D one at 512^3, `iters` x (applyFilter; copyOutToIn), input
    prod sin(2 pi x) made on the GPU.  16 B per cell-update compulsory
    (read input, write output); the oracle port (bit-identical result,
    checked) is timed on the host for one iteration as the CPU figure."""
    import time
    import numpy as np
    from fidibench_amd import Filter
    out = {}
    for name, nd, n in (("lap2d_8000", 2, 8000), ("lap3d_512", 3, 512)):
        st = {tuple([0] * nd): -2.0 * nd}
        for a in range(nd):
            for s in (1, -1):
                o = [0] * nd
                o[a] = s
                st[tuple(o)] = 1.0
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
                              "kernel": ("filter_march_kernel (rows, 2-D)" if nd == 2 else
                                         "filter_march_kernel (planes, 3-D)"),
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
    """The checkpointed gradient at 1024x1024x768 x 400 steps: BF16-storage
    forward with FP32 adjoint, against the same gradient in FP64; time
    ratio to the plain BF16 forward over the same steps."""
    import torch
    from fidibench_amd import _lib as L, configs as K
    from fidibench_amd.workloads import advect, compare, init_field, misfit_gradient
    w = K.misfit_large(shape, steps)
    wt = K.Workload("misfit_large_target", w.num_cells, steps, (0.8, 0.3, -0.5), w.init_kind,
                    w.init_params, dt_velocity=w.dt_velocity)
    target = advect(init_field(wt, L.F64), wt.velocity, wt, steps)
    free = torch.cuda.mem_get_info()[0]
    g64 = misfit_gradient(w, target, L.F64, L.F64, int(free * 0.7))
    lam64, gv64 = g64.pop("grad_f0"), g64.pop("grad_v")
    free = torch.cuda.mem_get_info()[0]
    g16 = misfit_gradient(w, target, L.BF16_SR, L.F32, int(free * 0.8))
    lam16, gv16 = g16.pop("grad_f0"), g16.pop("grad_v")
    err = compare(lam16, lam64)
    del lam16, lam64
    f0 = init_field(w, L.BF16)
    sync()
    gpu_busy()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    advect(f0, w.velocity, w, steps, 1)
    e1.record()
    sync()
    fwd_s = e0.elapsed_time(e1) * 1e-3
    return {"workload": (f"misfit_large: {shape[0]}x{shape[1]}x{shape[2]}, {steps} steps, "
                         "8 Gaussian bubbles a*exp(-|x-c|^2/w^2), v=(0.9,0.25,-0.45), target "
                         "at v*=(0.8,0.3,-0.5)"),
            "bf16_fp32": {**g16, "ratio_to_plain_bf16_forward": g16["seconds"] / fwd_s},
            "fp64": g64, "plain_bf16_forward_s": fwd_s,
            "grad_f0_error_vs_fp64": err,
            "grad_v_rel_error_vs_fp64": [abs(a - b) / abs(b) for a, b in zip(gv16, gv64)],
            "grad_v_fp64": gv64}


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
    best = None

    # Every storage type at every time-blocking depth k (k fused steps per
    # launch; k = 1 is one launch per step); results are identical for all
    # k, and the FP64 runs are checked against the reference's checksum.
    ks = [int(x) for x in a.time_blocks.split(",")]
    best_of = {}
    for dt_name in ("f64", "f32", "bf16"):
        if dt_name not in only:
            continue
        for k in ks:
            ms, launches, up = run_forward(w, dt_name, a.steps, a.warmup, k)
            cu = w.ntot * a.steps / (ms * 1e-3)
            entry = {"cell_updates_per_s": cu, "ms": ms, "launches": launches,
                     "time_block": k,
                     "hbm_frac_compulsory": cu * 2 * BYTES[dt_name] / HBM_PEAK}
            if dt_name == "f64" and a.steps == 100 and a.n == 512:
                entry["checksum"] = up.checksum()
            res[f"{dt_name}_k{k}"] = entry
            if dt_name not in best_of or cu > best_of[dt_name][0]:
                best_of[dt_name] = (cu, k, ms, launches)
            up.close()
    best = best_of.get("f64")
    for dt_name in ("f32", "bf16"):
        if dt_name in best_of and best:
            res[f"{dt_name}_best_speedup_vs_f64_best"] = best_of[dt_name][0] / best[0]
    if "mixed" in only:
        res["mixed_gauss1024"] = mixed_precision()
    if "fwdadj" in only:
        # The ratio is to the plain one-step forward over the same steps.
        fwd_ms = res.get("f64_k1", {}).get("ms")
        res["fwd_adj_f64"] = forward_adjoint(w, a.steps, "f64", fwd_ms)
        f32_ms = res.get("f32_k1", {}).get("ms")
        res["fwd_adj_f32"] = forward_adjoint(w, a.steps, "f32", f32_ms)
        bf16_ms = res.get("bf16_k1", {}).get("ms")
        res["fwd_adj_bf16"] = forward_adjoint(w, a.steps, "bf16", bf16_ms)
        res["adjoint_step"] = adjoint_step_rates(w)
    if "adjlarge" in only:
        res["adjoint_large"] = large_adjoint()
    if "laplacian" in only:
        res["laplacian"] = laplacian(with_cpu=not a.no_cpu)

    parity = None
    try:
        with open(os.path.join(ROOT, "tests", "golden", "manifest.json")) as fh:
            gold = json.load(fh)["large"].get("sphere_noise512")
        got = [v.get("checksum") for k, v in res.items() if k.startswith("f64_k")]
        if gold and got and all(g is not None for g in got):
            parity = {"reference_checksum": gold["checksum"],
                      "checksums": [repr(g) for g in got],
                      "bit_exact": all(repr(g) == gold["checksum"] for g in got)}
    except Exception:
        pass

    out = {"metric": METRIC, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "device_warmup_ms": WARM_MS}
    if best:
        cu, k, ms, launches = best
        out["value"] = cu
        out["ms_per_step"] = ms / a.steps
        per_launch_ms = ms / launches
        steps_per_launch = a.steps / launches
        alg_bytes = 16.0 * w.ntot * steps_per_launch
        achieved = alg_bytes / (per_launch_ms * 1e-3)
        out["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                           "unit": "GB/s", "frac": achieved / HBM_PEAK,
                           "traffic": pmc_traffic(f"f64_k{k}"),
                           "kernel": (f"tb_kernel<double> (FP64 forward, {k} fused steps per "
                                      "launch)" if k > 1 else
                                      "rows_kernel<double> (FP64 forward, 1 step per launch)"),
                           "steps_per_launch": steps_per_launch,
                           "note": ("algorithmic bytes count 16 B per cell-update (read t, "
                                    "write t+1) for every fused step, so frac > 1 is possible "
                                    "with k > 1; traffic is the measured HBM bytes per launch"),
                           "algorithmic_bytes_per_launch": alg_bytes,
                           "avg_launch_ms": per_launch_ms}
    out["config"] = {"workload": f"sphere_noise{a.n}: {a.n}^3 cells, {a.steps} steps, FP64, "
                                 "CFL 0.5, seeded mixed-sign velocity, sphere+0.01 noise",
                     "num_cells": list(w.num_cells), "velocity": list(w.velocity), "dt": w.dt}
    out["variants"] = res
    out["parity"] = parity
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
        if out["cpu_baseline"].get("value") and out.get("value"):
            out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
