"""Writes the golden fixtures under tests/golden/ from the REAL reference.

The reference's own upwind.cxx is compiled where it lies by oracle/Makefile
into oracle/_ref/libupwind_ref.so (an include-harness, oracle/ref_harness.cpp),
so this script only runs where the reference checkout exists.  Fixtures are
data only: input fields, output fields, checksums (repr, i.e. %.17g),
std, SHA-256 of the final field's little-endian bytes, VTK files written by
the reference's saveVTK, and the reference executable's stdout.

    python tests/golden/gen_golden.py [--large]

--large adds the 128^3 and 512^3 x 100 runs (a few minutes on 8 cores).
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from fidibench_amd import configs as K  # noqa: E402

L1 = (1.0, 1.0, 1.0)


def sha(f: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(f, dtype="<f8").tobytes()).hexdigest()


def ref_main_dt(n, v, lengths):
    """dt rule of the reference executable (upwind.cxx:185-192)."""
    dt = sys.float_info.max
    for j in range(len(n)):
        dx = lengths[j] / n[j]
        val = 0.1 * dx / v[j]
        dt = val if val < dt else dt
    return dt


# name, shape, velocity, init ("uniform", seed) / ("delta",) / ("sphere", r, amp, seed),
# steps, dt rule ("cfl" | "main" | float)
SMALL = [
    ("c456_mixed", (4, 5, 6), (0.7, -0.4, 0.2), ("uniform", 11), 3, "cfl"),
    ("c888_delta_main", (8, 8, 8), (1.0, 1.0, 1.0), ("delta",), 5, "main"),
    ("c161220", (16, 12, 20), (-0.3, 0.9, -0.6), ("uniform", 12), 7, "cfl"),
    ("c322440_sphere", (32, 24, 40), (0.7, -0.4, 0.2), ("sphere", 0.3, 0.1, 13), 7,
     0.5 / 32 / 1.3),
    ("c13_7_65", (13, 7, 65), (0.25, 0.5, -0.75), ("uniform", 14), 5, "cfl"),
    ("c3_130_67", (3, 130, 67), (-0.5, -0.2, 0.9), ("uniform", 15), 4, "cfl"),
    ("c1_7_9", (1, 7, 9), (0.3, 0.6, -0.1), ("uniform", 16), 6, "cfl"),
    ("c_zero_v", (6, 5, 7), (0.0, -0.0, 0.5), ("uniform", 17), 5, "cfl"),
    ("c2_2_128", (2, 2, 128), (0.4, -0.4, 0.9), ("uniform", 18), 8, "cfl"),
    ("c5_9_200", (5, 9, 200), (-0.3, 0.3, -0.3), ("uniform", 19), 3, "cfl"),
    ("c70_3_129", (70, 3, 129), (0.2, 0.2, 0.2), ("uniform", 20), 4, "cfl"),
    ("c_neg_dt_main", (6, 6, 6), (-1.0, 1.0, 1.0), ("uniform", 21), 3, "main"),
    ("c2d_33_17", (33, 17), (0.5, -0.8), ("uniform", 22), 9, "cfl"),
    ("c1d_101", (101,), (-0.6,), ("uniform", 23), 11, "cfl"),
    ("c3d_36_uniformv", (36, 36, 36), (1.0, 1.0, 1.0), ("sphere", 0.15, 0.0, 0), 12, "cfl"),
]


def make_init(spec, n):
    kind = spec[0]
    if kind == "uniform":
        return O.uniform(n, 0.0, 1.0, spec[1])
    if kind == "delta":
        f = np.zeros(n)
        f.flat[0] = 1.0
        return f
    if kind == "sphere":
        return O.sphere_noise(n, L1, (0.5, 0.5, 0.5), spec[1], spec[2], spec[3])
    raise ValueError(kind)


def dt_for(rule, n, v):
    lengths = (1.0,) * len(n)
    if rule == "cfl":
        return K.cfl_dt(n, v, lengths)
    if rule == "main":
        return ref_main_dt(n, v, lengths)
    return float(rule)


def small_cases(manifest):
    for name, n, v, init, steps, rule in SMALL:
        lengths = (1.0,) * len(n)
        dt = dt_for(rule, n, v)
        f0 = make_init(init, n)
        fT, cs, sd = O.ref_run(f0, v, lengths, dt, steps, use_ctor_field=(init[0] == "delta"))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), f0=f0, fT=fT)
        manifest["small"][name] = dict(shape=list(n), velocity=list(v), lengths=list(lengths),
                                       dt=repr(dt), steps=steps, init=list(init),
                                       checksum=repr(cs), std=repr(sd), sha256=sha(fT))
        print("small", name, cs)


def vtk_cases(manifest):
    for name, n, v in [("vtk3d", (3, 4, 5), (0.3, -0.2, 0.1)), ("vtk2d", (4, 6), (0.5, 0.5)),
                       ("vtk1d", (7,), (-0.4,))]:
        lengths = tuple(1.0 + 0.5 * j for j in range(len(n)))
        f0 = O.uniform(n, -1.0, 1.0, 31)
        path = os.path.join(HERE, f"{name}.vtk")
        dt = K.cfl_dt(n, v, lengths)
        fT, cs, _ = O.ref_run(f0, v, lengths, dt, 2, vtk_path=path)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), f0=f0, fT=fT)
        manifest["vtk"][name] = dict(shape=list(n), velocity=list(v), lengths=list(lengths),
                                     dt=repr(dt), steps=2, file=f"{name}.vtk")
        print("vtk", name)


CLI_ARGS = [
    [],
    ["-numCells", "16", "-numSteps", "3", "-std"],
    ["32", "10"],
    ["-numSteps", "5", "-numCells", "24", "-numCells", "20"],
    ["-std", "-std", "-numCells", "8", "-numSteps", "2"],
    ["-numCells", "12", "-numSteps", "0"],
    ["-bogus"],
    ["-h"],
]


def cli_cases(manifest):
    code = ("import sys, ctypes as C; sys.path.insert(0, %r); from oracle import oracle as O; "
            "a=[b'upwindCxx']+[x.encode() for x in sys.argv[1:]]; "
            "arr=(C.c_char_p*len(a))(*a); O.ref().ref_upwind_main(len(a), arr)") % ROOT
    out = []
    for args in CLI_ARGS:
        env = dict(os.environ, OMP_NUM_THREADS="1")
        p = subprocess.run([sys.executable, "-c", code] + args, capture_output=True, text=True,
                           env=env, timeout=600)
        out.append(dict(args=args, stdout=p.stdout, stderr=p.stderr))
        print("cli", args)
    manifest["cli"] = out


# name, shape, velocity, field ("small:<case>" = that case's f0, "special",
# "ctor" = the constructor's f[0]=1 field), steps, dt rule
PRINT = [
    ("print_c456", (4, 5, 6), (0.7, -0.4, 0.2), "small:c456_mixed", 3, "cfl"),
    ("print_special", (1, 2, 5), (0.3, 0.2, 0.1), "special", 0, "cfl"),
    ("print_ctor", (2, 2, 3), (1.0, 1.0, 1.0), "ctor", 2, "main"),
    ("print_1d", (9,), (-0.6,), "small:c1d_101", 4, "cfl"),
]


def print_field(spec, n):
    if spec == "ctor":
        return None
    if spec == "special":
        return np.array([0.0, -0.0, 1e-310, 1e300, -1.5e-7, 123456789.0, np.inf, -np.inf,
                         np.nan, 0.1]).reshape(n)
    with np.load(os.path.join(HERE, spec.split(":", 1)[1] + ".npz"), allow_pickle=False) as z:
        return np.ascontiguousarray(z["f0"], dtype=np.float64).ravel()[: int(np.prod(n))] \
            .reshape(n)


def print_cases(manifest):
    """stdout of the reference's print() (upwind.cxx:105-109), captured in a
    child process (the harness writes to the C++ std::cout)."""
    code = ("import sys, json, ctypes as C, numpy as np; sys.path.insert(0, %r); "
            "from oracle import oracle as O; a = json.loads(sys.argv[1]); "
            "f = None if a['f'] is None else np.load(a['f'], allow_pickle=False); "
            "n = a['n']; U = lambda x: (C.c_uint64 * len(x))(*x); D = lambda x: (C.c_double * len(x))(*x); "
            "p = None if f is None else np.ascontiguousarray(f, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double)); "
            "O.ref().ref_upwind_print(len(n), U(n), D(a['v']), D([1.0] * len(n)), a['dt'], a['steps'], p)"
            ) % ROOT
    manifest["print"] = {}
    for name, n, v, spec, steps, rule in PRINT:
        dt = dt_for(rule, n, v)
        f0 = print_field(spec, n)
        fpath = None
        if f0 is not None:
            np.savez_compressed(os.path.join(HERE, f"{name}.npz"), f0=f0)
            fpath = os.path.join(HERE, f"{name}.npz")
        arg = {"n": list(n), "v": list(v), "dt": dt, "steps": steps,
               "f": None if fpath is None else fpath.replace(".npz", ".npy")}
        if fpath is not None:
            np.save(arg["f"], f0)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        p = subprocess.run([sys.executable, "-c", code, json.dumps(arg)], capture_output=True,
                           text=True, env=env, timeout=600, check=True)
        if fpath is not None:
            os.remove(arg["f"])
        manifest["print"][name] = dict(shape=list(n), velocity=list(v), dt=repr(dt), steps=steps,
                                       field=spec, stdout=p.stdout)
        print("print", name, len(p.stdout.splitlines()), "lines")


def large_cases(manifest):
    runs = []
    n = (128, 128, 128)
    f0 = np.zeros(n)
    f0.flat[0] = 1.0
    runs.append(("ref_default128", n, (1.0, 1.0, 1.0), ref_main_dt(n, (1, 1, 1), L1), 10, f0,
                 True))
    for w in (K.sphere128(), K.misfit_target(), K.sphere_noise512()):
        p = w.init_params
        f0 = O.sphere_noise(w.num_cells, w.lengths, p[:3], p[3], p[4], w.init_seed)
        runs.append((w.name, w.num_cells, w.velocity, w.dt, w.steps, f0, False))
    for name, n, v, dt, steps, f0, ctor in runs:
        fT, cs, sd = O.ref_run(f0, v, L1, dt, steps, use_ctor_field=ctor)
        manifest["large"][name] = dict(shape=list(n), velocity=list(v), dt=repr(dt),
                                       steps=steps, checksum=repr(cs), std=repr(sd),
                                       sha256=sha(fT), sha256_f0=sha(f0),
                                       ctor_field=ctor)
        print("large", name, repr(cs))
        del fT


def main():
    if not O.have_ref():
        O.build()
    path = os.path.join(HERE, "manifest.json")
    manifest = {"small": {}, "vtk": {}, "cli": [], "large": {}}
    if os.path.exists(path):
        with open(path) as fh:
            manifest.update(json.load(fh))
    manifest["small"] = {}
    small_cases(manifest)
    manifest["vtk"] = {}
    vtk_cases(manifest)
    cli_cases(manifest)
    print_cases(manifest)
    if "--large" in sys.argv:
        manifest["large"] = {}
        large_cases(manifest)
    manifest["generator"] = ("tests/golden/gen_golden.py via oracle/_ref/libupwind_ref.so "
                             "(reference upwind.cxx compiled g++ -O3 -fopenmp -ffp-contract=off)")
    with open(path, "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
