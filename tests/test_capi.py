"""The C ABI on the host: exported symbols, host-side helpers (checksum, std,
saveVTK bytes), the checkpoint schedule, argument errors and the absence of
fused multiply-adds in the FP64 kernels.  Runs without a GPU."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import sys
from math import comb

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_case

HEADERS = [os.path.join(ROOT, "include", "fidibench_amd", h)
           for h in ("fdb_upwind.h", "fdb_filter.h")]


def _lib():
    from fidibench_amd import _lib as L
    return L


def declared_functions():
    names = set()
    for h in HEADERS:
        src = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(fdb_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    L = _lib()
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_library_has_gfx950_code_and_no_torch_dependency():
    L = _lib()
    out = subprocess.run(["ldd", L.LIB_PATH], capture_output=True, text=True).stdout
    assert "libtorch" not in out and "libc10" not in out
    blob = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_host_checksum_and_std_match_reference(manifest):
    L = _lib()
    for name, m in manifest["small"].items():
        _, fT = load_case(name)
        f = np.ascontiguousarray(fT, dtype=np.float64)
        p = f.ctypes.data_as(C.POINTER(C.c_double))
        assert repr(L.lib().fdb_host_checksum(p, f.size)) == m["checksum"], name
        assert repr(L.lib().fdb_host_std(p, f.size)) == m["std"], name


@pytest.mark.parametrize("name", ["vtk3d", "vtk2d", "vtk1d"])
def test_save_vtk_bytes_match_reference(manifest, tmp_path, name):
    L = _lib()
    m = manifest["vtk"][name]
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        fT = np.ascontiguousarray(z["fT"])
    out = tmp_path / "o.vtk"
    L.check(L.lib().fdb_host_save_vtk(len(m["shape"]), L.u64s(m["shape"]), L.dbls(m["lengths"]),
                                      fT.ctypes.data_as(C.POINTER(C.c_double)),
                                      str(out).encode()))
    assert out.read_bytes() == open(os.path.join(GOLDEN, m["file"]), "rb").read()


def test_save_vtk_unwritable_path_is_io_error():
    L = _lib()
    f = np.zeros(8)
    rc = L.lib().fdb_host_save_vtk(1, L.u64s([8]), L.dbls([1.0]),
                                   f.ctypes.data_as(C.POINTER(C.c_double)),
                                   b"/nonexistent_dir/x.vtk")
    assert rc == L.ERR_IO and b"cannot open" in L.lib().fdb_last_error()


# ------------------------------------------------------ checkpoint schedule

def _plan(T, snaps):
    L = _lib()
    cap = 20 * T * T + 100
    acts = (C.c_int * cap)()
    args = (C.c_int * cap)()
    fw = C.c_int64()
    n = L.lib().fdb_ckpt_plan(T, snaps, acts, args, cap, C.byref(fw))
    assert 0 <= n <= cap
    return [(acts[i], args[i]) for i in range(n)], fw.value


def _simulate(T, snaps, plan):
    """Run the schedule on integer 'states' and check every adjoint step
    sees the right state, once, in reverse order, within the snapshot
    budget.  Returns the number of forward steps."""
    L = _lib()
    stack = [0]
    work = 0
    fwd = 0
    seen = []
    for a, arg in plan:
        if a == L.CK_ADVANCE:
            assert arg >= 1
            work += arg
            fwd += arg
        elif a == L.CK_STORE:
            stack.append(work)
            assert len(stack) - 1 <= snaps
        elif a == L.CK_RESTORE:
            work = stack[-1]
        elif a == L.CK_POP:
            stack.pop()
            assert stack
        elif a == L.CK_ADJOINT:
            assert work == arg
            seen.append(arg)
        else:
            raise AssertionError(a)
    assert seen == list(range(T - 1, -1, -1))
    return fwd


def _brute(T, snaps, memo={}):
    """Optimal forward-step count of the recursive schedule, by DP."""
    if (T, snaps) in memo:
        return memo[(T, snaps)]
    if T == 1:
        r = 0
    elif snaps == 0:
        r = T * (T - 1) // 2
    else:
        r = min(m + _brute(T - m, snaps - 1) + _brute(m, snaps) for m in range(1, T))
    memo[(T, snaps)] = r
    return r


@pytest.mark.parametrize("T", [1, 2, 3, 7, 16, 33, 50])
@pytest.mark.parametrize("snaps", [0, 1, 2, 3, 5, 60])
def test_checkpoint_plan_is_valid_and_optimal(T, snaps):
    plan, fw = _plan(T, snaps)
    assert _simulate(T, min(snaps, max(T - 1, 0)), plan) == fw
    assert fw == _brute(T, min(snaps, T - 1))
    if 0 < snaps < T - 1:
        s = snaps + 1
        r = 0
        while comb(s + r, s) < T:
            r += 1
        assert fw == r * T - comb(s + r, s + 1)


def test_large_plan_repetition_two():
    """400 steps with 165 snapshots (the 1024x1024x768 BF16 case) needs r=2."""
    plan, fw = _plan(400, 165)
    assert _simulate(400, 165, plan) == fw
    assert fw == 2 * 400 - comb(168, 167)


# ------------------------------------------------------------ arguments

def test_invalid_arguments_are_reported():
    L = _lib()
    h = C.c_void_p()
    rc = L.lib().fdb_upwind_create(4, L.dbls([1] * 4), L.dbls([1] * 4), L.u64s([2] * 4), 0,
                                   C.byref(h))
    assert rc == L.ERR_INVALID and b"ndims" in L.lib().fdb_last_error()
    rc = L.lib().fdb_upwind_create(3, L.dbls([1] * 3), L.dbls([1] * 3), L.u64s([2, 0, 2]), 0,
                                   C.byref(h))
    assert rc == L.ERR_INVALID
    rc = L.lib().fdb_upwind_create(3, L.dbls([1] * 3), L.dbls([1] * 3), L.u64s([2] * 3), 7,
                                   C.byref(h))
    assert rc == L.ERR_INVALID and b"dtype" in L.lib().fdb_last_error()
    rc = L.lib().fdb_upwind_advect(None, 3, 0.1)
    assert rc == L.ERR_INVALID
    rc = L.lib().fdb_tape_create(3, L.u64s([4, 4, 4]), L.dbls([1] * 3), L.dbls([1] * 3), 0.1, 0,
                                 0, 0, 1 << 20, C.byref(h))
    assert rc == L.ERR_INVALID


def test_missing_library_raises(monkeypatch, tmp_path):
    """No fallback: without the built library every entry point raises."""
    from fidibench_amd import Filter, Upwind
    L = _lib()
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "libfdb_upwind.so"))
    monkeypatch.setattr(L, "_lib", None)
    with pytest.raises(FileNotFoundError, match="is missing"):
        L.lib()
    with pytest.raises(FileNotFoundError):
        Upwind((1.0, 1.0, 1.0), (1.0, 1.0, 1.0), (8, 8, 8))
    with pytest.raises(FileNotFoundError):
        Filter((8, 8), (0.0, 0.0), (1.0, 1.0), {(0, 0): 1.0})


def test_create_without_gpu_is_nodev():
    torch = pytest.importorskip("torch")
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    L = _lib()
    h = C.c_void_p()
    rc = L.lib().fdb_upwind_create(3, L.dbls([1] * 3), L.dbls([1] * 3), L.u64s([8] * 3), 0,
                                   C.byref(h))
    assert rc == L.ERR_NODEV


# --------------------------------------------------------- code generation

def test_fp64_kernels_have_no_fused_multiply_add(tmp_path):
    """-ffp-contract=off must hold: no v_fma_f64 / v_fmac_f64 in the FP64
    one-step kernels (forward and adjoint) nor in any FP64 instance of the
    temporally blocked kernel (plain, store-all and fused-adjoint modes),
    and no packed FP32 instructions anywhere."""
    src = os.path.join(ROOT, "fidibench_amd", "csrc", "fdb_kernels.hip")
    asm = tmp_path / "k.s"
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
                    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "--cuda-device-only", "-S", "-o", str(asm), src],
                   check=True, capture_output=True)
    text = asm.read_text()
    bodies = re.findall(r"^(_ZN3fdb\S*(?:rows|march|tb)_kernelId\S*):[^\n]*\n(.*?)s_endpgm",
                        text, re.M | re.S)
    rows = [n for n, _ in bodies if "rows_kernel" in n]
    fused = [n for n, _ in bodies if "tb_kernel" in n]
    assert len(rows) >= 4 and len(fused) >= 30, (len(rows), len(fused))
    for name, body in bodies:
        assert not re.search(r"v_fmac?_f64", body), name
    assert not re.search(r"v_pk_(add|mul|fma)_f32", text)


# ------------------------------------------------------- tuning variables

_CREATE = r"""
import ctypes as C, sys
sys.path.insert(0, %(root)r)
from fidibench_amd import _lib as L
h = C.c_void_p()
rc = L.lib().fdb_upwind_create(3, L.dbls([1] * 3), L.dbls([1] * 3), L.u64s([8] * 3), 0,
                               C.byref(h))
print(rc, L.lib().fdb_last_error().decode())
if rc == 0:
    L.lib().fdb_upwind_destroy(h)
"""


def _create_with_env(extra):
    env = dict(os.environ, **extra)
    out = subprocess.run([sys.executable, "-c", _CREATE % {"root": ROOT}], env=env,
                         capture_output=True, text=True, timeout=300, check=True)
    rc, _, msg = out.stdout.strip().splitlines()[-1].partition(" ")
    return int(rc), msg


@pytest.mark.parametrize("var,val", [
    ("FDB_TB_CFG_8_6", "7,3,0"),        # no (7 rows, K=6, 3 waves) instance
    ("FDB_TB_CFG_4_8", "9,8"),           # too few fields
    ("FDB_ROWS_CFG_8", "16,2,0,1"),      # no 16-row instance
    ("FDB_ROWS_CFG", "8,2,0,0"),         # cached (non-nontemporal) stores: not instantiated
    ("FDB_ADJ_CFG_4_2", "4,4,0,1,x"),    # trailing garbage
])
def test_bad_tuning_variable_is_invalid_argument(var, val):
    """A tuning variable naming a launch shape without a kernel is an
    argument error, reported before any device work (no GPU needed)."""
    L = _lib()
    rc, msg = _create_with_env({var: val})
    assert rc == L.ERR_INVALID and var in msg, (rc, msg)


def test_valid_tuning_variable_is_accepted():
    L = _lib()
    rc, msg = _create_with_env({"FDB_TB_CFG_8_6": "9,8,1024", "FDB_TB_CFG_8_4": "9,4,0",
                                "FDB_ROWS_CFG_4": "8,1,0,1"})
    assert rc in (L.OK, L.ERR_NODEV), (rc, msg)


def test_byte_count_overflow_is_invalid_argument():
    """ntot * itemsize beyond size_t is rejected before any allocation
    (2^62 cells of 8 bytes would wrap to 0 bytes)."""
    L = _lib()
    lib = L.lib()
    n = [1 << 21, 1 << 21, 1 << 20]
    h = C.c_void_p()
    assert lib.fdb_upwind_create(3, L.dbls([1] * 3), L.dbls([1] * 3), L.u64s(n), L.F64,
                                 C.byref(h)) == L.ERR_INVALID
    assert b"overflows" in lib.fdb_last_error()
    assert lib.fdb_tape_create(3, L.u64s(n), L.dbls([1] * 3), L.dbls([1] * 3), 1e-3, 4, L.F64,
                               L.F64, 1 << 30, C.byref(h)) == L.ERR_INVALID
    assert b"overflows" in lib.fdb_last_error()
    offs = (C.c_int * 3)(0, 0, 0)
    w = (C.c_double * 1)(1.0)
    assert lib.fdb_filter_create(3, L.u64s([(1 << 31) - 1, (1 << 31) - 1, 1]), L.dbls([0] * 3),
                                 L.dbls([1] * 3), 1, offs, w, C.byref(h)) == L.ERR_INVALID
    assert b"overflows" in lib.fdb_last_error()


# ------------------------------------------------------------ raw headers

def _raw_bytes(ndims=3, dtype=0, n=(4, 5, 6), steps=7, seed=99, payload=None, magic=b"FDBFLD01"):
    import struct
    head = magic + struct.pack("<II3QQQQ", ndims, dtype, *n, steps, seed, 0)
    assert len(head) == 64
    size = {0: 8, 1: 4, 2: 2, 3: 2}.get(dtype, 8)
    cells = 1
    for x in n:
        cells *= x
    return head + (payload if payload is not None else bytes(cells * size))


@pytest.mark.parametrize("kind,expect", [
    ("good", 0), ("good_bf16sr", 0), ("good_1d", 0), ("truncated", -4), ("longer", -4),
    ("foreign", -4), ("short_header", -4), ("bad_ndims", -4), ("bad_dtype", -4),
    ("zero_axis", -4), ("extra_axis", -4), ("huge_axes", -4), ("missing", -4)])
def test_raw_header_checks_without_a_device(tmp_path, kind, expect):
    """fdb_host_raw_info applies load_raw's header checks on the host: a
    missing, foreign, malformed or wrongly sized file is FDB_ERR_IO."""
    L = _lib()
    data = {
        "good": _raw_bytes(),
        "good_bf16sr": _raw_bytes(dtype=3, n=(3, 2, 1), seed=2**63 + 5),
        "good_1d": _raw_bytes(ndims=1, dtype=1, n=(10, 1, 1)),
        "truncated": _raw_bytes()[:-1],
        "longer": _raw_bytes() + b"\0",
        "foreign": b"PK\x03\x04" + _raw_bytes()[4:],
        "short_header": b"FDBFLD01" + bytes(20),
        "bad_ndims": _raw_bytes(ndims=4),
        "bad_dtype": _raw_bytes(dtype=7),
        "zero_axis": _raw_bytes(n=(4, 0, 6), payload=b""),
        "extra_axis": _raw_bytes(ndims=2, n=(4, 5, 6)),
        "huge_axes": _raw_bytes(n=(1 << 40, 1 << 30, 1), payload=b""),
    }.get(kind)
    path = tmp_path / "f.raw"
    if data is not None:
        path.write_bytes(data)
    nd, dt = C.c_int(-1), C.c_int(-1)
    n = (C.c_uint64 * 3)()
    st, sd = C.c_uint64(), C.c_uint64()
    rc = L.lib().fdb_host_raw_info(str(path).encode(), C.byref(nd), C.byref(dt), n, C.byref(st),
                                   C.byref(sd))
    assert rc == expect, (kind, rc, L.lib().fdb_last_error())
    if kind == "good":
        assert (nd.value, dt.value, list(n), st.value, sd.value) == (3, 0, [4, 5, 6], 7, 99)
    if kind == "good_bf16sr":
        assert (dt.value, sd.value) == (3, 2**63 + 5)
    if kind == "good_1d":
        assert (nd.value, n[0]) == (1, 10)


@pytest.mark.parametrize("snaps", [1, 2, 3, 5])
def test_plan_400_steps_few_snapshots(snaps):
    """The 400-step schedule with few snapshots: valid (every adjoint sees
    its state, within the budget) and at the closed-form binomial optimum
    r*T - C(s+r, s+1), s = snaps + 1."""
    T = 400
    plan, fw = _plan(T, snaps)
    assert _simulate(T, snaps, plan) == fw
    s = snaps + 1
    r = 0
    while comb(s + r, s) < T:
        r += 1
    assert fw == r * T - comb(s + r, s + 1), (snaps, r, fw)


# ------------------------------------------------------------ pass planner

SN512 = ((512, 512, 512), (0.7, -0.4, 0.2), (1.0, 1.0, 1.0))


def _passes(dtype, nsteps, tb, geom=SN512, dt=1e-4):
    L = _lib()
    shape, v, lens = geom
    return L.plan_passes(shape, v, lens, dt, dtype, nsteps, tb)


def test_auto_plan_of_twenty_fp64_steps_is_seven_seven_six():
    """20 FP64 steps at 512^3: 7+7+6 (three passes of near-best depth per
    step) rather than 8+8+4, whose 4-step pass costs as much as a 6-step
    one (launch times measured on MI355X)."""
    L = _lib()
    assert sorted(_passes(L.F64, 20, 0), reverse=True) == [7, 7, 6]


def test_auto_plan_scales_launch_times_to_larger_grids():
    """From 512^3 cells up the planner scales each depth's launch time by the
    pass's length on the grid (rounds of resident blocks x planes marched
    per block): at 1024^3 FP64 runs mostly 8-step passes (679 us per step
    measured on MI355X against 710 for 7-step ones), at 512^3 7-step passes
    (86 us per step against 91)."""
    L = _lib()
    big = ((1024, 1024, 1024), (0.7, -0.4, 0.2), (1.0, 1.0, 1.0))
    ks = _passes(L.F64, 100, 0, big)
    assert sum(ks) == 100 and ks.count(8) >= 9, ks
    ks = _passes(L.F64, 100, 0)
    assert ks.count(7) >= 10, ks


def test_auto_plan_on_small_grids_follows_measured_times():
    """Below 512^3 the planner interpolates launch times measured at 16^3,
    128^3 and 256^3 (latency-bound marches of few planes): 4-step passes at
    16^3 and 21x70x130 (FP64 2.5 us per step against 3.3 for 7-step ones),
    FP32 4-step passes at 128^3 (4.1 us per step against 5.6 for 8-step
    ones), bf16 with stochastic rounding one step per launch there (5.9 us
    against 6.2 for 4-step passes)."""
    L = _lib()
    v, lens = (0.3, -0.45, 0.2), (1.0, 1.0, 1.0)
    assert _passes(L.F64, 20, 0, ((16, 16, 16), v, lens)) == [4] * 5
    assert _passes(L.F64, 20, 0, ((21, 70, 130), v, lens)) == [4] * 5
    assert _passes(L.F32, 100, 0, ((128, 128, 128), v, lens)) == [4] * 25
    assert _passes(L.BF16_SR, 10, 0, ((128, 128, 128), v, lens)) == [1] * 10


@pytest.mark.parametrize("dtype", ["F64", "F32", "BF16", "BF16_SR"])
@pytest.mark.parametrize("nsteps", [1, 2, 5, 9, 13, 20, 47, 48, 49, 100, 400])
def test_auto_plan_covers_every_step_with_fused_depths(dtype, nsteps):
    L = _lib()
    ks = _passes(getattr(L, dtype), nsteps, 0)
    assert sum(ks) == nsteps
    allowed = {1, 2, 4, 8} if dtype == "BF16_SR" else {1, 2, 3, 4, 6, 7, 8}
    assert set(ks) <= allowed, ks
    if nsteps >= 8:  # long runs are mostly fused passes of each type's best depth
        deep = 4 if dtype == "BF16_SR" else 6  # (bf16-SR: 4 steps, 87 us per step at 512^3)
        assert sum(k for k in ks if k >= deep) >= nsteps - 8, ks


def test_auto_plan_takes_few_launches():
    """The automatic split of n FP64 steps at 512^3 takes at most as many
    launches as passes of 6 steps would (a one-step launch, about half a
    fused pass's time, may close it: 9 steps = 8 + 1)."""
    L = _lib()
    for n in range(9, 120):
        auto = _passes(L.F64, n, 0)
        assert len(auto) <= -(-n // 6), (n, auto)
    assert _passes(L.F64, 9, 0) == [8, 1]


def test_fixed_time_block_plan_splits_unsupported_depths():
    """time_block k: passes of min(left, k) steps, each launched as the
    deepest supported fused depth first (no 5-step kernel: 4 + 1)."""
    L = _lib()
    assert _passes(L.F64, 12, 5) == [4, 1, 4, 1, 2]
    assert _passes(L.F64, 10, 1) == [1] * 10
    assert _passes(L.F64, 17, 8) == [8, 8, 1]


def test_plan_on_a_grid_without_fused_passes_is_one_step_launches():
    """A 1-D grid has no fused form: every step is a one-step launch."""
    L = _lib()
    geom = ((1000,), (0.5,), (1.0,))
    assert _passes(L.F64, 9, 0, geom) == [1] * 9


@pytest.mark.parametrize("tb", [-1, 9])
def test_plan_time_block_out_of_range_is_invalid(tb):
    L = _lib()
    with pytest.raises(L.FdbError) as e:
        _passes(L.F64, 10, tb)
    assert e.value.code == L.ERR_INVALID
