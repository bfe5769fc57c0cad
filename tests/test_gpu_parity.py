"""FP64 parity of the HIP path with the reference (bit for bit).

Golden vectors come from the reference's own upwind.cxx (tests/golden);
the oracle (oracle/upwind_oracle.c, itself pinned to those vectors by
test_oracle.py) checks cases the fixtures do not hold.  Everything runs
through the C ABI (libfdb_upwind.so) via fidibench_amd.Upwind.
"""
from __future__ import annotations

import hashlib
import itertools

import numpy as np
import pytest

from conftest import bits, load_case

pytestmark = pytest.mark.gpu


def _upwind(shape, v, lengths, dtype="f64"):
    from fidibench_amd import Upwind
    return Upwind(v, lengths, shape, dtype=dtype)


def _sha(f):
    return hashlib.sha256(np.ascontiguousarray(f, dtype="<f8").tobytes()).hexdigest()


def test_library_is_native():
    import fidibench_amd
    from fidibench_amd import _lib
    assert fidibench_amd.lib()._name == _lib.LIB_PATH


@pytest.mark.parametrize("time_block", [1, 3])
def test_small_goldens_bit_exact(manifest, time_block):
    for name, m in manifest["small"].items():
        f0, fT = load_case(name)
        up = _upwind(m["shape"], m["velocity"], m["lengths"])
        up.set_time_block(time_block)
        if m["init"][0] != "delta":  # delta: keep the ctor's own f[0]=1 field
            up.set_field(f0)
        up.advect(m["steps"], float(m["dt"]))
        got = up.get_field()
        assert np.array_equal(bits(got), bits(fT)), name
        assert repr(up.checksum()) == m["checksum"], name
        assert repr(up.std()) == m["std"], name
        up.close()


def test_ctor_field_is_reference_delta():
    up = _upwind((5, 6, 7), (1, 1, 1), (1, 1, 1))
    f = up.get_field()
    assert f.flat[0] == 1.0 and np.count_nonzero(f) == 1


def test_one_step_overload_and_zero_steps(manifest):
    m = manifest["small"]["c161220"]
    f0, _ = load_case("c161220")
    a = _upwind(m["shape"], m["velocity"], m["lengths"])
    b = _upwind(m["shape"], m["velocity"], m["lengths"])
    a.set_field(f0)
    b.set_field(f0)
    dt = float(m["dt"])
    for _ in range(3):
        a.advect(dt)  # advect(double): one step
    b.advect(3, dt)
    assert np.array_equal(bits(a.get_field()), bits(b.get_field()))
    b.advect(0, dt)
    b.advect(-4, dt)  # the reference loop runs no step for <= 0
    assert np.array_equal(bits(a.get_field()), bits(b.get_field()))


@pytest.mark.parametrize("shape,v,steps", [
    ((17, 33, 130), (0.31, -0.27, 0.4), 6),
    ((64, 64, 64), (-0.9, -0.05, -0.05), 9),
    ((9, 1, 257), (0.5, 0.1, -0.4), 5),
    ((40, 50, 2), (-0.2, 0.7, 0.1), 4),
])
def test_against_oracle_other_shapes(oracle, shape, v, steps):
    from fidibench_amd import configs as K
    f0 = oracle.uniform(shape, -1.0, 2.0, seed=sum(shape))
    dt = K.cfl_dt(shape, v)
    ref = oracle.advect(f0, v, (1, 1, 1), dt, steps)
    up = _upwind(shape, v, (1, 1, 1))
    up.set_field(f0)
    up.advect(steps, dt)
    assert np.array_equal(bits(up.get_field()), bits(ref))
    assert up.checksum() == oracle.checksum(ref)


@pytest.mark.parametrize("signs", list(itertools.product((1, -1), repeat=3)))
def test_every_upwind_branch(oracle, signs):
    shape = (12, 10, 70)
    v = tuple(s * m for s, m in zip(signs, (0.3, 0.45, 0.2)))
    from fidibench_amd import configs as K
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=99)
    dt = K.cfl_dt(shape, v)
    ref = oracle.advect(f0, v, (1, 1, 1), dt, 5)
    up = _upwind(shape, v, (1, 1, 1))
    up.set_field(f0)
    up.advect(5, dt)
    assert np.array_equal(bits(up.get_field()), bits(ref))


def test_non_unit_lengths(oracle):
    shape, v, L = (10, 12, 14), (0.4, -0.3, 0.6), (2.0, 0.5, 1.25)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=7)
    dt = 0.001
    ref = oracle.advect(f0, v, L, dt, 7)
    up = _upwind(shape, v, L)
    up.set_field(f0)
    up.advect(7, dt)
    assert np.array_equal(bits(up.get_field()), bits(ref))


def test_nan_and_inf_propagate_like_reference(oracle):
    shape, v = (6, 7, 8), (0.3, -0.2, 0.1)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=3)
    f0[2, 3, 4] = np.nan
    f0[4, 1, 7] = np.inf
    dt = 0.01
    ref = oracle.advect(f0, v, (1, 1, 1), dt, 3)
    up = _upwind(shape, v, (1, 1, 1))
    up.set_field(f0)
    up.advect(3, dt)
    got = up.get_field()
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    m = ~np.isnan(ref)
    assert np.array_equal(bits(got[m]), bits(ref[m]))


def test_init_kernel_matches_host_generator(oracle):
    from fidibench_amd import configs as K
    from fidibench_amd import Upwind
    for shape in [(16, 20, 24), (33, 7, 65)]:
        up = Upwind((1, 1, 1), (1, 1, 1), shape)
        up.init_field(K.INIT_SPHERE, (0.5, 0.5, 0.5, 0.3, 0.01), seed=12345)
        host = oracle.sphere_noise(shape, (1, 1, 1), (0.5, 0.5, 0.5), 0.3, 0.01, 12345)
        assert np.array_equal(bits(up.get_field()), bits(host))
        up.init_field(K.INIT_UNIFORM, (-1.0, 3.0), seed=5)
        assert np.array_equal(bits(up.get_field()), bits(oracle.uniform(shape, -1.0, 3.0, 5)))


def _large(manifest, name):
    if name not in manifest.get("large", {}):
        pytest.skip(f"golden {name} not generated")
    return manifest["large"][name]


def test_reference_default_run_128(manifest):
    """upwindCxx defaults: 128^3, 10 steps, v=1, dt=0.1/128, f[0]=1."""
    m = _large(manifest, "ref_default128")
    up = _upwind(m["shape"], m["velocity"], (1, 1, 1))
    up.advect(m["steps"], float(m["dt"]))
    assert repr(up.checksum()) == m["checksum"]
    assert repr(up.std()) == m["std"]
    assert _sha(up.get_field()) == m["sha256"]


@pytest.mark.parametrize("name", ["sphere128", "misfit_target256", "sphere_noise512"])
def test_large_workloads_bit_exact(manifest, name):
    """Full-size runs (up to 512^3 x 100 steps) against the reference's hash
    and checksum, through the device initial-field generator."""
    from fidibench_amd import configs as K
    m = _large(manifest, name)
    w = {"sphere128": K.sphere128(), "misfit_target256": K.misfit_target(),
         "sphere_noise512": K.sphere_noise512()}[name]
    assert tuple(m["shape"]) == w.num_cells and repr(w.dt) == m["dt"]
    up = _upwind(w.num_cells, w.velocity, w.lengths)
    up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
    assert _sha(up.get_field()) == m["sha256_f0"]
    up.set_time_block(4)
    up.advect(w.steps, w.dt)
    assert repr(up.checksum()) == m["checksum"]
    assert _sha(up.get_field()) == m["sha256"]


def test_advect_dev_matches_handle(oracle):
    """The device-pointer entry point on torch-owned buffers."""
    import ctypes as C
    import torch
    from fidibench_amd import _lib as L
    shape, v = (20, 24, 130), (-0.4, 0.35, 0.2)
    from fidibench_amd import configs as K
    dt = K.cfl_dt(shape, v)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=41)
    ref = oracle.advect(f0, v, (1, 1, 1), dt, 7)
    a = torch.from_numpy(f0).cuda()
    b = torch.empty_like(a)
    res = C.c_int()
    L.check(L.lib().fdb_advect_dev(3, L.u64s(shape), L.dbls(v), L.dbls((1, 1, 1)), dt, L.F64,
                                   a.data_ptr(), b.data_ptr(), 7, 2,
                                   torch.cuda.current_stream().cuda_stream, C.byref(res)))
    torch.cuda.synchronize()
    out = (b if res.value else a).cpu().numpy()
    assert np.array_equal(bits(out), bits(ref))


@pytest.mark.parametrize("name", ["vtk3d", "vtk2d", "vtk1d"])
def test_save_vtk_handle_matches_reference_bytes(manifest, tmp_path, name):
    """Upwind.saveVTK after advect, byte-compared with the reference's file."""
    import os
    from conftest import GOLDEN
    m = manifest["vtk"][name]
    f0, _ = load_case(name)
    up = _upwind(m["shape"], m["velocity"], m["lengths"])
    up.set_field(f0)
    up.advect(m["steps"], float(m["dt"]))
    out = tmp_path / "x.vtk"
    up.saveVTK(str(out))
    assert out.read_bytes() == open(os.path.join(GOLDEN, m["file"]), "rb").read()


@pytest.mark.parametrize("name", ["print_c456", "print_special", "print_ctor", "print_1d"])
def test_print_matches_reference_stdout(manifest, capfd, name):
    """print() (upwind.cxx:105-109): "i f[i]" lines through the default
    ostream formatting, byte for byte against the reference's own print()
    captured by tests/golden/gen_golden.py -- after advect steps, on the
    constructor's f[0]=1 field, and on -0, a denormal, +-Inf and NaN."""
    import os
    from conftest import GOLDEN
    from fidibench_amd import Upwind
    m = manifest["print"][name]
    n = tuple(m["shape"])
    with Upwind(m["velocity"], (1.0,) * len(n), n) as up:
        if m["field"] != "ctor":
            with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
                up.set_field(np.ascontiguousarray(z["f0"], dtype=np.float64))
        up.advect(m["steps"], float(m["dt"]))
        capfd.readouterr()
        up.print()
        out = capfd.readouterr().out
    assert out == m["stdout"]
