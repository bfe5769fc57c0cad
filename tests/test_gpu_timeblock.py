"""Temporally blocked passes (several steps per launch) against the oracle.

A pass of K fused steps must give the same bits as K separate steps: each
cell value at each level is produced by the same operations on the same
inputs.  Shapes include partial tiles on every axis, wrap-around at both
ends, every upwind sign branch and step counts that leave a remainder.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu

SHAPES = [(9, 40, 100), (13, 25, 64), (6, 17, 130), (5, 48, 67)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("signs", list(itertools.product((1, -1), repeat=3)))
def test_time_blocked_fp64_bit_exact(oracle, shape, signs):
    from fidibench_amd import Upwind, configs as K
    v = tuple(s * m for s, m in zip(signs, (0.3, 0.45, 0.2)))
    dt = K.cfl_dt(shape, v)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=sum(shape) + 3)
    steps = 11
    ref = oracle.advect(f0, v, (1, 1, 1), dt, steps)
    for k in (2, 3, 4, 6, 7, 8):
        up = Upwind(v, (1, 1, 1), shape)
        up.set_time_block(k)
        up.set_field(f0)
        ms, launches = up.advect_timed(steps, dt)
        assert np.array_equal(bits(up.get_field()), bits(ref)), (shape, signs, k)
        assert launches < steps  # fused passes actually ran
        up.close()


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16", "bf16sr"])
@pytest.mark.parametrize("steps", [9, 23])
def test_auto_time_block_runs_the_plan_bit_exact(dtype, steps):
    """time_block 0 (automatic): advect launches exactly the planner's
    passes (fdb_plan_passes) and ends on the same bits as one launch per
    step; the partition changes no value."""
    from fidibench_amd import Upwind, configs as K
    shape = (21, 70, 130)
    v = (0.3, -0.45, 0.2)
    dt = K.cfl_dt(shape, v)
    f0 = np.random.default_rng(steps).random(shape)
    one = Upwind(v, (1, 1, 1), shape, dtype=dtype)
    one.set_field(f0)
    one.advect(steps, dt)
    auto = Upwind(v, (1, 1, 1), shape, dtype=dtype)
    auto.set_field(f0)
    auto.set_time_block(0)
    plan = auto.plan_passes(steps, dt, 0)
    ms, launches = auto.advect_timed(steps, dt)
    assert sum(plan) == steps and launches == len(plan), (plan, launches)
    assert max(plan) > 1  # fused passes were planned
    assert np.array_equal(bits(auto.get_field()), bits(one.get_field())), (dtype, steps, plan)
    one.close()
    auto.close()


@pytest.mark.parametrize("shape", SHAPES[:2])
@pytest.mark.parametrize("signs", list(itertools.product((1, -1), repeat=3)))
def test_time_blocked_bf16_bit_exact(oracle, shape, signs):
    """bf16 storage, FP32 arithmetic, every level rounded to nearest even
    through bf16: the same bits as the oracle's one-step restatement."""
    from fidibench_amd import Upwind, configs as K
    v = tuple(s * m for s, m in zip(signs, (0.3, 0.45, 0.2)))
    dt = K.cfl_dt(shape, v)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=sum(shape) + 5)
    for k in (2, 3, 4, 6, 7, 8):
        up = Upwind(v, (1, 1, 1), shape, dtype="bf16")
        up.set_field(f0)
        start = up.get_field()
        ref = oracle.advect(start, v, (1, 1, 1), dt, 11)  # uint16 bits: bf16
        up.set_time_block(k)
        ms, launches = up.advect_timed(11, dt)
        assert np.array_equal(bits(up.get_field()), bits(ref)), (shape, signs, k)
        assert launches < 11  # fused passes actually ran
        up.close()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_time_blocked_reduced_precision(oracle, dtype):
    from fidibench_amd import Upwind, configs as K
    shape, v = (11, 36, 96), (-0.35, 0.25, -0.3)
    dt = K.cfl_dt(shape, v)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=8)
    up = Upwind(v, (1, 1, 1), shape, dtype=dtype)
    up.set_field(f0)
    start = up.get_field()
    ref = oracle.advect(start, v, (1, 1, 1), dt, 10)
    up.set_time_block(4)
    up.advect(10, dt)
    assert np.array_equal(bits(up.get_field()), bits(ref))


def test_time_blocked_large_checksum(manifest):
    """512^3 x 100 with 4 fused steps per pass against the reference."""
    from fidibench_amd import Upwind, configs as K
    m = manifest["large"]["sphere_noise512"]
    w = K.sphere_noise512()
    up = Upwind(w.velocity, w.lengths, w.num_cells)
    up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
    up.set_time_block(4)
    ms, launches = up.advect_timed(w.steps, w.dt)
    assert launches == 25
    assert repr(up.checksum()) == m["checksum"]


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("k", [4, 6])
@pytest.mark.parametrize("n0", [7, 8])
@pytest.mark.parametrize("blocks", ["1", "100000"])
def test_two_plane_march_both_parities(oracle, monkeypatch, dtype, k, n0, blocks):
    """K <= 6 marches two planes per loop trip with a one-plane tail.  A
    pass marches chunk + K planes; one chunk (chunk = n0) or one plane per
    chunk (chunk = 1) gives both an even and an odd count across these
    cases, and every one must equal the oracle bit for bit."""
    from fidibench_amd import Upwind, configs as K
    size = {"f64": 8, "f32": 4, "bf16": 2}[dtype]
    monkeypatch.setenv(f"FDB_TB_CFG_{size}_{k}", f"9,{8 if k >= 6 else 4},{blocks}")
    shape = (n0, 30, 70)
    v = (-0.35, 0.4, -0.25)
    dt = K.cfl_dt(shape, v)
    up = Upwind(v, (1, 1, 1), shape, dtype=dtype)
    up.set_field(oracle.uniform(shape, 0.0, 1.0, seed=n0 + k))
    start = up.get_field()
    ref = oracle.advect(start, v, (1, 1, 1), dt, 2 * k + 1)
    up.set_time_block(k)
    ms, launches = up.advect_timed(2 * k + 1, dt)
    assert launches < 2 * k + 1
    assert np.array_equal(bits(up.get_field()), bits(ref)), (dtype, k, n0, blocks)
    up.close()


# Odd row lengths: bf16 rows then start on 2-byte (and FP32 rows on 4-byte)
# boundaries inside the fused kernel's raw buffer accesses; n2 = 1 and 3
# make every lane of a tile wrap onto the same few columns.
ODD = [(5, 48, 67), (6, 17, 131), (4, 21, 3), (7, 9, 1)]


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("shape", ODD)
@pytest.mark.parametrize("signs", [(1, 1, 1), (-1, 1, -1), (1, -1, 1), (-1, -1, -1)])
def test_time_blocked_odd_rows_bit_exact(oracle, dtype, shape, signs):
    from fidibench_amd import Upwind, configs as K
    v = tuple(s * m for s, m in zip(signs, (0.3, 0.45, 0.2)))
    dt = K.cfl_dt(shape, v)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=sum(shape) + 11)
    for k in (2, 3, 4, 6, 7, 8):
        up = Upwind(v, (1, 1, 1), shape, dtype=dtype)
        up.set_field(f0)
        start = up.get_field()
        ref = oracle.advect(start, v, (1, 1, 1), dt, 11)
        up.set_time_block(k)
        ms, launches = up.advect_timed(11, dt)
        assert np.array_equal(bits(up.get_field()), bits(ref)), (dtype, shape, signs, k)
        assert launches < 11
        up.close()


@pytest.mark.parametrize("time_block", [1, 8])
def test_device_pointer_bf16_sr_continues_one_run(oracle, time_block):
    """fdb_advect_dev_seeded with FDB_BF16_SR keys step i of a call by
    (seed, step0 + i): two calls of 4 and 5 steps equal one 9-step oracle
    run with that seed, bit for bit, for any time_block."""
    import ctypes as C
    import torch
    from fidibench_amd import _lib as L, configs as K
    shape, v, seed = (9, 20, 70), (0.4, -0.3, 0.25), 987654321
    dt = K.cfl_dt(shape, v)
    f0 = oracle.f32_to_bf16(oracle.uniform(shape, 0.0, 1.0, seed=4)).reshape(shape)
    ref = oracle.advect_bf16_sr(f0, v, (1, 1, 1), dt, 9, seed=seed, step0=0)
    a = torch.from_numpy(f0.view(np.int16).copy()).cuda()
    b = torch.empty_like(a)
    res = C.c_int()
    s = torch.cuda.current_stream().cuda_stream
    step0 = 0
    for n in (4, 5):
        L.check(L.lib().fdb_advect_dev_seeded(3, L.u64s(shape), L.dbls(v), L.dbls((1, 1, 1)), dt,
                                              L.BF16_SR, a.data_ptr(), b.data_ptr(), n,
                                              time_block, seed, step0, s, C.byref(res)))
        if res.value:
            a, b = b, a
        step0 += n
    torch.cuda.synchronize()
    got = a.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, ref)
    # the unseeded entry point is seed 0, step 0
    ref0 = oracle.advect_bf16_sr(f0, v, (1, 1, 1), dt, 3, seed=0, step0=0)
    a = torch.from_numpy(f0.view(np.int16).copy()).cuda()
    L.check(L.lib().fdb_advect_dev(3, L.u64s(shape), L.dbls(v), L.dbls((1, 1, 1)), dt, L.BF16_SR,
                                   a.data_ptr(), b.data_ptr(), 3, time_block, s, C.byref(res)))
    torch.cuda.synchronize()
    got0 = (b if res.value else a).cpu().numpy().view(np.uint16)
    assert np.array_equal(got0, ref0)


@pytest.mark.parametrize("shape", [(9, 40, 100), (13, 25, 64), (6, 17, 130), (5, 48, 67),
                                   (4, 21, 2), (7, 66, 33), (3, 2, 50)])
@pytest.mark.parametrize("signs", list(itertools.product((1, -1), repeat=3)))
def test_time_blocked_bf16_sr_bit_exact(oracle, shape, signs):
    """Stochastically rounded bf16 in fused passes (K = 2, 4, 8 with an even
    axis-1 length, the axis along which cells pair; any row length, odd ones
    included): every level draws the bits of its absolute step, so any
    time-blocking depth equals the oracle's one-step restatement bit for
    bit; odd axis-1 lengths take one-step launches."""
    from fidibench_amd import Upwind, configs as K
    v = tuple(s * m for s, m in zip(signs, (0.3, 0.45, 0.2)))
    dt = K.cfl_dt(shape, v)
    f0 = oracle.uniform(shape, 0.0, 1.0, seed=sum(shape) + 23)
    seed = 0xDEADBEEF12345 + sum(shape)
    for k in (1, 2, 4, 6, 8):
        up = Upwind(v, (1, 1, 1), shape, dtype="bf16sr")
        up.set_seed(seed)
        up.set_field(f0)
        start = up.get_field()
        ref = oracle.advect_bf16_sr(start, v, (1, 1, 1), dt, 11, seed=seed, step0=0)
        up.set_time_block(k)
        ms, launches = up.advect_timed(5, dt)  # a run in two calls: steps 0-4, 5-10
        ms, launches2 = up.advect_timed(6, dt)
        assert np.array_equal(bits(up.get_field()), bits(ref)), (shape, signs, k)
        if k >= 2 and shape[1] % 2 == 0 and shape[1] > k:
            assert launches + launches2 < 11  # fused passes ran
        up.close()
