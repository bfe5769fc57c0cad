"""FP32 and BF16-storage variants of the forward kernel.

Each variant is bit-exact with the oracle's restatement of the same
arithmetic (float math, coefficient rounded once, bf16 rounded to nearest
even on store); its error against the FP64 path -- max-abs, L2, max-rel
(max|f - f64| / max|f64|) -- and the drift of the compensated FP64 mass
are stated and bounded.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import bits, load_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_reduced_precision_bit_exact_vs_oracle(manifest, oracle, dtype):
    from fidibench_amd import Upwind
    for name in ["c161220", "c13_7_65", "c3_130_67", "c2d_33_17", "c1d_101", "c5_9_200",
                 "c70_3_129"]:
        m = manifest["small"][name]
        f0, _ = load_case(name)
        up = Upwind(m["velocity"], m["lengths"], m["shape"], dtype=dtype)
        up.set_field(f0)  # converted on the host: f64 -> f32 (-> bf16, ties to even)
        start = up.get_field()
        ref = oracle.advect(start, m["velocity"], m["lengths"], float(m["dt"]), m["steps"])
        up.advect(m["steps"], float(m["dt"]))
        assert np.array_equal(bits(up.get_field()), bits(ref)), (name, dtype)
        up.close()


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_special_values_match_oracle(oracle, dtype):
    """Subnormals, values near the bf16/f32 overflow threshold, rounding
    ties, +-Inf and NaN: the device rounding (v_cvt_pk_bf16_f32 for bf16)
    must agree with the oracle's software round-to-nearest-even.  NaN
    payloads differ between CPU and GPU arithmetic, so NaN positions are
    compared, and every other value bit for bit."""
    from fidibench_amd import Upwind
    shape, v = (6, 8, 64), (0.35, -0.25, 0.3)
    rng = np.random.default_rng(11)
    f = rng.standard_normal(shape)
    flat = f.reshape(-1)
    specials = [1e-39, -3e-40, 1e-45, 3.38e38, -3.39e38, 3.3895e38, 1.00390625, 1.01171875,
                1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 65504.0, 2.0 ** -126, -(2.0 ** -133)]
    flat[5:5 + len(specials)] = specials
    flat[300] = np.inf
    flat[1000] = -np.inf
    flat[2000] = np.nan
    up = Upwind(v, (1, 1, 1), shape, dtype=dtype)
    up.set_field(f)
    start = up.get_field()
    ref = oracle.advect(start, v, (1, 1, 1), 0.01, 3)
    up.advect(3, 0.01)
    got = up.get_field()
    wide = (lambda a: oracle.bf16_to_f32(a)) if dtype == "bf16" else (lambda a: a)
    nan_g, nan_r = np.isnan(wide(got)), np.isnan(wide(ref))
    assert np.array_equal(nan_g, nan_r)
    assert np.array_equal(bits(got)[~nan_g], bits(ref)[~nan_r])


@pytest.mark.parametrize("seed", [0, 12345])
def test_bf16_stochastic_rounding_bit_exact_vs_oracle(oracle, seed):
    """Stochastically rounded bf16 steps (dtype bf16sr) against the
    oracle's restatement of the same hash, including the step count
    carried across advect calls."""
    from fidibench_amd import Upwind
    for shape, v in [((9, 12, 130), (0.3, -0.45, 0.2)), ((5, 33, 64), (-0.6, 0.1, -0.25)),
                     ((40, 7), (0.5, -0.3)), ((101,), (-0.7,))]:
        from fidibench_amd import configs as K
        dt = K.cfl_dt(shape, v, (1.0,) * len(shape))
        f0 = oracle.uniform(shape, 0.0, 1.0, seed=4)
        up = Upwind(v, (1.0,) * len(shape), shape, dtype="bf16sr")
        up.set_field(f0)
        up.set_seed(seed)
        start = up.get_field()
        up.advect(3, dt)
        up.advect(4, dt)  # steps 3..6 continue the stream
        ref = oracle.advect_bf16_sr(start, v, (1.0,) * len(shape), dt, 7, seed=seed, step0=0)
        assert np.array_equal(up.get_field(), ref), shape
        up.close()


def test_bf16_stochastic_rounding_accuracy():
    """Gaussian bubble, 192^3 x 200: stochastic rounding removes the
    stagnation bias of round-to-nearest bf16 storage."""
    from fidibench_amd import Upwind, configs as K
    w = K.gauss1024(n=192, steps=200)
    out = {}
    for dt_name in ("f64", "bf16", "bf16sr"):
        up = Upwind(w.velocity, w.lengths, w.num_cells, dtype=dt_name)
        up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
        m0 = up.mass()
        up.advect(w.steps, w.dt)
        out[dt_name] = (up.get_field("f64"), abs(up.mass() - m0) / abs(m0))
        up.close()
    ref = out["f64"][0]
    e_rne, e_sr = _errors(out["bf16"][0], ref), _errors(out["bf16sr"][0], ref)
    print("bf16 RNE", e_rne, "drift", out["bf16"][1], "| bf16 SR", e_sr, "drift", out["bf16sr"][1])
    assert e_sr["l2"] < 0.5 * e_rne["l2"]
    assert out["bf16sr"][1] < 0.1 * out["bf16"][1]


def _errors(x, ref):
    d = x.astype(np.float64) - ref
    return {"max_abs": float(np.abs(d).max()), "l2": float(np.sqrt(np.mean(d * d))),
            "max_rel": float(np.abs(d).max() / np.abs(ref).max())}


def test_mixed_precision_errors_gauss():
    """The Gaussian-bubble mixed-precision workload at 192^3 x 200 steps."""
    from fidibench_amd import Upwind, configs as K
    w = K.gauss1024(n=192, steps=200)
    out = {}
    for dt_name in ("f64", "f32", "bf16"):
        up = Upwind(w.velocity, w.lengths, w.num_cells, dtype=dt_name)
        up.init_field(w.init_kind, w.init_params, seed=w.init_seed)
        m0 = up.mass()
        up.advect(w.steps, w.dt)
        out[dt_name] = (up.get_field("f64"), m0, up.mass())
        up.close()
    ref = out["f64"][0]
    e32 = _errors(out["f32"][0], ref)
    e16 = _errors(out["bf16"][0], ref)
    drift = {k: abs(m1 - m0) / abs(m0) for k, (f, m0, m1) in out.items()}
    print("f32", e32, "bf16", e16, "mass drift", drift)
    # FP32: the stated bar.  BF16 storage keeps 8 significant bits: a step
    # whose update is below half an ulp of the stored value is lost, so
    # smooth regions stagnate and the error after 200 steps is O(10%) of
    # the peak (measured 0.164 here); bounded, not a target.
    assert e32["max_rel"] <= 1e-5
    assert e16["max_rel"] <= 0.25
    assert drift["f64"] < 1e-13 and drift["f32"] < 1e-5 and drift["bf16"] < 0.1


def test_mass_is_compensated_sum():
    from fidibench_amd import Upwind
    shape = (33, 40, 64)
    rng = np.random.default_rng(3)
    f = rng.standard_normal(shape) * 1e8
    f[::3] += 1e16
    f[1::3] -= 1e16
    up = Upwind((1, 1, 1), (1, 1, 1), shape)
    up.set_field(f)
    exact = math.fsum(f.ravel())
    assert abs(up.mass() - exact) <= 2 * abs(exact) * 2 ** -52


def test_set_get_field_conversions():
    from fidibench_amd import Upwind
    from oracle import oracle as O
    shape = (4, 5, 64)
    x = np.random.default_rng(0).standard_normal(shape)
    up = Upwind((1, 1, 1), (1, 1, 1), shape, dtype="bf16")
    up.set_field(x)
    h = up.get_field()
    assert h.dtype == np.uint16
    assert np.array_equal(h.ravel(), O.f32_to_bf16(x.astype(np.float32)))
    assert np.array_equal(up.get_field("f32").ravel(), O.bf16_to_f32(h).ravel())
