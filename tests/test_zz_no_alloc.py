"""fdb_adjoint_step_dev_noalloc (partial sums in a fixed device array, no
allocation) against fdb_adjoint_step_dev (partial sums in a hipMalloc'd
buffer): the same lam bits and the same velocity-gradient sums, on two
calls in a row with the second on a smaller grid (stale partials of the
first call must not leak into the second), for the storage pairings of
the gradient; a grid needing more partial slots than the array holds is an
argument error before anything is launched."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TDT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}


def _step(name, shape, v, dt, lam_kind, lam_next, f_kind, ft):
    from fidibench_amd import _lib as L
    code = {"f64": L.F64, "f32": L.F32, "bf16": L.BF16}
    lam = torch.empty_like(lam_next)
    acc = torch.zeros(6, dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    L.check(getattr(L.lib(), name)(3, L.u64s(shape), L.dbls(v), L.dbls((1, 1, 1)), dt,
                                   code[lam_kind], lam_next.data_ptr(), lam.data_ptr(),
                                   code[f_kind], None if ft is None else ft.data_ptr(),
                                   None if ft is None else acc.data_ptr(), s))
    torch.cuda.synchronize()
    return lam, acc.cpu().numpy()


def _bits(t):
    return t.view({8: torch.int64, 4: torch.int32, 2: torch.int16}[t.element_size()])


@pytest.mark.parametrize("lam_kind,f_kind", [("f64", "f64"), ("f32", "f32"), ("f32", "bf16"),
                                             ("bf16", "bf16"), ("f64", None)])
def test_noalloc_matches_allocating_entry_point(lam_kind, f_kind):
    g = torch.Generator(device="cuda").manual_seed(5)
    for shape, v, dt in (((40, 96, 200), (0.45, -0.3, 0.2), 0.003),
                         ((9, 13, 70), (-0.35, 0.4, -0.25), 0.011)):
        lam_next = torch.randn(shape, generator=g, device="cuda",
                               dtype=torch.float64).to(TDT[lam_kind])
        ft = None if f_kind is None else torch.rand(shape, generator=g, device="cuda",
                                                    dtype=torch.float64).to(TDT[f_kind])
        ref_lam, ref_acc = _step("fdb_adjoint_step_dev", shape, v, dt, lam_kind, lam_next,
                                 f_kind or lam_kind, ft)
        lam, acc = _step("fdb_adjoint_step_dev_noalloc", shape, v, dt, lam_kind, lam_next,
                         f_kind or lam_kind, ft)
        assert torch.equal(_bits(lam), _bits(ref_lam)), (shape, lam_kind, f_kind)
        assert np.array_equal(acc.view(np.int64), ref_acc.view(np.int64)), (acc, ref_acc)
        if ft is not None:
            assert np.all(acc[0::2] != 0.0)  # the products were summed


def test_noalloc_rejects_grid_beyond_its_scratch():
    """16384 x 16384 planes need 131072 partial slots (FP64 tiles of 128 x
    16 cells) against the array's 65536: FDB_ERR_INVALID before any launch
    (the buffers are real, of the full size, so nothing could go out of
    bounds either way)."""
    from fidibench_amd import _lib as L
    shape = (1, 16384, 16384)
    a = torch.zeros(shape, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    f = torch.zeros_like(a)
    acc = torch.zeros(6, dtype=torch.float64, device="cuda")
    rc = L.lib().fdb_adjoint_step_dev_noalloc(3, L.u64s(shape), L.dbls((0.3, 0.2, -0.1)),
                                              L.dbls((1, 1, 1)), 1e-5, L.F64, a.data_ptr(),
                                              b.data_ptr(), L.F64, f.data_ptr(), acc.data_ptr(),
                                              torch.cuda.current_stream().cuda_stream)
    assert rc == L.ERR_INVALID
    assert b"131072" in L.lib().fdb_last_error()
