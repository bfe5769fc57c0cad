"""Failures a caller can force from the host end in status codes, not in a
fault: allocations beyond the device's memory (FDB_ERR_NOMEM) for the
upwind handle, the tape and the stencil filter."""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


def test_field_larger_than_hbm_is_nomem():
    from fidibench_amd import FdbError, Upwind, _lib as L
    with pytest.raises(FdbError) as e:
        Upwind((0.1, 0.1, 0.1), (1, 1, 1), (8192, 8192, 8192))  # 4.4 TB per FP64 field
    assert e.value.code == L.ERR_NOMEM


def test_filter_larger_than_hbm_is_nomem():
    from fidibench_amd import FdbError, Filter, _lib as L
    lap = {(0, 0, 0): -6.0, (1, 0, 0): 1.0, (-1, 0, 0): 1.0, (0, 1, 0): 1.0, (0, -1, 0): 1.0,
           (0, 0, 1): 1.0, (0, 0, -1): 1.0}
    with pytest.raises(FdbError) as e:
        Filter((8192, 8192, 8192), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), lap)
    assert e.value.code == L.ERR_NOMEM


def test_tape_budget_below_working_set_is_nomem():
    from fidibench_amd import FdbError, _lib as L
    from fidibench_amd.autograd import Tape
    with pytest.raises(FdbError) as e:
        Tape((64, 64, 64), (0.3, 0.2, 0.1), (1, 1, 1), 0.001, 10, L.F64, L.F64, 1 << 20)
    assert e.value.code == L.ERR_NOMEM
