"""fidibench_amd -- MI355X-native upwind advection (pletzer/fidibench upwind/cxx path).

The compute path is libfdb_upwind.so (HIP kernels for gfx950 behind a C ABI,
include/fidibench_amd/fdb_upwind.h and fdb_filter.h); this package is its
Python host side:

* ``Upwind``            mirror of the reference's Upwind<NDIMS> class
* ``Filter``            mirror of the reference's Filter class (one process)
* ``configs``           seeded synthetic workloads
* ``autograd``          torch.autograd.Function over the adjoint (imports torch)
"""
from ._lib import FdbError, lib  # noqa: F401
from .filter import Filter  # noqa: F401
from .upwind import Upwind  # noqa: F401

__all__ = ["Upwind", "Filter", "FdbError", "lib"]
