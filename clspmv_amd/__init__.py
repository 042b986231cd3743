"""clspmv_amd -- MI355X-native SpMV engine with clSpMV's nine formats and
the Cocktail analyzer.  Compute lives in libclspmv.so (C++ host data model,
hand-written gfx950 HIP kernels, C ABI in include/clspmv.h); this package is
the Python host mirror used by tests and benchmarks."""
from . import _lib  # noqa: F401

__all__ = ["_lib", "matrix"]
