"""MI355X-native Ant Colony Optimization solver (gfx950 HIP kernels behind a C ABI).

The product is libacogpu_amd.so (acogpu_amd/_build/), its C ABI
(include/acogpu.h), the C++ Colony/AntSystem/RankBasedAntSystem classes and
the acouav command line built on it.  This Python package is a thin host-side
binding used by tests and benchmarks.
"""
from . import capi, instances  # noqa: F401

__all__ = ["capi", "instances"]
