"""MI355X-native monodomain tissue time-step (Chaste monodomain hot path).

The compute lives in ``chaste_amd/lib/libcmx.so`` (hand-written gfx950 HIP
kernels behind the C ABI declared in ``include/cmx.h``).  This package holds
the build recipe, a ctypes binding used by the tests and benchmark, and the
CellML code generator.
"""
from .build import build_library, library_path  # noqa: F401
