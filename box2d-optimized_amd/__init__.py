"""MI355X-native b2World::Step engine.

The product is lib/libbox2d_amd.so (HIP kernels for gfx950 + the b2g C ABI in
include/b2g.h + the drop-in C++ b2World API in host/). `engine` is a thin
ctypes binding to it; `scenes` generates the seeded workloads.
"""
from . import scenes  # noqa: F401
from . import engine  # noqa: F401
