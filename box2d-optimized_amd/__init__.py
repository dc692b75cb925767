"""MI355X-native b2World::Step engine (HIP kernels behind the b2g C ABI)."""
from . import scenes  # noqa: F401
