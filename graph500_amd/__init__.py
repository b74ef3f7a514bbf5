"""MI355X-native Graph500 BFS hot path (Kronecker generation, CSR build,
direction-optimizing BFS) behind a C ABI; see include/g5.h."""
from .g5 import G5Context, G5Error, G5Server, load_library, EDGE_DTYPE, QUERY_DTYPE  # noqa: F401
from . import autotune  # noqa: F401
