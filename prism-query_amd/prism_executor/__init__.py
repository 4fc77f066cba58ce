"""MI355X (gfx950) drop-in for prism-executor's hash join, hash aggregation and
fused filter/project operators (native/prism-executor/src/{hash_join,
hash_aggregate,filter_project}.rs).  All compute runs in libprism_hip.so."""
from ._lib import PrismError, load  # noqa: F401

__all__ = ["PrismError", "load"]
