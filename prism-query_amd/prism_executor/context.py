"""The process-wide device context used by the operator modules."""
from __future__ import annotations

import threading

_ctx = None
_lock = threading.Lock()


def context():
    """The lazily created prism_ctx on device 0 (HIP is required)."""
    global _ctx
    with _lock:
        if _ctx is None:
            from .device import Context
            _ctx = Context(0)
        return _ctx
