"""Helpers shared by the GPU tests."""
import numpy as np

import oracle as orc


def to_dev(a):
    from prism_executor.context import context
    from prism_executor.device import DeviceArray
    return DeviceArray.from_numpy(context(), np.ascontiguousarray(a))


def exact_sum(a) -> int:
    """Exact integer sum of an int64 numpy array (< 2^31 elements)."""
    a = np.asarray(a, dtype=np.int64)
    return int((a >> 32).sum()) * (1 << 32) + int((a & 0xFFFFFFFF).sum())


def i64_batch(*cols):
    return orc.Batch([orc.Col(orc.INT64, np.ascontiguousarray(c, np.int64)) for c in cols])


def say(*a):
    print("[case]", *a, flush=True)
