"""Group-by plans over every row-program opcode, key type and null pattern:
the kernels generated per plan (hipRTC) and the prebuilt interpreting kernels
give the same groups (first rows, key hashes, row counts), the same exact
integer sums and counts, the same MIN/MAX, and float sums within 1e-12
relative.  Both engines are checked against the oracle on fixed plans in
test_gpu_groupby.py; this file covers the plan space between them."""
import math

import numpy as np
import pytest

from gpu_util import say, to_dev

from prism_executor import _lib as L
from prism_executor import device as D

pytestmark = pytest.mark.gpu

I64, I32, F64, UTF8, BOOL, DATE = (L.TYPE_INT64, L.TYPE_INT32, L.TYPE_FLOAT64, L.TYPE_UTF8, L.TYPE_BOOL,
                                   L.TYPE_DATE32)
EQ, NE, LT, LE, GT, GE = (L.CMP[k] for k in ("EQ", "NE", "LT", "LE", "GT", "GE"))
N = 300_000
WORDS = [b"", b"a", b"abc", b"xabcx", b"zz", b"Abc", b"abcdefghij", b"\xc3\xa9t\xc3\xa9"]
TYPES = [I64, I32, F64, UTF8, BOOL, DATE]


def _bits(mask):
    return np.packbits(mask, bitorder="little")


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(77)
    c0 = rng.integers(-50, 50, N).astype(np.int64)
    c1 = rng.integers(-5, 6, N).astype(np.int32)
    c2 = np.round(rng.standard_normal(N) * 100, 2)
    c2[::997] = np.nan
    c2[5::1013] = 3.0  # integral values take the exact lane
    widx = rng.integers(0, len(WORDS), N)
    strs = [WORDS[i] for i in widx]
    offs = np.zeros(N + 1, np.int32)
    offs[1:] = np.cumsum([len(s) for s in strs])
    sbytes = np.frombuffer(b"".join(strs), np.uint8) if offs[-1] else np.zeros(1, np.uint8)
    c4 = rng.integers(0, 2, N).astype(bool)
    c5 = rng.integers(0, 12, N).astype(np.int32)
    valid = [rng.random(N) > p for p in (0.05, 0.1, 0.02, 0.08, 0.1, 0.03)]
    dev = {
        "c0": to_dev(c0), "c1": to_dev(c1), "c2": to_dev(c2), "s": to_dev(sbytes), "so": to_dev(offs),
        "c4": to_dev(_bits(c4)), "c5": to_dev(c5), "v": [to_dev(_bits(m)) for m in valid],
    }
    return dev


def columns(dev, nullable):
    """Column descriptors; column c carries its validity bitmap when bit c of `nullable` is set."""
    v = [dev["v"][c] if (nullable >> c) & 1 else None for c in range(6)]
    return [D.column(I64, dev["c0"], validity=v[0]), D.column(I32, dev["c1"], validity=v[1]),
            D.column(F64, dev["c2"], validity=v[2]), D.column(UTF8, dev["s"], dev["so"], validity=v[3]),
            D.column(BOOL, dev["c4"], validity=v[4]), D.column(DATE, dev["c5"], validity=v[5])]


# (name, keys, filter program, value programs, aggregates, strings, nullable mask)
PLANS = [
    ("int keys, int arithmetic", [1], [("LOAD_I", 0), ("CONST_I", 0, -20), ("CMP_I", GT)],
     [[("LOAD_I", 0), ("CONST_I", 0, 3), ("MUL_I",), ("LOAD_I", 1), ("SUB_I",)], [("LOAD_I", 0), ("NEG_I",)]],
     [("SUM", 0), ("COUNT", 0), ("MIN", 1), ("MAX", 0), ("AVG", 1)], [], 0),
    ("utf8 + bool keys with nulls, LIKE filter", [3, 4], [("LIKE", 3, 0), ("LOAD_I", 1), ("CONST_I", 0, 0),
                                                          ("CMP_I", GE), ("OR",)],
     [[("LOAD_I", 0)]], [("SUM", 0), ("COUNT", 0), ("MAX", 0)], [b"%abc%"], 0b111111),
    ("date key, float values, float compare", [5], [("LOAD_F", 2), ("CONST_F", 0, 0, -50.0), ("CMP_F", GT)],
     [[("LOAD_F", 2), ("LOAD_I", 1), ("I2F",), ("MUL_F",), ("CONST_F", 0, 0, 0.5), ("ADD_F",)], [("LOAD_F", 2)]],
     [("SUM", 0), ("MIN", 1), ("MAX", 1), ("AVG", 0)], [], 0b100100),
    ("float + int32 keys, ILIKE and NOT", [2, 1], [("ILIKE", 3, 0), ("NOT",)],
     [[("LOAD_I", 0), ("CONST_I", 0, 7), ("ADD_I",)]], [("SUM", 0), ("COUNT", 0)], [b"ABC%"], 0b001011),
    ("SELECT / ZIP / null constant", [1], None,
     [[("LOAD_I", 0), ("CONST_I", 0, 0), ("CMP_I", LT), ("LOAD_I", 0), ("NEG_I",), ("LOAD_I", 0), ("SELECT",)],
      [("CONST_NULL",), ("LOAD_I", 4), ("LOAD_I", 1), ("ZIP",)],
      [("LOAD_I", 0), ("CONST_I", 0, 10), ("CMP_I", EQ), ("LOAD_F", 2), ("CONST_NULL",), ("SELECT",)]],
     [("SUM", 0), ("SUM", 1), ("MIN", 2), ("COUNT", 0)], [], 0b000001),
    ("IS NULL / IS NOT NULL, string compare", [4], [("IS_NOT_NULL", 0), ("CMP_STR", 3, 0 | (GE << 8)), ("AND",),
                                                    ("IS_NULL", 2), ("OR",)],
     [[("LOAD_I", 5), ("LOAD_I", 1), ("ADD_I",)]], [("SUM", 0), ("MAX", 0), ("MIN", 0)], [b"abc"], 0b000111),
    ("global aggregates, nulls skipped", [], [("LOAD_I", 1), ("CONST_I", 0, 3), ("CMP_I", NE)],
     [[("LOAD_I", 0)], [("LOAD_F", 2)]], [("COUNT", 0), ("SUM", 0), ("AVG", 1), ("MIN", 1), ("MAX", 0)], [],
     0b000101),
    ("many groups: int64 x int32 keys", [0, 1], None, [[("LOAD_I", 5)]], [("SUM", 0), ("COUNT", 0)], [], 0),
    ("utf8 key, general LIKE pattern", [3], [("LIKE", 3, 0), ("NOT",)], [[("LOAD_I", 1)]],
     [("SUM", 0), ("MAX", 0)], [b"a_c%"], 0b001000),
]


@pytest.mark.parametrize("plan", PLANS, ids=[p[0] for p in PLANS])
def test_generated_equals_prebuilt(ctx, data, monkeypatch, plan):
    name, keys, filt, values, aggs, strings, nullable = plan
    say("engines", name)
    cols = columns(data, nullable)
    res = {}
    for engine in ("generated", "prebuilt"):
        monkeypatch.setenv("PRISM_JIT", "1" if engine == "generated" else "0")
        g = D.GroupBy(ctx, TYPES, keys, values, aggs, filter_prog=filt, strings=strings)
        g.update(cols, N, 0)
        res[engine] = g.read(g.finish())
        g.close()
    a, b = res["generated"], res["prebuilt"]
    assert a["first_row"].tolist() == b["first_row"].tolist()
    assert a["key_hash"].tolist() == b["key_hash"].tolist()
    assert a["row_count"].tolist() == b["row_count"].tolist()
    assert len(a["first_row"]) > 0
    for i, (fn, _) in enumerate(aggs):
        assert a["sum128"][i] == b["sum128"][i], (name, fn)
        assert np.array_equal(a["i64"][i], b["i64"][i]), (name, fn)
        x, y = a["f64"][i], b["f64"][i]
        if fn in ("MIN", "MAX", "COUNT"):
            assert np.array_equal(x, y, equal_nan=True), (name, fn)
        else:
            for u, w in zip(x.tolist(), y.tolist()):
                assert (math.isnan(u) and math.isnan(w)) or u == w or abs(u - w) <= 1e-12 * max(abs(u), abs(w)), \
                    (name, fn, u, w)
    st = D.jit_stats()
    assert st["failures"] == 0, st["last_error"]
