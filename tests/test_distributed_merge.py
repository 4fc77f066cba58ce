"""Distributed aggregate merge (prism-bench handler.rs:270-344, 397-526): the
oracle's restatement pinned by the reference's own reducer test
(handler.rs:1377-1460) and by the expansion's closed form; the C ABI's
expansion checked without a device."""
import json
import os

import pyarrow as pa

import oracle as orc
from prism_executor.distributed import build_final_merge_config, expand_worker_aggregates
from prism_executor.hash_aggregate import AggExpr, AggFunc

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))

ORIGINAL = [(1, "SUM"), (1, "COUNT"), (1, "AVG")]


def kat_partials():
    """The reference test's partial batches: status, agg_0 = SUM, agg_1 = COUNT,
    agg_2 / agg_3 = the SUM / COUNT an AVG expands into."""
    out = []
    for p in KATS["reducer_merge"]["partials"]:
        out.append(pa.record_batch([pa.array(p["status"]), pa.array(p["sum"], pa.float64()),
                                    pa.array(p["count"], pa.int64()), pa.array(p["sum"], pa.float64()),
                                    pa.array(p["count"], pa.int64())],
                                   names=["status", "agg_0", "agg_1", "agg_2", "agg_3"]))
    return out


def test_oracle_reducer_matches_reference_test():
    parts = kat_partials()
    first, vals = orc.reducer_merge(parts, 1, ORIGINAL)
    keys = [parts[b].column(0)[r].as_py() for b, r in first]
    assert keys == ["F", "O", "P"]
    exp = KATS["reducer_merge"]["expected"]
    for i, k in enumerate(keys):
        e = exp[k]
        e_avg = eval(e[2]) if isinstance(e[2], str) else e[2]
        assert (vals[0][i], vals[1][i], vals[2][i]) == (e[0], e[1], e_avg)
        assert isinstance(vals[1][i], int)


def test_oracle_expansion_closed_form():
    w, m = orc.expand_worker_aggregates([(3, "AVG"), (1, "MIN"), (2, "AVG"), (0, "COUNT_DISTINCT")])
    assert w == [(3, "SUM"), (3, "COUNT"), (1, "MIN"), (2, "SUM"), (2, "COUNT"), (0, "COUNT_DISTINCT")]
    assert m == [[0, 1], [2], [3, 4], [5]]


def test_abi_expansion_and_merge_config():
    orig = [AggExpr(1, AggFunc.Sum, "s"), AggExpr(1, AggFunc.Count, "c"), AggExpr(2, AggFunc.Avg, "a"),
            AggExpr(3, AggFunc.Min, "lo"), AggExpr(3, AggFunc.Max, "hi"), AggExpr(4, AggFunc.CountDistinct, "d")]
    w, m = expand_worker_aggregates(orig)
    assert [(a.column, a.func) for a in w] == [(1, AggFunc.Sum), (1, AggFunc.Count), (2, AggFunc.Sum),
                                              (2, AggFunc.Count), (3, AggFunc.Min), (3, AggFunc.Max),
                                              (4, AggFunc.CountDistinct)]
    assert [a.output_name for a in w] == [f"agg_{i}" for i in range(7)]
    assert m == [[0], [1], [2, 3], [4], [5], [6]]
    cfg = build_final_merge_config(2, w)
    assert list(cfg.group_by) == [0, 1]
    assert [(a.column, a.func) for a in cfg.aggregates] == [(2, AggFunc.Sum), (3, AggFunc.Sum), (4, AggFunc.Sum),
                                                          (5, AggFunc.Sum), (6, AggFunc.Min), (7, AggFunc.Max),
                                                          (8, AggFunc.Sum)]
