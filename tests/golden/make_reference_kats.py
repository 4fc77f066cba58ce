"""Writes reference_kats.json: the inputs and expected outputs that the
reference's own unit tests hold for this path, transcribed as data.

Sources (skn-dd/prism-query):
  hash_join      native/prism-executor/src/hash_join.rs:678-743
  hash_aggregate native/prism-executor/src/hash_aggregate.rs:534-660
  filter_project native/prism-executor/src/filter_project.rs:413-509
  string_ops     native/prism-executor/src/string_ops.rs:125-164
  fused join-agg native/prism-substrait/src/executor.rs:903-1080
  reducer merge  native/prism-bench/src/handler.rs:1377-1460
  integration    native/prism-executor/tests/integration.rs:332 (coast_count == 40_000)

The reference is Rust and cannot be built or run here; no vector below was
produced by executing it.  Expected values are the literals of the
reference's assertions (or, where an assertion is a row count, that count).
Run: python tests/golden/make_reference_kats.py
"""
import json
import os

KATS = {
    "hash_join": {
        "build": {"dept_id": [1, 2, 3], "dept_name": ["Engineering", "Sales", "Marketing"]},
        "probe": {"emp_id": [100, 101, 102, 103], "name": ["Alice", "Bob", "Carol", "Dave"],
                  "dept_id": [1, 2, 1, 99]},
        "probe_keys": [2], "build_keys": [0],
        "inner": {"rows": 3, "columns": 5},
        "left_semi": {"rows": 3, "columns": 3},
        "left_anti": {"rows": 1},
    },
    "hash_aggregate": {
        "sales": {"region": ["east", "west", "east", "west", "east"], "product": ["A", "B", "A", "A", "B"],
                  "amount": [100.0, 200.0, 150.0, 300.0, 50.0], "qty": [10, 20, 15, 30, 5]},
        "sum_by_region": {"group_by": [0], "aggregates": [[2, "SUM"]], "rows": 2, "columns": 2,
                          "total": 800.0},
        "count_and_avg": {"group_by": [0], "aggregates": [[2, "COUNT"], [2, "AVG"]], "rows": 2,
                          "columns": 3},
        "global_count_distinct_int64": {"values": [9007199254740992, 9007199254740993, 9007199254740992, None],
                                        "expected": 2},
        "grouped_count_distinct_utf8": {"region": ["east", "east", "east", "west", "west", "west"],
                                        "label": ["A", "A", "B", "A", None, "C"],
                                        "expected": {"east": 2, "west": 2}},
    },
    "filter_project": {
        "batch": {"id": [1, 2, 3, 4, 5], "name": ["alice", "bob", "carol", "dave", "eve"],
                  "value": [10.0, 20.0, 30.0, 40.0, 50.0]},
        "gt_id_3": {"ids": [4, 5]},
        "le_value_30_project_1_2": {"rows": 3, "columns": 2, "first": "alice", "third": "carol"},
        "ge2_and_le4": {"rows": 3},
        "literal_true": {"all_true": True},
        "if_then": {"id": [1, 2, 3], "label": ["a", "b", "c"], "expected": ["owner", "b", "c"]},
    },
    "string_ops": {
        "like_hel": {"values": ["hello", "world", "help", "held"], "pattern": "hel%",
                     "expected": [True, False, True, True]},
        "contains_world": {"values": ["Hello World", "Goodbye"], "pattern": "%World%",
                           "expected": [True, False]},
    },
    "fused_join_aggregate": {
        "lineitem": [{"l_orderkey": [0, 1], "l_extendedprice": [100.0, 200.0], "l_discount": [0.10, 0.20]},
                     {"l_orderkey": [2, 3], "l_extendedprice": [300.0, 400.0], "l_discount": [0.00, 0.25]}],
        "orders": {"o_orderkey": [0, 1, 2, 3], "o_orderstatus": ["F", "F", "O", "O"]},
        "expected_revenue_expr": {"F": "100.0 * 0.9 + 200.0 * 0.8", "O": "300.0 + 400.0 * 0.75"},
        "expected_count": {"F": 2, "O": 2},
        "avg": {"lineitem": {"l_orderkey": [0, 1, 2, 3], "l_extendedprice": [10.0, 20.0, 30.0, 40.0]},
                "expected": {"F": 15.0, "O": 35.0}},
    },
    "reducer_merge": {
        "partials": [{"status": ["F", "O"], "sum": [100.0, 50.0], "count": [2, 1]},
                     {"status": ["F", "P"], "sum": [30.0, 70.0], "count": [1, 2]}],
        "expected": {"F": [130.0, 3, "130.0 / 3.0"], "O": [50.0, 1, 50.0], "P": [70.0, 2, 35.0]},
    },
}

if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_kats.json")
    with open(out, "w") as f:
        json.dump(KATS, f, indent=1)
    print("wrote", out)
