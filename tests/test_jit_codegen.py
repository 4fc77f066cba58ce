"""Generated group-by kernels: every row-program opcode, key type, aggregate
lane and accumulation form yields source that hipRTC compiles for gfx950 (no
device needed).  What the kernels compute is checked on the GPU against the
oracle and against the prebuilt kernels (test_gpu_groupby.py)."""
import pytest

from prism_executor import _lib as L
from prism_executor import device as D

I64, I32, F64, UTF8, BOOL, DATE = (L.TYPE_INT64, L.TYPE_INT32, L.TYPE_FLOAT64, L.TYPE_UTF8, L.TYPE_BOOL,
                                   L.TYPE_DATE32)
LE, EQ, GT = L.CMP["LE"], L.CMP["EQ"], L.CMP["GT"]

PLANS = {
    # config-0 shape: int64 key, SUM/COUNT/MIN/MAX of an int64 value
    "int_key_all_aggs": dict(column_types=[I64, I64], keys=[0], values=[[("LOAD_I", 1)]],
                             aggs=[("SUM", 0), ("COUNT", 0), ("MIN", 0), ("MAX", 0)]),
    # config-2 shape: date filter, two Utf8 keys, exact int64 projection
    "filter_utf8_keys": dict(column_types=[I32, I64, I64, UTF8, UTF8], keys=[3, 4],
                             values=[[("LOAD_I", 1), ("CONST_I", 0, 100), ("LOAD_I", 2), ("SUB_I",), ("MUL_I",)]],
                             aggs=[("SUM", 0), ("COUNT", 0), ("AVG", 0)],
                             filter_prog=[("LOAD_I", 0), ("CONST_I", 0, 2470), ("CMP_I", LE)]),
    # float64 projection with every float opcode; float total-order compare
    "float_ops": dict(column_types=[F64, F64, DATE], keys=[2],
                      values=[[("LOAD_F", 0), ("LOAD_F", 1), ("MUL_F",), ("CONST_F", 0, 0, 0.5), ("ADD_F",),
                               ("LOAD_F", 1), ("DIV_F",), ("NEG_F",), ("CONST_F", 0, 0, 1.0), ("SUB_F",)],
                              [("LOAD_I", 2), ("I2F",)]],
                      aggs=[("SUM", 0), ("MIN", 1), ("MAX", 0), ("AVG", 1)],
                      filter_prog=[("LOAD_F", 0), ("CONST_F", 0, 0, 3.25), ("CMP_F", GT)]),
    # bool / float keys, integer opcodes, null tests, AND / OR / NOT
    "bool_float_keys": dict(column_types=[BOOL, F64, I64, I32], keys=[0, 1],
                            values=[[("LOAD_I", 2), ("NEG_I",), ("LOAD_I", 3), ("ADD_I",)]],
                            aggs=[("SUM", 0), ("COUNT", 0)],
                            filter_prog=[("IS_NOT_NULL", 2), ("LOAD_I", 0), ("NOT",), ("AND",), ("IS_NULL", 3),
                                         ("OR",)]),
    # strings: comparison, LIKE shapes (fast and general), ILIKE
    "string_predicates": dict(column_types=[UTF8, UTF8, I64], keys=[0],
                              values=[[("LOAD_I", 2)]], aggs=[("SUM", 0), ("MAX", 0)],
                              filter_prog=[("LIKE", 0, 0), ("ILIKE", 1, 1), ("OR",), ("CMP_STR", 0, 2 | (GT << 8)),
                                           ("AND",)],
                              strings=[b"%abc%", b"x_y%", b"m"]),
    # SELECT / ZIP with a null branch and mixed types
    "select_zip": dict(column_types=[I64, F64, I32], keys=[2],
                       values=[[("LOAD_I", 0), ("CONST_I", 0, 0), ("CMP_I", GT), ("LOAD_F", 1), ("CONST_NULL",),
                                ("SELECT",)],
                               [("CONST_I", 0, 7), ("LOAD_I", 0), ("CONST_I", 0, 5), ("CMP_I", EQ), ("LOAD_I", 2),
                                ("ZIP",)]],
                       aggs=[("SUM", 0), ("SUM", 1), ("MIN", 1)]),
    # global aggregates (no GROUP BY): COUNT of non-null values
    "global": dict(column_types=[I64, F64], keys=[],
                   values=[[("LOAD_I", 0)], [("LOAD_F", 1)]],
                   aggs=[("COUNT", 0), ("SUM", 0), ("AVG", 1), ("MIN", 0), ("MAX", 1), ("COUNT", 1)]),
}


@pytest.mark.parametrize("name", sorted(PLANS))
def test_generated_group_by_kernel_compiles(name):
    plan = PLANS[name]
    n_cols = len(plan["column_types"])
    # alternate the form and whether columns carry validity bitmaps
    form = sorted(PLANS).index(name) % 2
    nullable = (1 << n_cols) - 1 if form else 0
    src = D.jit_source(**plan, form=form, nullable=nullable, compile=True)
    assert "prism_gb_update" in src


def test_source_specialises_null_free_columns():
    plan = PLANS["filter_utf8_keys"]
    plain = D.jit_source(**plan, form=0, nullable=0)
    nulls = D.jit_source(**plan, form=0, nullable=0b11111)
    assert "bit_at(n" not in plain and "bit_at(n" in nulls
    # the filter constant is folded into the code
    assert "0x00000000000009a6ULL" in plain


def test_bad_form_is_rejected():
    with pytest.raises(L.PrismError) as e:
        D.jit_source(**PLANS["global"], form=7)
    assert e.value.kind == "InvalidArgument"
