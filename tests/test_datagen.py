"""The seeded generators: invariants the workloads rely on."""
import numpy as np

from prism_executor import datagen as dg


def test_permutation_is_a_bijection():
    for n in (1, 2, 3, 1000, 1 << 16, 1_000_000):
        p = dg.permutation_i64(n, 7, 10)
        assert np.array_equal(np.sort(p), np.arange(n))


def test_permutation_subrange_matches_full():
    full = dg.permutation_i64(100_000, 7, 10)
    part = dg.permutation_i64(100_000, 7, 10, start=31_337, count=1000)
    assert np.array_equal(full[31_337:32_337], part)


def test_uniform_ranges_and_subranges():
    x = dg.uniform_i64(200_000, 1, 2, -5, 11)
    assert x.min() == -5 and x.max() == 5
    assert np.array_equal(dg.uniform_i64(1000, 1, 2, -5, 11, start=500), x[500:1500])
    q = dg.uniform_i32(100_000, 1, 3, 1, 50)
    assert q.min() == 1 and q.max() == 50


def test_zipf_head_probability():
    V, s = 1 << 16, 1.1
    cdf = dg.zipf_cdf(V, s)
    assert np.all(np.diff(cdf.astype(np.float64)) >= 0) and cdf[-1] == np.uint64(2**64 - 1)
    x = dg.zipf_i64(400_000, 9, 1, cdf, 2)
    top = np.bincount(x).max() / len(x)
    H = np.sum(np.arange(1, V + 1, dtype=np.float64) ** -s)
    assert abs(top - 1.0 / H) < 0.01


def test_flags():
    b, off = dg.flag_utf8(10_000, 3, 25, dg.FLAG_A)
    assert set(np.unique(b).tolist()) == set(dg.FLAG_A)
    assert np.array_equal(off, np.arange(10_001))
