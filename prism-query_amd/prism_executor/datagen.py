"""Seeded synthetic columns (numpy twin of csrc/datagen.hip, bit-identical).

Counter-based and integer-only: value i of a stream depends only on
(seed, stream, i), so the device generator and this one agree at any size and
any sub-range can be generated on its own.  The workloads follow the shapes
named for this path (uniform keys/values, a seeded permutation of 0..n-1,
bounded Zipf keys, one-byte Utf8 flags as the reference's TPC-H generator
stores l_returnflag / l_linestatus, native/prism-bench/src/datagen.rs:109-110).
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
_C1 = np.uint64(0xBF58476D1CE4E5B9)
_C2 = np.uint64(0x94D049BB133111EB)


def _mix_int(z: int) -> int:
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _C1
    z = (z ^ (z >> np.uint64(27))) * _C2
    return z ^ (z >> np.uint64(31))


def stream_base(seed: int, stream: int) -> int:
    return _mix_int((seed * GOLDEN + stream + 1) & M64)


def rnd(seed: int, stream: int, start: int, n: int) -> np.ndarray:
    """rnd(seed, stream, i) for i in [start, start + n) as uint64."""
    base = np.uint64(stream_base(seed, stream))
    i = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(base + (i + np.uint64(1)) * np.uint64(GOLDEN))


def uniform_from(r: np.ndarray, rng: int) -> np.ndarray:
    assert 1 <= rng <= (1 << 32)
    with np.errstate(over="ignore"):
        return ((r >> np.uint64(32)) * np.uint64(rng)) >> np.uint64(32)


def uniform_i64(n: int, seed: int, stream: int, lo: int, rng: int, start: int = 0) -> np.ndarray:
    return uniform_from(rnd(seed, stream, start, n), rng).astype(np.int64) + np.int64(lo)


def uniform_i32(n: int, seed: int, stream: int, lo: int, rng: int, start: int = 0) -> np.ndarray:
    return (uniform_from(rnd(seed, stream, start, n), rng).astype(np.int64) + lo).astype(np.int32)


def bits_i64(n: int, seed: int, stream: int, start: int = 0) -> np.ndarray:
    return rnd(seed, stream, start, n).view(np.int64)


class Feistel:
    """Seeded bijection of [0, n): 4-round balanced Feistel network on the
    smallest even bit width covering n, with cycle walking."""

    def __init__(self, n: int, seed: int, stream: int):
        self.n = n
        bits = 2
        while (1 << bits) < n:
            bits += 1
        if bits & 1:
            bits += 1
        self.half = bits // 2
        self.mask = (1 << self.half) - 1
        base = stream_base(seed, stream)
        self.keys = [_mix_int((base + (r + 1) * GOLDEN) & M64) for r in range(4)]

    def _once(self, x: np.ndarray) -> np.ndarray:
        h, m = np.uint64(self.half), np.uint64(self.mask)
        L, R = x >> h, x & m
        for k in self.keys:
            F = _mix(R ^ np.uint64(k)) & m
            L, R = R, L ^ F
        return (L << h) | R

    def __call__(self, x: np.ndarray) -> np.ndarray:
        x = np.asarray(x, dtype=np.uint64)
        with np.errstate(over="ignore"):
            x = self._once(x)
            bad = x >= np.uint64(self.n)
            while bad.any():
                x[bad] = self._once(x[bad])
                bad = x >= np.uint64(self.n)
        return x


def permutation_i64(n: int, seed: int, stream: int, start: int = 0, count: int | None = None) -> np.ndarray:
    count = n - start if count is None else count
    f = Feistel(n, seed, stream)
    return f(np.arange(start, start + count, dtype=np.uint64)).astype(np.int64)


def zipf_cdf(n_values: int, s: float) -> np.ndarray:
    """u64 fixed-point CDF of a Zipf(s) law over ranks 1..n_values."""
    w = np.arange(1, n_values + 1, dtype=np.float64) ** (-s)
    c = np.cumsum(w)
    F = c / c[-1]
    q = np.floor(F * 18446744073709551616.0)
    q = np.minimum(q, 18446744073709549568.0)  # largest double below 2^64
    out = q.astype(np.uint64)
    out[-1] = np.uint64(M64)
    return out


def zipf_i64(n: int, seed: int, stream: int, cdf: np.ndarray, perm_stream: int, start: int = 0) -> np.ndarray:
    u = rnd(seed, stream, start, n)
    idx = np.searchsorted(cdf, u, side="right")
    idx = np.minimum(idx, len(cdf) - 1)
    return Feistel(len(cdf), seed, perm_stream)(idx.astype(np.uint64)).astype(np.int64)


def flag_utf8(n: int, seed: int, stream: int, alphabet: bytes, start: int = 0):
    """(bytes uint8[n], offsets int32[n+1]) of one-byte strings."""
    a = np.frombuffer(alphabet, dtype=np.uint8)
    idx = uniform_from(rnd(seed, stream, start, n), len(a)).astype(np.int64)
    return a[idx].copy(), np.arange(n + 1, dtype=np.int32)


# ---------------------------------------------------------------------------
# The workloads.  Seeds and streams are fixed here so every caller (tests,
# benchmark, smoke check) generates the same data.
# ---------------------------------------------------------------------------
SEED = 20240601
M = 1_000_000


def groupby_small(n: int = M, seed: int = SEED):
    """key int64 U[0,1000), value int64 U[0,10^6): SUM/COUNT/MIN/MAX by key."""
    return {"key": uniform_i64(n, seed, 1, 0, 1000), "value": uniform_i64(n, seed, 2, 0, 1_000_000)}


def join_inputs(n_build: int = M, n_probe: int = 64 * M, seed: int = SEED, start: int = 0,
                count: int | None = None):
    """Build: unique keys = seeded permutation of 0..n_build-1, int32 payload.
    Probe: keys U[0, 2*n_build) (about half match), int64 payload (full range)."""
    count = n_probe - start if count is None else count
    return {
        "build_key": permutation_i64(n_build, seed, 10),
        "build_payload": uniform_i32(n_build, seed, 11, -(1 << 31), 1 << 32),
        "probe_key": uniform_i64(count, seed, 12, 0, 2 * n_build, start=start),
        "probe_payload": bits_i64(count, seed, 13, start=start),
    }


FLAG_A = b"ANR"
FLAG_B = b"FO"


def filter_groupby(n: int, seed: int = SEED, start: int = 0):
    """date int32 U[0,2557), price int64 cents U[90000,10500000), discount int64
    in 0..10 inclusive, quantity int32 in 1..50 inclusive, flag A (3 values) and
    flag B (2 values) as one-byte Utf8."""
    a_bytes, a_off = flag_utf8(n, seed, 25, FLAG_A, start=start)
    b_bytes, b_off = flag_utf8(n, seed, 26, FLAG_B, start=start)
    return {
        "date": uniform_i32(n, seed, 20, 0, 2557, start=start),
        "price": uniform_i64(n, seed, 21, 90_000, 10_500_000 - 90_000, start=start),
        "discount": uniform_i64(n, seed, 22, 0, 11, start=start),
        "quantity": uniform_i32(n, seed, 23, 1, 50, start=start),
        "flag_a": (a_bytes, a_off),
        "flag_b": (b_bytes, b_off),
    }


def skewed_join(n_build: int, n_probe: int, n_values: int, n_attr: int, seed: int = SEED, s: float = 1.1,
                start: int = 0, count: int | None = None):
    """Build: keys Zipf(s) over n_values (seeded rank -> key permutation), int32
    attribute U[0, n_attr).  Probe: keys Zipf(s) over the same values, int64
    payload U[0, 10^6)."""
    count = n_probe - start if count is None else count
    cdf = zipf_cdf(n_values, s)
    return {
        "build_key": zipf_i64(n_build, seed, 30, cdf, 31),
        "build_attr": uniform_i32(n_build, seed, 32, 0, n_attr),
        "probe_key": zipf_i64(count, seed, 33, cdf, 31, start=start),
        "probe_payload": uniform_i64(count, seed, 34, 0, 1_000_000, start=start),
        "cdf": cdf,
    }


def lower_utf8(n: int, seed: int, stream: int, min_len: int, max_len: int):
    """(bytes uint8, offsets int32[n+1]) of random lowercase strings; bit-identical
    to prism_datagen_lower_utf8."""
    lens = (min_len + uniform_from(rnd(seed, stream, 0, n), max_len - min_len + 1)).astype(np.int64)
    off = np.zeros(n + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    total = int(off[-1])
    b = (ord("a") + uniform_from(rnd(seed, stream + 1, 0, total), 26)).astype(np.uint8)
    return b, off.astype(np.int32)
