"""Pins the closed-form answers used by the full-size GPU tests (CPU only):
they must equal a numpy brute force (sort + segmented reduce) and the oracle
on small workloads of the same generators, including n < D, n not a
multiple of D, 32- and 64-bit keys and a negative value offset."""
import numpy as np
import pytest

import oracle as O
from closed_form import columns, expected, fmix32, mix64

AGGS4 = ["count", ("sum", 0), ("min", 0), ("max", 0)]


def brute(keys, vals):
    o = np.argsort(keys, kind="stable")
    k, v = keys[o], vals[o].astype(np.int64)
    starts = np.flatnonzero(np.r_[True, k[1:] != k[:-1]])
    return k[starts], [np.diff(np.r_[starts, k.size]).astype(np.int64), np.add.reduceat(v, starts),
                       np.minimum.reduceat(v, starts).astype(np.int32), np.maximum.reduceat(v, starts).astype(np.int32)]


def test_mixers_match_published_values():
    # SplitMix64 (seed 0 stream): the finalizer of 0x9e3779b97f4a7c15
    assert int(mix64(np.array([0x9E3779B97F4A7C15], np.uint64))[0]) == 0xE220A8397B1DCDAF
    # fmix32 is a bijection: no two of 2^16 consecutive inputs collide
    assert np.unique(fmix32(np.arange(1 << 16, dtype=np.uint32))).size == 1 << 16
    assert int(fmix32(np.array([0], np.uint32))[0]) == 0


@pytest.mark.parametrize("n,D,bits,salt,lo", [
    (1, 1, 64, 0, 0), (1000, 7, 32, 5, 0), (100_000, 1024, 64, 0x5EED, 0),
    (65_537, 65_536, 32, 3, -(1 << 20)), (5000, 1 << 20, 64, 9, 17),
])
def test_closed_form_equals_brute_force_and_oracle(n, D, bits, salt, lo):
    kh, vh = columns(n, D, bits, salt, lo)
    ek, ea = expected(n, D, bits, salt, lo)
    bk, ba = brute(kh, vh)
    np.testing.assert_array_equal(ek, bk)
    for e, b in zip(ea, ba):
        np.testing.assert_array_equal(e, b)
    r = O.groupby(kh, [vh], AGGS4)
    np.testing.assert_array_equal(r["keys"], ek)
    for e, g in zip(ea, r["aggs"]):
        np.testing.assert_array_equal(e, g)
