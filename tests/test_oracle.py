"""Pins the CPU oracle (CPU only).

Each check is chosen so a plausible mistake (a dropped term, a wrong index or
sign, a transposed operand) fails it: golden fixtures made by the reference's
own generator, the reference's only known answer (group count ==
cardinality), closed forms, brute force on tiny inputs, published hash test
vectors and per-stage invariants of the restated reference algorithm.
"""
import glob
import math
import os
import random

import numpy as np
import pytest

import oracle as O

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "refgen_*.npz")))


# ---- golden fixtures from the reference generator -----------------------------
@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_golden_reference_workload(path):
    z = np.load(path, allow_pickle=False)
    keys, vals = z["keys"], z["vals"]
    r = O.groupby(keys, [vals], ["count", ("sum", 0)])
    # the reference's printed observable: result kv num == cardinality
    assert r["n_groups"] == int(z["cardinality"])
    np.testing.assert_array_equal(r["keys"], z["out_keys"])
    np.testing.assert_array_equal(r["aggs"][0], z["out_count"])
    np.testing.assert_array_equal((r["aggs"][1] & 0xFFFFFFFF).astype(np.uint32), z["out_sum_u32"])
    assert r["aggs"][0].sum() == keys.size


@pytest.mark.parametrize("path", GOLDEN, ids=os.path.basename)
def test_golden_restated_reference_path(path):
    """The restated two-phase algorithm (u32 wrap-around SUM) reproduces the
    fixture, with tiles and partitions small enough to exercise merging."""
    z = np.load(path, allow_pickle=False)
    keys, vals = z["keys"], z["vals"]
    n = keys.size
    tile = max(1000, n // 3)
    r = O.ref_run(keys, vals, tile_len=tile, P=17, capacity=tile, min_load=tile * 3 // 5, max_load=tile * 4 // 5)
    assert r["n_groups"] == int(z["cardinality"])
    o = np.argsort(r["keys"], kind="stable")
    np.testing.assert_array_equal(r["keys"][o], z["out_keys"])
    np.testing.assert_array_equal(r["count"][o].astype(np.int64), z["out_count"])
    np.testing.assert_array_equal((r["sum"][o] & 0xFFFFFFFF).astype(np.uint32), z["out_sum_u32"])
    # gapped layout: merged partitions tile the output in order
    assert np.all(r["par_count"] >= 0) and r["par_count"].sum() == r["n_groups"]


# ---- closed forms ------------------------------------------------------------------
@pytest.mark.parametrize("N,D", [(1, 1), (1000, 7), (100_000, 1024), (65_537, 65_536)])
def test_closed_form(N, D):
    i = np.arange(N, dtype=np.int64)
    keys = (i % D).astype(np.uint64)
    vals = i.astype(np.int64)
    r = O.groupby(keys, [vals], ["count", ("sum", 0), ("min", 0), ("max", 0)])
    k = np.arange(min(N, D), dtype=np.int64)
    c = (N - 1 - k) // D + 1
    np.testing.assert_array_equal(r["keys"], k.astype(np.uint64))
    np.testing.assert_array_equal(r["aggs"][0], c)
    np.testing.assert_array_equal(r["aggs"][1], k * c + D * c * (c - 1) // 2)
    np.testing.assert_array_equal(r["aggs"][2], k)
    np.testing.assert_array_equal(r["aggs"][3], k + D * (c - 1))


# ---- brute force on tiny inputs --------------------------------------------------------
def brute(keys, cols, aggs):
    d = {}
    for i, k in enumerate(keys.tolist()):
        d.setdefault(k, []).append(i)
    out_keys = sorted(d)
    outs = []
    for op, c in aggs:
        col = cols[c] if cols else None
        vals = []
        for k in out_keys:
            rows = d[k]
            if op == "count":
                vals.append(len(rows))
            else:
                xs = [col[j].item() for j in rows]
                if op == "sum":
                    vals.append(math.fsum(xs) if col.dtype == np.float64 else sum(xs))
                elif op == "min":
                    vals.append(min(xs))
                else:
                    vals.append(max(xs))
        outs.append(vals)
    return out_keys, outs


@pytest.mark.parametrize("seed", range(6))
def test_brute_force_tiny(seed):
    rng = random.Random(seed)
    n = rng.randint(1, 300)
    kd = np.uint32 if seed % 2 else np.uint64
    pool = [0, 1, 2**32 - 1 if kd is np.uint32 else 2**64 - 1, 7, 123456789]
    keys = np.array([rng.choice(pool + [rng.getrandbits(31)]) for _ in range(n)], dtype=kd)
    a = np.array([rng.choice([-(2**31), 2**31 - 1, 0, rng.randint(-1000, 1000)]) for _ in range(n)], dtype=np.int32)
    b = np.array([rng.getrandbits(63) - 2**62 for _ in range(n)], dtype=np.int64)
    f = np.array([rng.choice([0.5, -0.25, 1e3 * rng.random(), 3.0]) for _ in range(n)], dtype=np.float64)
    spec = [("count", 0), ("sum", 0), ("min", 0), ("max", 0), ("sum", 1), ("min", 1), ("max", 2), ("sum", 2)]
    r = O.groupby(keys, [a, b, f], [(op, c) for op, c in spec])
    bk, bo = brute(keys, [a, b, f], spec)
    assert r["keys"].tolist() == bk
    for (op, c), got, exp in zip(spec, r["aggs"], bo):
        if op == "sum" and c == 1:  # int64 sum wraps mod 2^64
            exp = [((x + 2**63) % 2**64) - 2**63 for x in exp]
        if op == "sum" and c == 2:
            np.testing.assert_allclose(got, exp, rtol=1e-15, atol=0)
        else:
            assert got.tolist() == exp


# ---- fp64: compensated sum ------------------------------------------------------------------
def test_neumaier_catastrophic_cancellation():
    x = np.array([1.0, 1e100, 1.0, -1e100] * 1000, dtype=np.float64)
    assert O.neumaier(x) == 2000.0
    assert math.fsum(x) == 2000.0
    rng = np.random.default_rng(5)
    y = rng.normal(0, 1e3, 200_000)
    assert abs(O.neumaier(y) - math.fsum(y)) <= 1e-12 * abs(math.fsum(y))


def test_fp64_integer_valued_sum_exact():
    x = np.arange(-5000, 5001, dtype=np.float64) * 3.0 + 0.5
    r = O.groupby(np.zeros(x.size, np.uint32), [x], [("sum", 0)])
    assert r["aggs"][0][0] == math.fsum(x.tolist())


# ---- prefix sums --------------------------------------------------------------------------------
def test_scan_textbook():
    x = np.array([3, 0, 5, 1, 2**40], dtype=np.uint64)
    np.testing.assert_array_equal(O.scan(x, exclusive=True), [0, 3, 3, 8, 9])
    np.testing.assert_array_equal(O.scan(x, exclusive=False), [3, 3, 8, 9, 9 + 2**40])


# ---- hash and generators: published test vectors ---------------------------------------------------
def test_murmur3_x86_32_vectors():
    # MurmurHash3_x86_32 published vectors (4-byte inputs)
    assert O.lib().tga_o_murmur3_u32(0, 0) == 0x2362F9DE
    assert O.lib().tga_o_murmur3_u32(int.from_bytes(b"aaaa", "little"), 0x9747B28C) == 0x5A97808A
    assert O.lib().tga_o_murmur3_u32(int.from_bytes(b"abcd", "little"), 0x9747B28C) == 0xF0478627


def test_splitmix64_reference_stream():
    # SplitMix64 seeded with 0: first outputs of the published generator
    assert O.lib().tga_o_splitmix(0, 0) == 0xE220A8397B1DCDAF
    assert O.lib().tga_o_splitmix(0, 1) == 0x6E789E6AA1B965F4


def test_generators_distinct_and_in_range():
    k = O.gen_keys_u32(200_000, 1000, 1, 0)
    assert np.unique(k).size == 1000
    k = O.gen_keys_u64(1_000_000, 1 << 16, 2, 9)
    assert np.unique(k).size == 1 << 16
    v = O.gen_vals_i32(100_000, -(1 << 30), 1 << 30, 3)
    assert v.min() >= -(1 << 30) and v.max() < (1 << 30)


# ---- per-stage invariants of the restated reference ------------------------------------------------
def test_phase1_invariants():
    n, tile, P = 50_000, 12_000, 13
    keys0 = O.gen_keys_u32(n, 3000, 4, 1)
    vals0 = O.gen_vals_i32(n, 0, 1 << 20, 5).view(np.uint32).copy()
    keys, vals = keys0.copy(), vals0.copy()
    segs = O.ref_phase1_u32(keys, vals, tile, P)
    total = 0
    for t in range((n + tile - 1) // tile):
        b, e = t * tile, min(n, (t + 1) * tile)
        tk = keys0[b:e]
        u = np.unique(tk)
        r = u.size
        # tile result is the distinct keys of the tile, written back at the tile offset
        out_k = keys[b:b + r]
        assert np.unique(out_k).size == r
        np.testing.assert_array_equal(np.sort(out_k), u)
        # multiset sum preserved (mod 2^32)
        assert (int(vals[b:b + r].astype(np.uint64).sum()) & 0xFFFFFFFF) == (int(vals0[b:e].astype(np.uint64).sum()) & 0xFFFFFFFF)
        total += r
    # segments: P per tile, each key lies in its partition, lengths add up
    assert len(segs) == P * ((n + tile - 1) // tile)
    assert sum(int(s[2]) for s in segs) == total
    for p, b, ln in segs.tolist():
        for key in keys[b:b + ln].tolist():
            assert O.partition_of_u32(key, P) == p


def test_merge_partition_properties():
    rng = np.random.default_rng(1)
    sizes = rng.integers(0, 1000, 300).astype(np.uint32)
    K = 6000
    m = O.merge_sizes(sizes, K)
    assert int(m.sum()) == int(sizes.sum())
    assert np.all(np.diff(m.astype(np.int64)) >= 0)   # smallest first, like the reference
    # no two of the merged partitions fit together under K (greedy stops)
    assert len(m) == 1 or int(m[0]) + int(m[1]) > K
    assert O.merge_sizes(np.array([5], np.uint32), 1).tolist() == [5]


def test_reference_overflow_rule():
    # more distinct keys in one merged partition than min_load with rows left
    n = 40_000
    keys = np.arange(n, dtype=np.uint32)
    vals = np.ones(n, np.uint32)
    with pytest.raises(OverflowError):
        O.ref_run(keys, vals, tile_len=40_000, P=1, capacity=10_000, min_load=6_000, max_load=8_000)
    r = O.ref_run(keys, vals, tile_len=40_000, P=1, capacity=100_000, min_load=60_000, max_load=80_000)
    assert r["n_groups"] == n


def test_restatement_widened_types_match_exact_oracle():
    n = 300_000
    k = O.gen_keys_u64(n, 40_000, 8, 3)
    v = O.gen_vals_i32(n, -(1 << 30), 1 << 30, 9)
    r = O.ref_run(k, v, tile_len=100_000, P=32, capacity=100_000, min_load=60_000, max_load=80_000)
    e = O.groupby(k, [v], ["count", ("sum", 0), ("min", 0), ("max", 0)])
    o = np.argsort(r["keys"], kind="stable")
    np.testing.assert_array_equal(r["keys"][o], e["keys"])
    np.testing.assert_array_equal(r["count"][o].astype(np.int64), e["aggs"][0])
    np.testing.assert_array_equal(r["sum"][o], e["aggs"][1])
    np.testing.assert_array_equal(r["min"][o], e["aggs"][2].astype(np.int64))
    np.testing.assert_array_equal(r["max"][o], e["aggs"][3].astype(np.int64))
