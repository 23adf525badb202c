"""GPU parity of the plain operator against the CPU oracle (bit-exact for
integers; fp64 SUM within 1e-12 relative of the compensated oracle and equal
to math.fsum).  Every call goes through the C ABI (libtga.so)."""
import math

import numpy as np
import pytest

import cpu_gpu_coprocess_groupby_agg_amd as tga
import oracle as O

pytestmark = pytest.mark.gpu

AGGS4 = ["count", ("sum", 0), ("min", 0), ("max", 0)]


def check_equal(got, ref):
    assert got["n_groups"] == ref["n_groups"]
    np.testing.assert_array_equal(got["keys"], ref["keys"])
    for g, r in zip(got["aggs"], ref["aggs"]):
        np.testing.assert_array_equal(g, r)


def test_generators_match_oracle(ctx):
    n = 100_003
    k32 = ctx.generate(tga.GEN_KEYS_U32, n, a=1 << 10, seed=11, salt=5).to_numpy()
    k64 = ctx.generate(tga.GEN_KEYS_U64, n, a=1 << 20, seed=12, salt=6).to_numpy()
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=13).to_numpy()
    np.testing.assert_array_equal(k32, O.gen_keys_u32(n, 1 << 10, 11, 5))
    np.testing.assert_array_equal(k64, O.gen_keys_u64(n, 1 << 20, 12, 6))
    np.testing.assert_array_equal(v, O.gen_vals_i32(n, -(1 << 30), 1 << 30, 13))


@pytest.mark.parametrize("path", [tga.PATH_AUTO, tga.PATH_DIRECT, tga.PATH_PARTITIONED])
def test_u32_keys_count_sum(ctx, path):
    n, d = 1 << 20, 1 << 10
    k = ctx.generate(tga.GEN_KEYS_U32, n, a=d, seed=1, salt=3)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 20)) & (2**64 - 1), b=1 << 21, seed=2)
    got = ctx.groupby(k, [v], ["count", ("sum", 0)], path=path)
    ref = O.groupby(k.to_numpy(), [v.to_numpy()], ["count", ("sum", 0)])
    check_equal(got, ref)
    assert got["n_groups"] == d


@pytest.mark.parametrize("path", [tga.PATH_AUTO, tga.PATH_PARTITIONED, tga.PATH_DIRECT])
def test_i64_keys_four_aggs(ctx, path):
    n, d = 1 << 22, 1 << 16
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=5, salt=9)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=6)
    kh = k.to_numpy().view(np.int64)
    vh = v.to_numpy()
    got = ctx.groupby(k, [v], AGGS4, path=path)
    ref = O.groupby(kh.view(np.uint64), [vh], AGGS4)
    check_equal(got, ref)


@pytest.mark.parametrize("layout", [1, 2])
@pytest.mark.parametrize("ncols", [1, 2])
def test_partition_layouts(ctx, layout, ncols):
    # both partitioned-copy layouts: partition-contiguous (histogram + column
    # scan, then the scatter to scanned offsets) and per-workgroup buckets
    # (count-free scatter, bucket lists); one column (the single-column
    # scatter with two register sets) and two (the multi-column one); enough
    # tiles that the bucket reserves refill, and a last tile that is ragged
    n, d = (1 << 22) + 4097, 150_000
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=41, salt=2)
    cols = [ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=42 + c) for c in range(ncols)]
    aggs = AGGS4 + ([("sum", 1), ("min", 1)] if ncols > 1 else [])
    got = ctx.groupby(k, cols, aggs, path=tga.PATH_PARTITIONED, scatter_layout=layout)
    assert ctx.last_plan()["scatter_layout"] == layout
    ref = O.groupby(k.to_numpy(), [c.to_numpy() for c in cols], aggs)
    check_equal(got, ref)
    assert got["n_groups"] == d


@pytest.mark.parametrize("wc", [1, 2])
@pytest.mark.parametrize("key_bits", [32, 64])
def test_write_combining(ctx, wc, key_bits):
    # bucket layout of (key, value) records with and without the LDS carry
    # of unfinished cache lines, at 2^9 partitions (the most the carry
    # buffer fits): 8-byte records in 8192-row tiles, 12-byte records in
    # 6144-row tiles; a ragged last tile
    n, d = (1 << 22) + 4097, 150_000
    gen = tga.GEN_KEYS_U64 if key_bits == 64 else tga.GEN_KEYS_U32
    k = ctx.generate(gen, n, a=d, seed=43, salt=3)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=44)
    got = ctx.groupby(k, [v], AGGS4, path=tga.PATH_PARTITIONED, partition_bits=9, write_combine=wc)
    plan = ctx.last_plan()
    assert plan["write_combine"] == wc and plan["scatter_layout"] == 2
    if wc == 1:
        assert plan["tile_rows"] == (8192 if key_bits == 32 else 6144)
    ref = O.groupby(k.to_numpy(), [v.to_numpy()], AGGS4)
    check_equal(got, ref)
    assert got["n_groups"] == d


@pytest.mark.parametrize("bits", [12, 13])
@pytest.mark.parametrize("key_bits,hot", [(32, False), (64, False), (64, True)])
def test_two_pass_partitioning(ctx, key_bits, hot, bits):
    # 2^12 (2^13) partitions: records written whole-line into 2^8 (2^9)
    # partitions, each split again into 16; with a hot key holding half the
    # rows its first-pass partition spans several second-pass work items
    n, d = (1 << 22) + 333, 1 << 19
    gen = tga.GEN_KEYS_U64 if key_bits == 64 else tga.GEN_KEYS_U32
    k = ctx.generate(gen, n, a=d, seed=51, salt=4).to_numpy()
    if hot:
        k[::2] = 987654321
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=52).to_numpy()
    got = ctx.groupby(ctx.upload(k), [ctx.upload(v)], AGGS4, path=tga.PATH_PARTITIONED, partition_bits=bits)
    plan = ctx.last_plan()
    assert plan["partition_bits"] == bits and plan["write_combine"] == 1
    check_equal(got, O.groupby(k, [v], AGGS4))


def test_split_partitions_without_key_sample(ctx):
    # 40% of the rows on each of two keys (partitions 5 and 35 of the 64 the
    # tuner picks: tables a quarter full would take 16, and 2^20 rows keep
    # 16K rows in each of 64) and a group-count hint, so no key sample is
    # taken: both partitions are split into chunks, and the work plan sizes
    # the global table for their groups on the GPU, so one attempt does
    n = 1 << 20
    rng = np.random.default_rng(81)
    k = rng.integers(0, 1 << 62, size=20000, dtype=np.uint64)[rng.integers(0, 20000, size=n)]
    i = np.arange(n)
    k[i % 5 < 2] = 5
    k[(i % 5 >= 2) & (i % 5 < 4)] = 9
    v = rng.integers(-99, 100, size=n).astype(np.int32)
    got = ctx.groupby(k, [v], AGGS4, path=tga.PATH_PARTITIONED, est_groups=20002)
    plan = ctx.last_plan()
    assert plan["partition_bits"] == 6 and plan["attempts"] == 1
    check_equal(got, O.groupby(k, [v], AGGS4))


def test_mid_size_query_takes_a_partition_per_cu(ctx):
    # 2^22 rows over 2^14 groups: tables a quarter full would take 2^4
    # partitions, and the aggregation (a workgroup per partition) would run
    # on 16 CUs; the tuner raises the fanout to a partition per CU while
    # each partition keeps 16K rows (2^22 / 2^14 = 2^8 partitions at most)
    n, d = 1 << 22, 1 << 14
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=91, salt=6)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=92)
    got = ctx.groupby(k, [v], AGGS4, est_groups=d)
    plan = ctx.last_plan()
    cus = ctx.device_info()["cus"]
    want = min(max(1, (cus - 1).bit_length()), 8)
    assert plan["path"] == tga.PATH_PARTITIONED and plan["partition_bits"] == want and plan["attempts"] == 1, plan
    check_equal(got, O.groupby(k.to_numpy(), [v.to_numpy()], AGGS4))
    assert got["n_groups"] == d


def test_spills_overflow_planned_global_table(ctx):
    # 20000 groups over 4 partitions with 256-slot LDS tables: nearly every
    # group spills to the global table.  The plan sees its partitions hold far
    # more groups than their tables and takes the full-size global table from
    # the start (sized for split partitions only, it would overflow and the
    # call run again), so one attempt does
    n = 1 << 20
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=20000, seed=83, salt=5)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=84)
    got = ctx.groupby(k, [v], AGGS4, path=tga.PATH_PARTITIONED, partition_bits=2, table_slots=256, est_groups=20000)
    plan = ctx.last_plan()
    assert plan["partition_bits"] == 2 and plan["attempts"] == 1
    check_equal(got, O.groupby(k.to_numpy(), [v.to_numpy()], AGGS4))


def test_partitioned_small_tables_many_partitions(ctx):
    n, d = 1 << 21, 200_000
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=15, salt=1)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=16)
    got = ctx.groupby(k, [v], AGGS4, path=tga.PATH_PARTITIONED, partition_bits=10, table_slots=1024)
    ref = O.groupby(k.to_numpy(), [v.to_numpy()], AGGS4)
    check_equal(got, ref)


def test_edge_empty_and_single(ctx):
    k = np.zeros(0, np.uint32)
    v = np.zeros(0, np.int32)
    got = ctx.groupby(k, [v], ["count", ("sum", 0)])
    assert got["n_groups"] == 0
    got = ctx.groupby(np.array([7], np.uint32), [np.array([-3], np.int32)], AGGS4)
    assert got["n_groups"] == 1 and got["aggs"][0][0] == 1 and got["aggs"][1][0] == -3
    assert got["aggs"][2][0] == -3 and got["aggs"][3][0] == -3


@pytest.mark.parametrize("path", [tga.PATH_DIRECT, tga.PATH_PARTITIONED])
def test_single_key_max_contention(ctx, path):
    n = 1 << 20
    k = np.full(n, 42, np.uint64)
    v = O.gen_vals_i32(n, -(1 << 31), (1 << 31) - 1, 3)
    got = ctx.groupby(k, [v], AGGS4, path=path)
    ref = O.groupby(k, [v], AGGS4)
    check_equal(got, ref)


def test_all_distinct_and_sentinel_keys(ctx):
    n = 300_000
    k = np.arange(n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    k[:5] = np.uint64(2**64 - 1)  # the empty-slot marker as a real key
    k[5] = 0
    v = O.gen_vals_i32(n, -(1 << 31), (1 << 31) - 1, 4)
    v[:3] = np.iinfo(np.int32).min
    v[3] = np.iinfo(np.int32).max
    for path in (tga.PATH_DIRECT, tga.PATH_PARTITIONED):
        got = ctx.groupby(k, [v], AGGS4, path=path)
        ref = O.groupby(k, [v], AGGS4)
        check_equal(got, ref)
    k32 = (np.arange(n, dtype=np.uint64) % 1000).astype(np.uint32)
    k32[::7] = 0xFFFFFFFF
    got = ctx.groupby(k32, [v], AGGS4)
    check_equal(got, O.groupby(k32, [v], AGGS4))


def test_global_table_overflow_retry(ctx):
    n, d = 1 << 20, 100_000
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=21, salt=2)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=0, b=1000, seed=22)
    got = ctx.groupby(k, [v], AGGS4, est_groups=100)  # far too low: table must grow
    check_equal(got, O.groupby(k.to_numpy(), [v.to_numpy()], AGGS4))


def test_lds_spill_path(ctx):
    # more groups than the forced LDS table holds: rows spill to the global table
    n, d = 1 << 20, 20_000
    k = ctx.generate(tga.GEN_KEYS_U32, n, a=d, seed=31, salt=2)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=32)
    got = ctx.groupby(k, [v], AGGS4, path=tga.PATH_DIRECT, table_slots=256, est_groups=d)
    check_equal(got, O.groupby(k.to_numpy(), [v.to_numpy()], AGGS4))


def test_multi_column_dtypes(ctx):
    n = 200_000
    k = O.gen_keys_u32(n, 5000, 41, 0)
    a = O.gen_vals_i32(n, -(1 << 31), (1 << 31) - 1, 42)
    b = O.gen_keys_u64(n, 1 << 40, 43, 7)            # u64 column
    c = b.view(np.int64).copy()                        # i64 column
    d = a.view(np.uint32).copy()                       # u32 column
    aggs = [("sum", 0), ("min", 1), ("max", 1), ("min", 2), ("max", 2), ("sum", 3), ("max", 3), "count"]
    got = ctx.groupby(k, [a, b, c, d], aggs)
    check_equal(got, O.groupby(k, [a, b, c, d], aggs))


def test_fp64_sum_exact(ctx):
    n, d = 1 << 20, 3000
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=51, salt=3)
    x = ctx.generate(tga.GEN_NORMAL_F64, n, x=0.0, y=1e3, seed=52)
    kh, xh = k.to_numpy(), x.to_numpy()
    got = ctx.groupby(k, [x], ["count", ("sum", 0), ("min", 0), ("max", 0)])
    ref = O.groupby(kh, [xh], ["count", ("sum", 0), ("min", 0), ("max", 0)])
    np.testing.assert_array_equal(got["keys"], ref["keys"])
    np.testing.assert_array_equal(got["aggs"][0], ref["aggs"][0])
    np.testing.assert_array_equal(got["aggs"][2], ref["aggs"][2])
    np.testing.assert_array_equal(got["aggs"][3], ref["aggs"][3])
    s, r = got["aggs"][1], ref["aggs"][1]
    # tolerance of the compensated oracle: 1e-12 relative
    assert np.all(np.abs(s - r) <= 1e-12 * np.maximum(np.abs(r), 1e-300))
    # the GPU sum is the exact sum rounded once: equal to math.fsum
    order = np.argsort(kh, kind="stable")
    ks = kh[order]
    xs = xh[order]
    starts = np.r_[0, np.flatnonzero(np.diff(ks)) + 1, ks.size]
    for i in range(0, len(starts) - 1, 97):
        assert s[i] == math.fsum(xs[starts[i]:starts[i + 1]])


def test_scan_matches_prefix_sum(ctx):
    for n in (1, 4095, 4096, 4097, 1_000_003):
        x = (np.arange(n, dtype=np.uint64) * 2654435761 % 1000).astype(np.uint32)
        d_in = ctx.upload(x)
        d_out = ctx.empty(n, tga.U32)
        ctx.scan_exclusive_u32(d_in, d_out)
        exp = (O.scan(x.astype(np.uint64)) & 0xFFFFFFFF).astype(np.uint32)
        np.testing.assert_array_equal(d_out.to_numpy(), exp)


@pytest.mark.parametrize("path", [tga.PATH_AUTO, tga.PATH_DIRECT, tga.PATH_PARTITIONED])
def test_zipf_hot_keys_int_and_fp64(ctx, path):
    """Config-3 shape at reduced size: Zipf(1.1) keys (hot-key contention),
    an int32 column (SUM bit-exact) and an fp64 N(0, 1e3) column (SUM within
    1e-12 of the compensated oracle, equal to math.fsum)."""
    n, D = 1 << 22, 1 << 18
    k = ctx.generate(tga.GEN_ZIPF_U64, n, a=D, x=1.1, seed=61, salt=7)
    a = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=62)
    f = ctx.generate(tga.GEN_NORMAL_F64, n, x=0.0, y=1e3, seed=63)
    kh, ah, fh = k.to_numpy(), a.to_numpy(), f.to_numpy()
    aggs = ["count", ("sum", 0), ("sum", 1)]
    got = ctx.groupby(k, [a, f], aggs, path=path)
    ref = O.groupby(kh, [ah, fh], aggs)
    assert got["n_groups"] == ref["n_groups"]
    np.testing.assert_array_equal(got["keys"], ref["keys"])
    np.testing.assert_array_equal(got["aggs"][0], ref["aggs"][0])
    np.testing.assert_array_equal(got["aggs"][1], ref["aggs"][1])
    s, r = got["aggs"][2], ref["aggs"][2]
    assert np.all(np.abs(s - r) <= 1e-12 * np.maximum(np.abs(r), 1e-300))
    # the hottest key holds a large share of the rows: its fp64 sum is exact
    top = int(np.argmax(got["aggs"][0]))
    assert got["aggs"][0][top] > n // 20
    assert s[top] == math.fsum(fh[kh == got["keys"][top]].tolist())


@pytest.mark.parametrize("schema,bits", [("i32_f64", 7), ("i32_f64", 9), ("i32_f64", 12), ("i32_f64", 14),
                                         ("u64", 11), ("f64_f64", 10), ("4xi32", 10), ("i32_f64", 0)])
def test_record_copy_several_value_words(ctx, schema, bits):
    """The partitioned copy as records of the key and several value words
    (columns in order at their own widths), written by the write-combining
    scatter in one pass or two (up to 2^14 partitions), read back by the
    aggregation: integer aggregates bit-exact, fp64 SUM within 1e-12 relative
    of the compensated oracle.  bits 0: the tuner's plan for Zipf(1.1) keys."""
    n = 1 << 22
    D = 1 << 20
    if bits == 0:
        k = ctx.generate(tga.GEN_ZIPF_U64, n, a=D, x=1.1, seed=81, salt=4)
    else:
        k = ctx.generate(tga.GEN_KEYS_U64, n, a=D, seed=81, salt=4)
    i32 = lambda s: ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=s)
    f64 = lambda s: ctx.generate(tga.GEN_NORMAL_F64, n, x=0.0, y=1e3, seed=s)
    if schema == "i32_f64":
        cols, aggs = [i32(82), f64(83)], ["count", ("sum", 0), ("sum", 1), ("min", 1), ("max", 0)]
    elif schema == "u64":
        cols, aggs = [ctx.generate(tga.GEN_KEYS_U64, n, a=1 << 40, seed=84, salt=1)], AGGS4
    elif schema == "f64_f64":
        cols, aggs = [f64(85), f64(86)], [("sum", 0), ("sum", 1), ("max", 0), ("min", 1), "count"]
    else:
        cols = [i32(87 + c) for c in range(4)]
        aggs = [("sum", 0), ("min", 1), ("max", 2), ("sum", 3), "count"]
    opts = {} if bits == 0 else {"path": tga.PATH_PARTITIONED, "partition_bits": bits, "est_groups": D}
    got = ctx.groupby(k, cols, aggs, **opts)
    plan = ctx.last_plan()
    assert plan["record_words"] > 1 and plan["write_combine"] == 1, plan
    if bits:
        assert plan["partition_bits"] == bits
    ref = O.groupby(k.to_numpy(), [c.to_numpy() for c in cols], aggs)
    assert got["n_groups"] == ref["n_groups"]
    np.testing.assert_array_equal(got["keys"], ref["keys"])
    for a, g, r in zip(aggs, got["aggs"], ref["aggs"]):
        if g.dtype == np.float64 and a != "count" and a[0] == "sum":
            assert np.all(np.abs(g - r) <= 1e-12 * np.maximum(np.abs(r), 1e-300))
        else:
            np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("bits", [32, 64])
@pytest.mark.parametrize("path", [tga.PATH_AUTO, tga.PATH_DIRECT, tga.PATH_PARTITIONED])
def test_closed_form_columns_every_group(ctx, bits, path):
    """Closed-form columns (key = a bijective mix of i mod D, value = lo + i):
    the device generators equal the host restatement, and every group's
    COUNT / SUM / MIN / MAX equals its closed form."""
    from closed_form import columns, expected
    for n, D, lo in ((1_000_003, 1 << 10, -(1 << 20)), (3_000_000, 50_000, 7)):
        kind = tga.GEN_MOD_U32 if bits == 32 else tga.GEN_MOD_U64
        k = ctx.generate(kind, n, a=D, salt=0xC0DE)
        v = ctx.generate(tga.GEN_SEQ_I32, n, a=lo & (2**64 - 1))
        if n <= 1_000_003:
            kh, vh = columns(n, D, bits, 0xC0DE, lo)
            np.testing.assert_array_equal(k.to_numpy(), kh)
            np.testing.assert_array_equal(v.to_numpy(), vh)
        got = ctx.groupby(k, [v], AGGS4, path=path, est_groups=D)
        ek, ea = expected(n, D, bits, 0xC0DE, lo)
        np.testing.assert_array_equal(got["keys"], ek)
        for g, e in zip(got["aggs"], ea):
            np.testing.assert_array_equal(g, e)


@pytest.mark.parametrize("path", [tga.PATH_DIRECT, tga.PATH_PARTITIONED])
@pytest.mark.parametrize("keys", ["one_hot_key", "zipf"])
def test_wave_fold_on(ctx, path, keys):
    """wave_fold=1: equal keys of a wave are folded into one lane (ballots on
    the slot bits) before the LDS atomics.  Bit-exact against the oracle on
    maximum contention (one key) and Zipf(1.1) hot keys."""
    n = 1 << 21
    if keys == "one_hot_key":
        kh = np.full(n, 0x1234_5678_9ABC, np.uint64)
        kh[::1000] = np.arange(kh[::1000].size, dtype=np.uint64)  # a few cold keys beside it
        k = ctx.upload(kh)
    else:
        k = ctx.generate(tga.GEN_ZIPF_U64, n, a=1 << 16, x=1.1, seed=71, salt=3)
        kh = k.to_numpy()
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=72)
    got = ctx.groupby(k, [v], AGGS4, path=path, wave_fold=1)
    assert ctx.last_plan()["wave_fold"] == 1
    check_equal(got, O.groupby(kh, [v.to_numpy()], AGGS4))
