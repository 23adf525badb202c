"""GPU tests of the batched persistent path and the serving loop: every query's
result equals the oracle and the plain operator run alone (bit-exact), across
graph replays and inside a mix of concurrent requests."""
import ctypes

import numpy as np
import pytest

import cpu_gpu_coprocess_groupby_agg_amd as tga
import oracle as O
from cpu_gpu_coprocess_groupby_agg_amd import lib

pytestmark = pytest.mark.gpu


def _queries(ctx, specs):
    """specs: list of (rows, distinct, key_bits, hint).  Returns device columns,
    the ctypes query array and host copies."""
    aggs = (tga.Agg * 4)(tga.Agg(tga.COUNT, 0), tga.Agg(tga.SUM, 0), tga.Agg(tga.MIN, 0), tga.Agg(tga.MAX, 0))
    qs = (tga.Query * len(specs))()
    keep, host = [], []
    for i, (n, d, bits, hint) in enumerate(specs):
        kind = tga.GEN_KEYS_U32 if bits == 32 else tga.GEN_KEYS_U64
        k = ctx.generate(kind, n, a=d, seed=100 + i, salt=i) if n else ctx.empty(0, tga.U32 if bits == 32 else tga.U64)
        v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=200 + i) if n \
            else ctx.empty(0, tga.I32)
        vc = (tga.Column * 1)(v.column())
        keep += [k, v, vc]
        qs[i].key = k.column()
        qs[i].values = ctypes.cast(vc, ctypes.POINTER(tga.Column))
        qs[i].n_values = 1
        qs[i].aggs = ctypes.cast(aggs, ctypes.POINTER(tga.Agg))
        qs[i].n_aggs = 4
        qs[i].n_rows = n
        qs[i].est_groups = d if hint else 0
        host.append((k.to_numpy(), v.to_numpy()))
    keep.append(aggs)
    return qs, keep, host


AGGS4 = ["count", ("sum", 0), ("min", 0), ("max", 0)]

SPECS = [
    (1 << 10, 16, 32, True), (5000, 5000, 32, True), (1 << 16, 1 << 12, 32, True),
    (1 << 18, 1 << 16, 32, True),   # large cardinality: partitioned items
    (1 << 20, 1 << 18, 32, True),   # partitioned
    (300_000, 100_000, 64, True),    # 64-bit keys, partitioned
    (1 << 17, 30, 64, False),        # estimate from a sample
    (0, 1, 32, True),                # empty query
    (1, 1, 32, True),
    (200_000, 50, 32, False),
    (100_000, 90_000, 32, False),
    (70_000, 64, 32, True),
    (1 << 20, 200, 32, True),        # direct, several row-range items: flushes and compaction
]


def _batch_result(b, i, nq_dtype_key, G):
    keys = np.empty(G, nq_dtype_key)
    outs = []
    if G:
        assert lib().tga_batch_copy_to_host(b, i, -1, keys.ctypes.data, G) == 0
    for a, dt in enumerate([np.int64, np.int64, np.int32, np.int32]):
        x = np.empty(G, dt)
        if G:
            assert lib().tga_batch_copy_to_host(b, i, a, x.ctypes.data, G) == 0
        outs.append(x)
    o = np.argsort(keys.view(np.uint32 if keys.itemsize == 4 else np.uint64), kind="stable")
    return keys[o], [x[o] for x in outs]


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_matches_oracle(ctx, mode):
    qs, keep, host = _queries(ctx, SPECS)
    b = ctypes.c_void_p()
    ctx._check(lib().tga_batch_create(ctx.h, qs, len(SPECS), mode, ctypes.byref(b)))
    try:
        refs = [O.groupby(k, [v], AGGS4) for k, v in host]
        for rep in range(3):  # graph replays must give the same results every time
            ctx._check(lib().tga_batch_run(b))
            for i, (n, d, bits, _) in enumerate(SPECS):
                G = ctypes.c_int64()
                ctx._check(lib().tga_batch_num_groups(b, i, ctypes.byref(G)))
                ref = refs[i]
                assert G.value == ref["n_groups"], (rep, i)
                k, outs = _batch_result(b, i, np.uint32 if bits == 32 else np.uint64, G.value)
                np.testing.assert_array_equal(k, ref["keys"])
                for x, r in zip(outs, ref["aggs"]):
                    np.testing.assert_array_equal(x, r)
            # eager work between replays
            ctx.groupby(host[2][0], [host[2][1]], ["count"])
    finally:
        lib().tga_batch_destroy(b)


def test_batch_profile_per_item_type(ctx):
    # tga_batch_profile: per item type (INIT, COUNT, DIRECT, SCAN, SCATTER,
    # PAGG, COMPACT) the items run and their workgroup-ms, for a batch built
    # with kernel timing on; TGA_ERR_UNSUPPORTED for one built without
    qs, keep, host = _queries(ctx, SPECS)
    n = len(SPECS)
    arrs = lambda: ((ctypes.c_double * 7)(), (ctypes.c_double * 7)(), (ctypes.c_int64 * 7)(),
                    (ctypes.c_double * 7)(), (ctypes.c_double * 8)())
    b = ctypes.c_void_p()
    ctx._check(lib().tga_batch_create(ctx.h, qs, n, 0, ctypes.byref(b)))
    try:
        ctx._check(lib().tga_batch_run(b))
        assert lib().tga_batch_profile(b, *arrs()) == tga.ERR_UNSUPPORTED
    finally:
        lib().tga_batch_destroy(b)
    ctx.profile(True)
    b = ctypes.c_void_p()
    try:
        ctx._check(lib().tga_batch_create(ctx.h, qs, n, 0, ctypes.byref(b)))
        ctx._check(lib().tga_batch_run(b))
        work, wait, items, fence, phase = arrs()
        ctx._check(lib().tga_batch_profile(b, work, wait, items, fence, phase))
        names = ["INIT", "COUNT", "DIRECT", "SCAN", "SCATTER", "PAGG", "COMPACT"]
        got = {t: items[i] for i, t in enumerate(names)}
        # every query clears its table and compacts it; partitioned queries
        # (three hinted ones in the mix, and a sampled one may be) take a
        # histogram scan each, tiles and partitions
        assert got["INIT"] >= n and got["COMPACT"] >= n and 3 <= got["SCAN"] <= 4, got
        assert got["DIRECT"] > 0 and got["COUNT"] == got["SCATTER"] > 0 and got["PAGG"] > 0, got
        for i, t in enumerate(names):
            assert work[i] >= 0 and wait[i] >= 0 and fence[i] >= 0
            if got[t]:
                assert work[i] > 0, t
        assert all(x >= 0 for x in phase)
    finally:
        if b.value:
            lib().tga_batch_destroy(b)
        ctx.profile(False)


def test_batch_underestimate_reruns(ctx):
    # hints far below the truth: table overflow is detected and the query
    # rerun alone, at every replay.  A second batch built between replays and
    # held across the next one takes its small buffers from the arena's free
    # lists: the first batch must never write into memory it released.
    specs = [(1 << 18, 100_000, 32, True), (20_000, 6000, 64, True), (70_000, 64, 32, True)]
    qs, keep, host = _queries(ctx, specs)
    qs[0].est_groups = 10
    qs[1].est_groups = 10
    refs = [O.groupby(k, [v], AGGS4) for k, v in host]
    b = ctypes.c_void_p()
    ctx._check(lib().tga_batch_create(ctx.h, qs, 2, 0, ctypes.byref(b)))
    try:
        for rep in range(3):
            held = ctypes.c_void_p()
            if rep:
                ctx._check(lib().tga_batch_create(ctx.h, ctypes.byref(qs[2]), 1, 0, ctypes.byref(held)))
                ctx._check(lib().tga_batch_run(held))
            ctx._check(lib().tga_batch_run(b))
            for i in range(2):
                G = ctypes.c_int64()
                ctx._check(lib().tga_batch_num_groups(b, i, ctypes.byref(G)))
                assert G.value == refs[i]["n_groups"], (rep, i)
                k, outs = _batch_result(b, i, np.uint32 if specs[i][2] == 32 else np.uint64, G.value)
                np.testing.assert_array_equal(k, refs[i]["keys"])
                for x, r in zip(outs, refs[i]["aggs"]):
                    np.testing.assert_array_equal(x, r)
            if rep:
                G = ctypes.c_int64()
                ctx._check(lib().tga_batch_num_groups(held, 0, ctypes.byref(G)))
                assert G.value == refs[2]["n_groups"]
                k, outs = _batch_result(held, 0, np.uint32, G.value)
                np.testing.assert_array_equal(k, refs[2]["keys"])
                for x, r in zip(outs, refs[2]["aggs"]):
                    np.testing.assert_array_equal(x, r)
                lib().tga_batch_destroy(held)
    finally:
        lib().tga_batch_destroy(b)


def _result_arrays(r, key_dtype):
    G = ctypes.c_int64()
    assert lib().tga_result_num_groups(r, ctypes.byref(G)) == 0
    G = G.value
    keys = np.empty(G, key_dtype)
    if G:
        assert lib().tga_result_copy_to_host(r, -1, keys.ctypes.data, G) == 0
    outs = []
    for a, dt in enumerate([np.int64, np.int64, np.int32, np.int32]):
        x = np.empty(G, dt)
        if G:
            assert lib().tga_result_copy_to_host(r, a, x.ctypes.data, G) == 0
        outs.append(x)
    o = np.argsort(keys.view(np.uint32 if keys.itemsize == 4 else np.uint64), kind="stable")
    lib().tga_result_free(r)
    return keys[o], [x[o] for x in outs]


@pytest.mark.parametrize("budget", [64 << 30, 96 << 20])
def test_server_mix_equals_alone(ctx, budget):
    specs = [s for s in SPECS if s[0] > 0] * 3
    if budget > 1 << 30:
        # requests of >= 2^24 rows run alone through the plain operator, beside the batched flights
        specs[5:5] = [(1 << 24, 1 << 16, 64, True)]
        specs += [(1 << 24, 5000, 32, False)]
    qs, keep, host = _queries(ctx, specs)
    srv = ctypes.c_void_p()
    ctx._check(lib().tga_server_create(ctx.h, budget, 8, ctypes.byref(srv)))
    try:
        tickets = []
        for i in range(len(specs)):
            t = ctypes.c_int64()
            ctx._check(lib().tga_server_submit(srv, ctypes.byref(qs[i]), ctypes.byref(t)))
            tickets.append(t.value)
        for i, t in enumerate(tickets):
            r = ctypes.c_void_p()
            ctx._check(lib().tga_server_wait(srv, t, ctypes.byref(r)))
            k, outs = _result_arrays(r, np.uint32 if specs[i][2] == 32 else np.uint64)
            alone = ctx.groupby(host[i][0], [host[i][1]], AGGS4)  # the same query, alone
            np.testing.assert_array_equal(k, alone["keys"])
            for x, y in zip(outs, alone["aggs"]):
                np.testing.assert_array_equal(x, y)
            ref = O.groupby(host[i][0], [host[i][1]], AGGS4)  # and the oracle
            np.testing.assert_array_equal(k, ref["keys"])
            for x, y in zip(outs, ref["aggs"]):
                np.testing.assert_array_equal(x, y)
        st = tga.ServerStats()
        ctx._check(lib().tga_server_stats_get(srv, ctypes.byref(st)))
        assert st.completed == len(specs) and st.p99_us >= st.p50_us > 0
        assert st.peak_bytes <= budget
    finally:
        lib().tga_server_destroy(srv)


def test_server_rejects_over_budget(ctx):
    qs, keep, host = _queries(ctx, [(1 << 20, 1 << 18, 32, True)])
    srv = ctypes.c_void_p()
    ctx._check(lib().tga_server_create(ctx.h, 1 << 20, 8, ctypes.byref(srv)))
    t = ctypes.c_int64()
    assert lib().tga_server_submit(srv, ctypes.byref(qs[0]), ctypes.byref(t)) == tga.ERR_BUDGET
    st = tga.ServerStats()
    ctx._check(lib().tga_server_stats_get(srv, ctypes.byref(st)))
    assert st.rejected == 1
    lib().tga_server_destroy(srv)


def test_server_poisson_arrivals_with_large_requests(ctx):
    """Requests arrive on a seeded Poisson schedule (mean gap 300 us) while
    earlier ones are on the GPU: 95% from the batched mix, 5% 2^26-row
    int64-key requests (closed-form columns).  Every served result equals
    the oracle (small requests) or its closed form (the large ones), and the
    same small request alone through the plain operator."""
    import time
    from closed_form import expected
    rng = np.random.default_rng(2026)
    small = [s for s in SPECS if s[0] > 0]
    qs_small, keep, host = _queries(ctx, small)
    refs = [O.groupby(k, [v], AGGS4) for k, v in host]
    big_n = 1 << 26
    bigs = []
    aggs4 = (tga.Agg * 4)(tga.Agg(tga.COUNT, 0), tga.Agg(tga.SUM, 0), tga.Agg(tga.MIN, 0), tga.Agg(tga.MAX, 0))
    for j, (d, salt, hint) in enumerate(((1 << 20, 501, True), (3 << 16, 502, False))):
        k = ctx.generate(tga.GEN_MOD_U64, big_n, a=d, salt=salt)
        v = ctx.generate(tga.GEN_SEQ_I32, big_n, a=j * 1000)
        vc = (tga.Column * 1)(v.column())
        q = tga.Query()
        q.key = k.column()
        q.values = ctypes.cast(vc, ctypes.POINTER(tga.Column))
        q.n_values = 1
        q.aggs = ctypes.cast(aggs4, ctypes.POINTER(tga.Agg))
        q.n_aggs = 4
        q.n_rows = big_n
        q.est_groups = d if hint else 0
        keep += [k, v, vc]
        bigs.append((q, expected(big_n, d, 64, salt, j * 1000)))
    nreq = 120
    kinds = []  # ("s", index) or ("b", index)
    for i in range(nreq):
        kinds.append(("b", int(rng.integers(len(bigs)))) if (rng.random() < 0.05 or i in (3, 40)) else
                     ("s", int(rng.integers(len(small)))))
    gaps = rng.exponential(300e-6, nreq)
    srv = ctypes.c_void_p()
    ctx._check(lib().tga_server_create(ctx.h, 64 << 30, 64, ctypes.byref(srv)))
    try:
        tickets = []
        t_next = time.perf_counter()
        for (kind, j), gap in zip(kinds, gaps):
            t_next += gap
            while time.perf_counter() < t_next:
                pass
            q = qs_small[j] if kind == "s" else bigs[j][0]
            t = ctypes.c_int64()
            ctx._check(lib().tga_server_submit(srv, ctypes.byref(q), ctypes.byref(t)))
            tickets.append(t.value)
        for (kind, j), t in zip(kinds, tickets):
            r = ctypes.c_void_p()
            ctx._check(lib().tga_server_wait(srv, t, ctypes.byref(r)))
            if kind == "s":
                k, outs = _result_arrays(r, np.uint32 if small[j][2] == 32 else np.uint64)
                ek, ea = refs[j]["keys"], refs[j]["aggs"]
            else:
                k, outs = _result_arrays(r, np.uint64)
                ek, ea = bigs[j][1]
            np.testing.assert_array_equal(k, ek)
            for x, y in zip(outs, ea):
                np.testing.assert_array_equal(x, y)
        st = tga.ServerStats()
        ctx._check(lib().tga_server_stats_get(srv, ctypes.byref(st)))
        assert st.completed == nreq and st.inflight == 0
    finally:
        lib().tga_server_destroy(srv)
