"""Error paths of the C ABI on the GPU, each forced from the host: invalid
arguments, fp64 SUM inputs outside the exact window (TGA_ERR_RANGE), a
caller-provided output smaller than the group count (TGA_ERR_CAPACITY, via
tga_groupby_into, whose normal result is checked against the oracle too) and
a served request that is still pending (TGA_ERR_NOT_READY).  Every kernel
launched here exits normally."""
import ctypes

import numpy as np
import pytest

import cpu_gpu_coprocess_groupby_agg_amd as tga
import oracle as O
from cpu_gpu_coprocess_groupby_agg_amd import lib

pytestmark = pytest.mark.gpu

AGGS4 = ["count", ("sum", 0), ("min", 0), ("max", 0)]


def _call(ctx, key_col, n, cols, aggs):
    """tga_groupby through ctypes with raw column descriptors; returns the status."""
    vc = (tga.Column * max(1, len(cols)))(*cols)
    ag = (tga.Agg * max(1, len(aggs)))(*[tga.Agg(o, c) for o, c in aggs])
    r = ctypes.c_void_p()
    st = lib().tga_groupby(ctx.h, ctypes.byref(key_col), vc, len(cols), n, ag, len(aggs), None, ctypes.byref(r))
    if r.value:
        lib().tga_result_free(r.value)
    return st


def test_invalid_arguments(ctx):
    n = 1000
    k = ctx.generate(tga.GEN_KEYS_U32, n, a=10, seed=1)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=0, b=100, seed=2)
    kc, vcol = k.column(), v.column()
    sum0 = [(tga.SUM, 0)]
    cases = {
        "negative rows": (kc, -1, [vcol], sum0),
        "no aggregates": (kc, n, [vcol], []),
        "aggregate column out of range": (kc, n, [vcol], [(tga.SUM, 3)]),
        "unknown aggregate op": (kc, n, [vcol], [(9, 0)]),
        "f64 key dtype": (tga.Column(k.ptr, tga.F64, 0), n, [vcol], sum0),
        "null key data": (tga.Column(None, tga.U32, 0), n, [vcol], sum0),
        "null value data": (kc, n, [tga.Column(None, tga.I32, 0)], sum0),
        "bad value dtype": (kc, n, [tga.Column(v.ptr, 77, 0)], sum0),
        "five value columns": (kc, n, [vcol] * 5, sum0),
        "nine aggregates": (kc, n, [vcol], [(tga.COUNT, 0)] * 9),
        "three fp64 sums": (kc, n, [tga.Column(v.ptr, tga.F64, 0)], [(tga.SUM, 0)] * 3),
    }
    for name, (key_col, rows, cols, aggs) in cases.items():
        st = _call(ctx, key_col, rows, cols, aggs)
        assert st == tga.ERR_INVALID_ARG, (name, st)
        assert lib().tga_last_error(ctx.h), name
    # null handles and malformed utility calls
    assert lib().tga_groupby(None, None, None, 0, 0, None, 0, None, None) == tga.ERR_INVALID_ARG
    assert lib().tga_result_free(None) == tga.ERR_INVALID_ARG
    assert lib().tga_scan_exclusive_u32(ctx.h, None, None, 5) == tga.ERR_INVALID_ARG
    assert lib().tga_device_copy(ctx.h, k.ptr, v.ptr, 15) == tga.ERR_INVALID_ARG  # not a multiple of 16
    assert lib().tga_generate(ctx.h, tga.GEN_KEYS_U32, k.ptr, 10, 0, 0, 0.0, 0.0, 1, 0) == tga.ERR_INVALID_ARG
    # the context still works after every rejected call
    got = ctx.groupby(k, [v], ["count", ("sum", 0)])
    ref = O.groupby(k.to_numpy(), [v.to_numpy()], ["count", ("sum", 0)])
    np.testing.assert_array_equal(got["keys"], ref["keys"])
    np.testing.assert_array_equal(got["aggs"][1], ref["aggs"][1])


@pytest.mark.parametrize("bad", [np.inf, -np.inf, np.nan, 2.0 ** 64, -(2.0 ** 65)])
@pytest.mark.parametrize("path", [tga.PATH_DIRECT, tga.PATH_PARTITIONED])
def test_fp64_sum_outside_exact_window(ctx, bad, path):
    n = 200_000
    kh = O.gen_keys_u64(n, 5000, 81, 1)
    xh = np.linspace(-1e3, 1e3, n)
    xh[n // 3] = bad
    with pytest.raises(tga.TgaError) as e:
        ctx.groupby(kh, [xh], ["count", ("sum", 0)], path=path)
    assert e.value.status == tga.ERR_RANGE
    # MIN / MAX of the same column take any finite or infinite value; only SUM is range-checked
    if np.isinf(bad):
        got = ctx.groupby(kh, [xh], ["count", ("min", 0), ("max", 0)], path=path)
        ref = O.groupby(kh, [xh], ["count", ("min", 0), ("max", 0)])
        for g, r in zip(got["aggs"], ref["aggs"]):
            np.testing.assert_array_equal(g, r)


def _into(ctx, k, v, cap, aggs=AGGS4):
    al = [(tga._OPS[a] if isinstance(a, str) else tga._OPS[a[0]], 0 if isinstance(a, str) else a[1]) for a in aggs]
    ag = (tga.Agg * len(al))(*[tga.Agg(o, c) for o, c in al])
    outk = ctx.empty(max(cap, 1), k.dtype)
    outs = [ctx.empty(max(cap, 1), tga.out_dtype(o, v.dtype)) for o, _ in al]
    ptrs = (ctypes.c_void_p * len(outs))(*[o.ptr for o in outs])
    vc = (tga.Column * 1)(v.column())
    kc = k.column()
    g = ctypes.c_int64(-1)
    st = lib().tga_groupby_into(ctx.h, ctypes.byref(kc), vc, 1, k.n, ag, len(al), None, outk.ptr, ptrs, cap,
                                ctypes.byref(g))
    return st, g.value, outk, outs


@pytest.mark.parametrize("n,d", [(1 << 20, 1000), (1 << 21, 200_000)])
def test_groupby_into_and_capacity(ctx, n, d):
    k = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=91, salt=4)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=92)
    ref = O.groupby(k.to_numpy(), [v.to_numpy()], AGGS4)
    G = ref["n_groups"]
    # exact capacity: the caller's buffers hold the result
    st, g, outk, outs = _into(ctx, k, v, G)
    assert st == tga.OK and g == G
    kk = outk.to_numpy()[:G]
    o = np.argsort(kk, kind="stable")
    np.testing.assert_array_equal(kk[o], ref["keys"])
    for x, r in zip(outs, ref["aggs"]):
        np.testing.assert_array_equal(x.to_numpy()[:G][o], r)
    # one row short: TGA_ERR_CAPACITY, the true count reported
    st, g, _, _ = _into(ctx, k, v, G - 1)
    assert st == tga.ERR_CAPACITY and g == G
    st, g, _, _ = _into(ctx, k, v, 0)
    assert st == tga.ERR_CAPACITY and g == G


def test_server_poll_not_ready_then_result(ctx):
    n, d = 1 << 26, 1 << 20
    from closed_form import expected
    k = ctx.generate(tga.GEN_MOD_U64, n, a=d, salt=77)
    v = ctx.generate(tga.GEN_SEQ_I32, n, a=0)
    aggs = (tga.Agg * 4)(tga.Agg(tga.COUNT, 0), tga.Agg(tga.SUM, 0), tga.Agg(tga.MIN, 0), tga.Agg(tga.MAX, 0))
    vc = (tga.Column * 1)(v.column())
    q = tga.Query()
    q.key = k.column()
    q.values = ctypes.cast(vc, ctypes.POINTER(tga.Column))
    q.n_values = 1
    q.aggs = ctypes.cast(aggs, ctypes.POINTER(tga.Agg))
    q.n_aggs = 4
    q.n_rows = n
    q.est_groups = d
    srv = ctypes.c_void_p()
    ctx._check(lib().tga_server_create(ctx.h, 16 << 30, 8, ctypes.byref(srv)))
    try:
        t = ctypes.c_int64()
        ctx._check(lib().tga_server_submit(srv, ctypes.byref(q), ctypes.byref(t)))
        r = ctypes.c_void_p()
        # a 2^26-row request cannot have finished microseconds after its submission
        assert lib().tga_server_poll(srv, t.value, ctypes.byref(r)) == tga.ERR_NOT_READY
        assert lib().tga_server_poll(srv, t.value + 100, ctypes.byref(r)) == tga.ERR_INVALID_ARG
        while (st := lib().tga_server_poll(srv, t.value, ctypes.byref(r))) == tga.ERR_NOT_READY:
            pass
        assert st == tga.OK
        res = ctx._fetch(r.value, tga.U64, [tga.I64, tga.I64, tga.I32, tga.I32], True)
        lib().tga_result_free(r.value)
        ek, ea = expected(n, d, 64, 77, 0)
        np.testing.assert_array_equal(res["keys"], ek)
        for g, e in zip(res["aggs"], ea):
            np.testing.assert_array_equal(g, e)
        # a taken result is gone: its ticket is no longer known
        assert lib().tga_server_poll(srv, t.value, ctypes.byref(r)) == tga.ERR_INVALID_ARG
    finally:
        lib().tga_server_destroy(srv)
