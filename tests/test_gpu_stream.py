"""Host-resident input streamed through HBM (the reference's tile-by-tile
data path, group_then_partition.cu:353-354, 413-418, 466-499): inputs larger
than the context's device budget, in pinned and in pageable host memory,
checked against closed forms (every group) and the oracle."""
import numpy as np
import pytest

import cpu_gpu_coprocess_groupby_agg_amd as tga
import oracle as O
from closed_form import columns, expected

pytestmark = pytest.mark.gpu

AGGS4 = ["count", ("sum", 0), ("min", 0), ("max", 0)]


@pytest.fixture(scope="module")
def small_ctx():
    c = tga.Context(0, budget_bytes=512 << 20)
    yield c
    c.close()


@pytest.mark.parametrize("pinned", [True, False])
def test_stream_larger_than_budget_every_group(small_ctx, pinned):
    n, d = 1 << 26, 1 << 20  # 805 MB of columns against a 512 MB budget
    kh, vh = columns(n, d, 64, 0x51, 0)
    if pinned:
        k, v = small_ctx.pinned_empty(n, np.uint64), small_ctx.pinned_empty(n, np.int32)
        k[:] = kh
        v[:] = vh
    else:
        k, v = kh, vh
    got = small_ctx.groupby(k, [v], AGGS4)
    plan = small_ctx.last_plan()
    assert plan["streamed_tiles"] > 1 and plan["stream_pinned"] == (1 if pinned else 0), plan
    ek, ea = expected(n, d, 64, 0x51, 0)
    assert got["n_groups"] == d
    np.testing.assert_array_equal(got["keys"], ek)
    for g, e in zip(got["aggs"], ea):
        np.testing.assert_array_equal(g, e)


def test_stream_fp64_and_hot_key_against_oracle(small_ctx):
    n = (1 << 23) + 777
    kh = O.gen_keys_u64(n, 300_000, 61, 5)
    kh[::7] = np.uint64(12345)  # a hot key
    ah = O.gen_vals_i32(n, -(1 << 31), (1 << 31) - 1, 62)
    fh = np.random.default_rng(63).normal(0.0, 1e3, n)
    k, a, f = (small_ctx.pinned_empty(n, x.dtype) for x in (kh, ah, fh))
    k[:], a[:], f[:] = kh, ah, fh
    aggs = ["count", ("sum", 0), ("min", 0), ("max", 0), ("sum", 1)]
    got = small_ctx.groupby(k, [a, f], aggs)
    assert small_ctx.last_plan()["streamed_tiles"] > 1
    ref = O.groupby(kh, [ah, fh], aggs)
    assert got["n_groups"] == ref["n_groups"]
    np.testing.assert_array_equal(got["keys"], ref["keys"])
    for g, r in zip(got["aggs"][:4], ref["aggs"][:4]):
        np.testing.assert_array_equal(g, r)
    s, r = got["aggs"][4], ref["aggs"][4]
    assert np.all(np.abs(s - r) <= 1e-12 * np.maximum(np.abs(r), 1e-300))


def test_stream_device_value_column_beside_host_key(ctx):
    # a host key column with a device-resident value column: only the key streams
    n, d = (1 << 22) + 3, 5000
    kh, vh = columns(n, d, 32, 7, -1000)
    v = ctx.upload(vh)
    got = ctx.groupby(kh, [v], AGGS4)
    assert ctx.last_plan()["streamed_tiles"] >= 1
    ek, ea = expected(n, d, 32, 7, -1000)
    np.testing.assert_array_equal(got["keys"], ek)
    for g, e in zip(got["aggs"], ea):
        np.testing.assert_array_equal(g, e)
