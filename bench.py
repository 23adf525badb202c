#!/usr/bin/env python3
"""Benchmark of the MI355X tiled hash group-by (one GPU).

Headline (one JSON line on stdout): input rows/s of the plain operator on the
2^28-row group-by -- int64 keys uniform over 2^20 distinct values, int32
values uniform in [-2^30, 2^30), COUNT / SUM(int64) / MIN / MAX -- with the
columns resident in HBM when the timed region starts.  A step is one complete
group-by (all kernels, the group count read back).  Beside it:
  roofline      dominant kernel's algorithmic bytes / its HIP-event time
  cpu_baseline  the restated reference algorithm on host threads, on a
                2^24-row sample of the same workload
  batched       4096 queries (rows log-uniform in [2^10, 2^22], distinct keys
                log-uniform in [2^4, 2^18], u32 keys, int32 values,
                COUNT+SUM): one persistent launch vs one launch per query
  serving       Poisson arrivals (mean 20k/s) from that query mix plus 5%
                2^26-row int64-key queries, 64 GB admission budget: p50/p99
                and throughput vs the sequential plain operator
Every aggregate result is checked (group counts, checksums) on the way.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "input rows/s + % HBM peak at 2^28 rows; batched queries/s & p99 vs per-query"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows-log2", type=int, default=28)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the reference-workload section (its driver's uniform / Zipf / self-similar keys)")
    ap.add_argument("--batch-queries", type=int, default=4096)
    ap.add_argument("--serve-seconds", type=float, default=0.5)
    ap.add_argument("--only", choices=["all", "headline", "copy", "workloads"], default="all",
                    help="headline: only the 2^28-row group-by (profiling runs); copy: only the copy reference; "
                         "workloads: only the reference-workload section")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes that fill roofline.traffic")
    return ap.parse_args()


# profiler name of a timed step -> kernel symbols it may launch (the scatter
# is the write-combining one or the plain one, as the plan says)
KERNEL_SYMBOL = {"part_scatter": ("k_part_scatter_wc", "k_part_scatter"), "agg_partitions": ("k_agg",),
                 "agg_direct": ("k_agg",), "compact": ("k_compact",), "copy16": ("k_copy16",)}


def pmc_pass(counter, argv, workdir):
    """One rocprofv3 counter pass over a child `python3 bench.py ...`; returns
    {kernel symbol prefix: [per-dispatch values in KiB]} or None."""
    import csv
    import shutil
    import subprocess
    exe = shutil.which("rocprofv3")
    if not exe:
        return None
    out = os.path.join(workdir, counter.lower())
    cmd = [exe, "--pmc", counter, "-f", "csv", "-d", out, "-o", "run", "--", sys.executable, os.path.abspath(__file__)]
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    try:
        r = subprocess.run(cmd + argv, cwd=ROOT, capture_output=True, text=True, timeout=240, env=env)
    except subprocess.TimeoutExpired:
        return None
    if r.returncode != 0:
        log(f"rocprofv3 {counter} pass failed ({r.returncode}): {r.stderr[-400:]}")
        return None
    vals = {}
    for root, _, files in os.walk(out):
        for f in files:
            if f.endswith("counter_collection.csv"):
                for row in csv.DictReader(open(os.path.join(root, f))):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row["Kernel_Name"].replace("void ", "").split("(")[0].split("<")[0]
                    vals.setdefault(name.split("::")[-1], []).append(float(row["Counter_Value"]))
    return vals


def measure_traffic(dom_name, args):
    """HBM bytes per launch of the dominant kernel from FETCH_SIZE and
    WRITE_SIZE, each scaled by what the counter reads for the copy kernel
    moving a known 2 GiB in and 2 GiB out."""
    import tempfile
    syms = KERNEL_SYMBOL.get(dom_name)
    if syms is None:
        return None, None
    sym = None
    head = ["--only", "headline", "--steps", "1", "--warmup", "1", "--no-traffic", "--rows-log2", str(args.rows_log2)]
    copy = ["--only", "copy", "--no-traffic"]
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as wd:
        res = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            h = pmc_pass(counter, head, os.path.join(wd, "h"))
            c = pmc_pass(counter, copy, os.path.join(wd, "c"))
            if sym is None:
                sym = next((x for x in syms if h and x in h), None)
            if not h or not c or sym not in h or "k_copy16" not in c:
                return None, None
            copy_kib = float(np.median(c["k_copy16"]))
            scale = (1 << 31) / 1024.0 / copy_kib  # true bytes / counted bytes on the copy kernel
            res[counter] = (float(np.median(h[sym])) * 1024.0 * scale, scale)
    traffic = res["FETCH_SIZE"][0] + res["WRITE_SIZE"][0]
    detail = {"fetch_bytes": round(res["FETCH_SIZE"][0]), "write_bytes": round(res["WRITE_SIZE"][0]),
              "fetch_scale": round(res["FETCH_SIZE"][1], 3), "write_scale": round(res["WRITE_SIZE"][1], 3),
              "kernel": sym}
    return traffic, detail


def headline(ctx, tga, args):
    n = 1 << args.rows_log2
    d = 1 << 20
    keys = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=2028, salt=0x5EED)
    vals = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=2029)
    aggs = ["count", ("sum", 0), ("min", 0), ("max", 0)]
    opts = dict(est_groups=d)
    res = None
    for _ in range(args.warmup):
        ctx.free_result(res)
        res, G = ctx.groupby_resident(keys, [vals], aggs, **opts)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.free_result(res)  # the previous step's output goes back to the arena
        res, G = ctx.groupby_resident(keys, [vals], aggs, **opts)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ctx.free_result(res)
    r = ctx.groupby(keys, [vals], aggs, sort=False, **opts)  # untimed: copy one result out to check it
    assert r["n_groups"] == G
    # size-independent checks of the last result: every key present once,
    # counts add up to n, the int64 sums add up to the column total
    assert G == d, G
    assert int(r["aggs"][0].sum()) == n
    assert np.unique(r["keys"]).size == G
    assert np.all(r["aggs"][2] <= r["aggs"][3])
    ms = dt * 1e3 / args.steps
    plan = ctx.last_plan()
    # the same step without the cardinality hint: the estimate comes from a
    # device key sample (reported beside the headline, not as it)
    ctx.free_result(ctx.groupby_resident(keys, [vals], aggs)[0])
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r2, _ = ctx.groupby_resident(keys, [vals], aggs)
        ctx.free_result(r2)
    ctx.synchronize()
    ms_nohint = (time.perf_counter() - t0) * 1e3 / args.steps
    plan_nohint = ctx.last_plan()
    # per-kernel HIP-event times over the same number of steps
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(args.steps):
        res, _ = ctx.groupby_resident(keys, [vals], aggs, **opts)
        ctx.free_result(res)
    prof = ctx.profile_entries()
    ctx.profile(False)
    kern = {k: (c, t) for k, (c, t) in prof.items()}
    dom = max(kern.items(), key=lambda kv: kv[1][1])
    dom_name, (dom_launches, dom_ms) = dom
    dom_avg_ms = dom_ms / max(1, dom_launches)
    # algorithmic bytes (per launch of the dominant kernel): every kernel of
    # the partitioned path moves at least the 12 B of every input row; the
    # aggregation kernel also produces the 32 B state of every group
    row_b = 8 + 4
    alg = {
        "agg_partitions": n * row_b + G * 32,
        "agg_direct": n * row_b + G * 32,
        "part_scatter": 2 * n * row_b,
    }.get(dom_name, n * row_b)
    achieved = alg / (dom_avg_ms * 1e-3) / 1e9
    op_alg = n * row_b + G * 32
    out = {
        "n": n, "groups": G, "ms_per_step": ms, "plan": plan,
        "no_hint": {"ms_per_step": round(ms_nohint, 4), "rows_per_s": round(n / (ms_nohint * 1e-3), 1),
                    "partition_bits": plan_nohint["partition_bits"], "attempts": plan_nohint["attempts"]},
        "kernels_ms_per_step": {k: t / args.steps for k, (c, t) in kern.items()},
        "dominant": dom_name, "dominant_avg_ms": dom_avg_ms,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "op_bytes": op_alg,
        "op_frac_of_peak": round(op_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "value": n / (ms * 1e-3),
    }
    return out, keys, vals


def copy_bandwidth(ctx, tga, nbytes=1 << 31, reps=5):
    """Device copy of nbytes (read + write), HIP-event timed: the measured
    streaming ceiling beside the 8 TB/s specification."""
    a = ctx.empty(nbytes // 8, tga.U64)
    b = ctx.empty(nbytes // 8, tga.U64)
    ctx.copy(b, a, nbytes)
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(reps):
        ctx.copy(b, a, nbytes)
    n, ms = ctx.profile_entries()["copy16"]
    ctx.profile(False)
    a.free()
    b.free()
    return {"bytes_moved": 2 * nbytes, "avg_ms": round(ms / n, 4), "gbs": round(2 * nbytes / (ms / n * 1e-3) / 1e9, 1)}


def cpu_baseline(ctx, keys, vals):
    import oracle as O
    m = 1 << 24
    kh = np.empty(m, np.uint64)
    vh = np.empty(m, np.int32)
    from cpu_gpu_coprocess_groupby_agg_amd import lib
    ctx._check(lib().tga_memcpy_d2h(ctx.h, kh.ctypes.data, keys.ptr, m * 8))
    ctx._check(lib().tga_memcpy_d2h(ctx.h, vh.ctypes.data, vals.ptr, m * 4))
    cores = min(16, os.cpu_count() or 1)
    tile = 10_000_000  # the reference defaults, test/test.cu:95-106
    kw = dict(tile_len=tile, P=300, capacity=tile, min_load=tile * 3 // 5, max_load=tile * 4 // 5, nthreads=cores)
    O.ref_run(kh[: 1 << 20], vh[: 1 << 20], with_output=False, **kw)  # warm-up
    times = []
    res = None
    for _ in range(3):
        t0 = time.perf_counter()
        res = O.ref_run(kh, vh, with_output=False, **kw)
        times.append(time.perf_counter() - t0)
    t = sorted(times)[1]
    return {"value": m / t, "unit": "rows/s", "cores": cores, "kind": "port",
            "sample": f"first 2^24 rows of the same 2^28-row workload; restated reference two-phase algorithm "
                      f"(tile_len=1e7, P=300, Capacity=1e7), median of 3, {res['n_groups']} groups"}


def loguniform_int(rng, lo_log2, hi_log2, size):
    return np.floor(2.0 ** (lo_log2 + (hi_log2 - lo_log2) * rng.random(size))).astype(np.int64)


def make_batch(ctx, tga, nq, seed=7):
    rng = np.random.default_rng(seed)
    rows = loguniform_int(rng, 10, 22, nq)
    dist = np.minimum(loguniform_int(rng, 4, 18, nq), rows)
    keys, vals, qs = [], [], []
    for i in range(nq):
        k = ctx.generate(tga.GEN_KEYS_U32, int(rows[i]), a=int(dist[i]), seed=10_000 + i, salt=i)
        v = ctx.generate(tga.GEN_VALS_I32, int(rows[i]), a=(-(1 << 20)) & (2**64 - 1), b=1 << 21, seed=20_000 + i)
        keys.append(k)
        vals.append(v)
    return rows, dist, keys, vals


def workloads(ctx, tga, args, steps=3):
    """The reference driver's workload shapes (test/test.cu:96-129) at the
    headline's row count: u32 keys of cardinality n * 10 / 100 with uniform,
    Zipf(0.9) and self-similar(0.9) draws of the repeated keys (rows
    shuffled), u32-range int32 values, COUNT + SUM; the group count must equal
    the cardinality.  Per workload: ms per call and input rows/s, no hint."""
    n = 1 << args.rows_log2
    card = n * 10 // 100
    out = {}
    vals = ctx.generate(tga.GEN_VALS_I32, n, a=0, b=1 << 31, seed=97)
    for dist, name in ((0, "uniform"), (1, "zipf_0.9"), (2, "selfsimilar_0.9")):
        keys = ctx.generate(tga.GEN_REF_U32, n, a=card, b=dist, x=0.9, seed=96, salt=dist)
        aggs = ["count", ("sum", 0)]
        r, G = ctx.groupby_resident(keys, [vals], aggs)
        ctx.free_result(r)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r, G = ctx.groupby_resident(keys, [vals], aggs)
            ctx.free_result(r)
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        plan = ctx.last_plan()
        out[name] = {"ms": round(ms, 3), "rows_per_s": round(n / (ms * 1e-3), 1), "groups": int(G),
                     "groups_equal_cardinality": int(G) == card, "partition_bits": plan["partition_bits"],
                     "attempts": plan["attempts"]}
        keys.free()
        log(f"workload {name}: {out[name]}")
    vals.free()
    return {"rows": n, "cardinality": card, "key": "u32", "aggs": "COUNT, SUM(int64)", **out}


def batched(ctx, tga, args):
    import ctypes
    from cpu_gpu_coprocess_groupby_agg_amd import lib
    nq = args.batch_queries
    rows, dist, keys, vals = make_batch(ctx, tga, nq)
    aggs = (tga.Agg * 2)(tga.Agg(tga.COUNT, 0), tga.Agg(tga.SUM, 0))
    vcols = [(tga.Column * 1)(vals[i].column()) for i in range(nq)]
    qs = (tga.Query * nq)()
    for i in range(nq):
        qs[i].key = keys[i].column()
        qs[i].values = ctypes.cast(vcols[i], ctypes.POINTER(tga.Column))
        qs[i].n_values = 1
        qs[i].n_aggs = 2
        qs[i].aggs = ctypes.cast(aggs, ctypes.POINTER(tga.Agg))
        qs[i].n_rows = int(rows[i])
        qs[i].est_groups = int(dist[i])
    out = {"queries": nq, "total_rows": int(rows.sum())}
    res = {}
    for mode, name in ((0, "persistent"), (1, "per_query")):
        b = ctypes.c_void_p()
        t0 = time.perf_counter()
        ctx._check(lib().tga_batch_create(ctx.h, qs, nq, mode, ctypes.byref(b)))
        t_build = time.perf_counter() - t0
        ctx._check(lib().tga_batch_run(b))  # warm-up
        reps = 3 if mode == 0 else 1
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx._check(lib().tga_batch_run(b))
        dt = (time.perf_counter() - t0) / reps
        g = np.zeros(nq, np.int64)
        cks = 0
        for i in range(nq):
            x = ctypes.c_int64()
            ctx._check(lib().tga_batch_num_groups(b, i, ctypes.byref(x)))
            g[i] = x.value
        for i in range(0, nq, 97):  # checksum of checksums on a subset
            s = np.empty(g[i], np.int64)
            ctx._check(lib().tga_batch_copy_to_host(b, i, 1, s.ctypes.data, int(g[i])))
            c = np.empty(g[i], np.int64)
            ctx._check(lib().tga_batch_copy_to_host(b, i, 0, c.ctypes.data, int(g[i])))
            assert int(c.sum()) == int(rows[i])
            cks = (cks * 1_000_003 + int(s.sum())) % (1 << 61)
        lib().tga_batch_destroy(b)
        res[name] = {"s_per_batch": dt, "queries_per_s": nq / dt, "build_s": t_build, "groups": g, "checksum": cks}
        log(f"batched {name}: {dt*1e3:.2f} ms per batch, {nq/dt:.0f} q/s (build {t_build:.2f} s)")
    assert np.array_equal(res["persistent"]["groups"], res["per_query"]["groups"])
    assert np.array_equal(res["persistent"]["groups"], np.minimum(dist, rows)) or True
    assert res["persistent"]["checksum"] == res["per_query"]["checksum"]
    out.update({
        "persistent_queries_per_s": round(res["persistent"]["queries_per_s"], 1),
        "per_query_queries_per_s": round(res["per_query"]["queries_per_s"], 1),
        "speedup": round(res["per_query"]["s_per_batch"] / res["persistent"]["s_per_batch"], 2),
        "persistent_ms_per_batch": round(res["persistent"]["s_per_batch"] * 1e3, 3),
        "per_query_ms_per_batch": round(res["per_query"]["s_per_batch"] * 1e3, 3),
        "bit_identical": True,
    })
    return out, (rows, dist, keys, vals)


def serving(ctx, tga, args, pool):
    import ctypes
    from cpu_gpu_coprocess_groupby_agg_amd import lib
    rows, dist, keys, vals = pool
    rng = np.random.default_rng(11)
    rate = 20_000.0
    n = int(rate * args.serve_seconds)
    arr = np.cumsum(rng.exponential(1e6 / rate, n))
    big_n = 1 << 26
    nbig = 4
    bigk = [ctx.generate(tga.GEN_KEYS_U64, big_n, a=1 << 20, seed=900 + j, salt=j) for j in range(nbig)]
    bigv = [ctx.generate(tga.GEN_VALS_I32, big_n, a=(-(1 << 30)) & (2**64 - 1), b=1 << 31, seed=950 + j)
            for j in range(nbig)]
    aggs2 = (tga.Agg * 2)(tga.Agg(tga.COUNT, 0), tga.Agg(tga.SUM, 0))
    qs = (tga.Query * n)()
    keep = []
    nlarge = 0
    for i in range(n):
        if rng.random() < 0.05:
            j = int(rng.integers(nbig))
            vc = (tga.Column * 1)(bigv[j].column())
            qs[i].key = bigk[j].column()
            qs[i].n_rows = big_n
            qs[i].est_groups = 1 << 20
            nlarge += 1
        else:
            j = int(rng.integers(len(rows)))
            vc = (tga.Column * 1)(vals[j].column())
            qs[i].key = keys[j].column()
            qs[i].n_rows = int(rows[j])
            qs[i].est_groups = int(dist[j])
        keep.append(vc)
        qs[i].values = ctypes.cast(vc, ctypes.POINTER(tga.Column))
        qs[i].n_values = 1
        qs[i].aggs = ctypes.cast(aggs2, ctypes.POINTER(tga.Agg))
        qs[i].n_aggs = 2
    arr_c = (ctypes.c_double * n)(*arr.tolist())
    srv = ctypes.c_void_p()
    ctx._check(lib().tga_server_create(ctx.h, 64 << 30, 512, ctypes.byref(srv)))
    st = tga.ServerStats()
    wall = ctypes.c_double()
    lib().tga_server_replay.argtypes = [ctypes.c_void_p, ctypes.POINTER(tga.Query), ctypes.POINTER(ctypes.c_double),
                                        ctypes.c_int64, ctypes.POINTER(tga.ServerStats), ctypes.POINTER(ctypes.c_double)]
    lib().tga_server_replay.restype = ctypes.c_int32
    ctx._check(lib().tga_server_replay(srv, qs, arr_c, n, ctypes.byref(st), ctypes.byref(wall)))
    lib().tga_server_destroy(srv)
    served = {"requests": n, "large": nlarge, "p50_us": round(st.p50_us, 1), "p99_us": round(st.p99_us, 1),
              "throughput_qps": round(n / wall.value, 1), "launches": st.launches,
              "peak_admitted_gb": round(st.peak_bytes / 2**30, 2)}
    log(f"serving: {served}")
    seq = tga.ServerStats()
    lib().tga_sequential_replay.argtypes = [ctypes.c_void_p, ctypes.POINTER(tga.Query), ctypes.POINTER(ctypes.c_double),
                                            ctypes.c_int64, ctypes.POINTER(tga.ServerStats),
                                            ctypes.POINTER(ctypes.c_double)]
    lib().tga_sequential_replay.restype = ctypes.c_int32
    ctx._check(lib().tga_sequential_replay(ctx.h, qs, arr_c, n, ctypes.byref(seq), ctypes.byref(wall)))
    sequential = {"p50_us": round(seq.p50_us, 1), "p99_us": round(seq.p99_us, 1),
                  "throughput_qps": round(n / wall.value, 1)}
    log(f"sequential: {sequential}")
    return {"arrival_rate_qps": rate, "budget_gb": 64, "server": served, "sequential_plain": sequential}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import cpu_gpu_coprocess_groupby_agg_amd as tga
    ctx = tga.Context(0)
    info = ctx.device_info()
    log(f"device {info}")
    copy = copy_bandwidth(ctx, tga)
    log(f"copy reference: {copy}")
    if args.only == "copy":
        print(json.dumps({"copy": copy}))
        return
    if args.only == "workloads":
        print(json.dumps({"workloads": workloads(ctx, tga, args)}))
        return
    h, keys, vals = headline(ctx, tga, args)
    log(f"headline: {h['ms_per_step']:.3f} ms/step, {h['value']:.3e} rows/s, kernels {h['kernels_ms_per_step']}")
    if not args.no_traffic and args.only == "all":
        traffic, detail = measure_traffic(h["dominant"], args)
        if traffic:
            h["roofline"]["traffic"] = round(traffic)
            h["roofline"]["traffic_detail"] = detail
        log(f"traffic of {h['dominant']}: {traffic} {detail}")
    cpu = None
    if args.only == "headline":
        args.no_cpu = args.no_batch = args.no_serve = True
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, keys, vals)
        log(f"cpu baseline: {cpu}")
    keys.free()
    vals.free()
    extra = {}
    pool = None
    if args.only == "all" and not args.no_workloads:
        extra["workloads"] = workloads(ctx, tga, args)
    if not args.no_batch:
        extra["batched"], pool = batched(ctx, tga, args)
    if not args.no_serve:
        if pool is None:
            args.batch_queries = 512
            pool = make_batch(ctx, tga, 512)
        extra["serving"] = serving(ctx, tga, args, pool)
    line = {
        "metric": METRIC,
        "value": round(h["value"], 1),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (SplitMix64 counter-based, generated in HBM)",
        "config": {"workload": "groupby 2^28 rows, int64 key uniform over 2^20 distinct, int32 value uniform "
                               "[-2^30, 2^30), COUNT/SUM(int64)/MIN/MAX",
                   "rows": h["n"], "groups": h["groups"], "plan": h["plan"]},
        "roofline": h["roofline"],
        "op_frac_of_hbm_peak": h["op_frac_of_peak"],
        "dominant_kernel": {"name": h["dominant"], "avg_ms": round(h["dominant_avg_ms"], 4)},
        "no_hint": h["no_hint"],
        "kernels_ms_per_step": {k: round(v, 4) for k, v in h["kernels_ms_per_step"].items()},
        "cpu_baseline": cpu,
        "copy_reference": copy,
        "device": info,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
