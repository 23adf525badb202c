#!/usr/bin/env python3
"""Benchmark of the MI355X tiled hash group-by (one GPU).

Headline (one JSON line on stdout): input rows/s of the plain operator on the
2^28-row group-by -- int64 keys uniform over 2^20 distinct values, int32
values uniform in [-2^30, 2^30), COUNT / SUM(int64) / MIN / MAX -- with the
columns resident in HBM when the timed region starts and NO cardinality hint
(the autotuner estimates the group count itself).  A step is one complete
group-by: estimate, plan, every kernel, the group count read back.  Beside it:
  hinted        the same step with the exact group count as a hint
  roofline      dominant kernel: algorithmic bytes per launch (12 B per input
                row it processes, + 32 B per group for the aggregation) / its
                HIP-event time; traffic = measured HBM bytes of that kernel
  cpu_baseline  the restated reference algorithm (tile pre-aggregation, hash
                partitioning, merge_partition, per-partition aggregation) on
                host threads, over the full 2^28-row input of the headline
  config1       2^24 rows, u32 keys over 2^10, int32 values: GPU vs the
                restated reference on the CPU, bit-exact, both timed
  config3       2^29 rows, Zipf(1.1) int64 keys over 2^24, int32 + fp64
                columns, COUNT + SUM(int64) + SUM(fp64): timed, roofline
  batched       4096 queries (rows log-uniform in [2^10, 2^22], distinct keys
                log-uniform in [2^4, 2^18], u32 keys, int32 values,
                COUNT+SUM): one persistent launch vs one launch per query,
                every query's results compared between the two
  serving       Poisson arrivals (mean 20k/s) from that query mix plus 5%
                2^26-row int64-key queries, 64 GB admission budget: p50/p99
                and throughput vs the sequential plain operator, every
                request's group count checked; a sweep of arrival rates
                names the highest one served with p99 <= 50 ms
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "input rows/s + % HBM peak at 2^28 rows; batched queries/s & p99 vs per-query"
M64 = 2**64 - 1


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
    ap.add_argument("--no-configs", action="store_true", help="skip the config-1 and config-3 sections")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the reference-workload section (its driver's uniform / Zipf / self-similar keys)")
    ap.add_argument("--batch-queries", type=int, default=4096)
    ap.add_argument("--batch-modes", choices=["both", "persistent"], default="both",
                    help="persistent: skip the one-launch-per-query comparison (for counter runs)")
    ap.add_argument("--batch-profile", action="store_true",
                    help="also report the persistent kernel's per-item-type clock profile")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-rate", type=float, default=20_000.0)
    ap.add_argument("--serve-sweep", type=lambda x: [float(v) for v in x.split(",")],
                    default=[5_000.0, 10_000.0, 15_000.0, 17_500.0, 20_000.0])
    ap.add_argument("--only", choices=["all", "headline", "copy", "workloads", "config1", "config3", "batch", "serve",
                                       "stream", "coprocess"],
                    default="all",
                    help="run one section (profiling runs); headline: only the 2^28-row group-by; copy: only the "
                         "copy reference")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes that fill roofline.traffic")
    return ap.parse_args()


# profiler name of a timed step -> kernel symbols it may launch
KERNEL_SYMBOL = {"part_scatter": ("k_part_scatter_wc", "k_part_scatter"), "part_scatter2": ("k_part_scatter_wc",),
                 "agg_partitions": ("k_agg",), "agg_direct": ("k_agg",), "compact": ("k_compact",),
                 "copy16": ("k_copy16",), "slice_distinct": ("k_slice_distinct",)}


def alg_bytes(kernel, n, G, row_b, state_b, key_b):
    """Algorithmic bytes of one launch: the input bytes of the rows it
    processes (the scatter's write of the partitioned copy is traffic, not
    algorithmic bytes), plus the group state the aggregation produces."""
    if kernel in ("agg_partitions", "agg_direct"):
        return n * row_b + G * state_b
    if kernel == "slice_distinct":
        return n * key_b
    return n * row_b


# ---------------------------------------------------------------------------
# measurement helpers
# ---------------------------------------------------------------------------
def time_steps(ctx, keys, cols, aggs, steps, warmup, **opts):
    """Mean ms per complete group-by call (result left in HBM and released)."""
    G = None
    for _ in range(warmup):
        r, G = ctx.groupby_resident(keys, cols, aggs, **opts)
        ctx.free_result(r)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r, G = ctx.groupby_resident(keys, cols, aggs, **opts)
        ctx.free_result(r)
    ctx.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, G, ctx.last_plan()


def kernel_times(ctx, keys, cols, aggs, steps, **opts):
    """{step name: (launches, total ms)} from HIP events on the context stream."""
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        r, _ = ctx.groupby_resident(keys, cols, aggs, **opts)
        ctx.free_result(r)
    prof = ctx.profile_entries()
    ctx.profile(False)
    return prof


def roofline(prof, steps, n, G, row_b, state_b, key_b):
    dom_name, (dom_launches, dom_ms) = max(prof.items(), key=lambda kv: kv[1][1])
    avg = dom_ms / max(1, dom_launches)
    alg = alg_bytes(dom_name, n, G, row_b, state_b, key_b)
    achieved = alg / (avg * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": dom_name, "avg_ms": round(avg, 4), "algorithmic_bytes": int(alg)}, \
        {k: round(t / steps, 4) for k, (c, t) in prof.items()}


def pmc_pass(counter, argv, workdir):
    """One rocprofv3 counter pass over a child `python3 bench.py ...`; returns
    {kernel symbol prefix: [per-dispatch values in KiB]} or None."""
    import csv
    import shutil
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


def host_info():
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), "")
    except Exception:
        pass
    return {"host_cores": os.cpu_count(), "cpu_model": model}


def cpu_threads():
    # the host share of one GPU on the box is 16 cores; never more than it has
    return min(16, os.cpu_count() or 1)


def ref_kw(n):
    tile = 10_000_000  # the reference defaults, test/test.cu:95-106
    return dict(tile_len=tile, P=300, capacity=tile, min_load=tile * 3 // 5, max_load=tile * 4 // 5)


# ---------------------------------------------------------------------------
# sections
# ---------------------------------------------------------------------------
def headline(ctx, tga, args):
    n = 1 << args.rows_log2
    d = 1 << 20
    keys = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=2028, salt=0x5EED)
    vals = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & M64, b=1 << 31, seed=2029)
    aggs = ["count", ("sum", 0), ("min", 0), ("max", 0)]
    ms, G, plan = time_steps(ctx, keys, [vals], aggs, args.steps, args.warmup)
    ms_hint, G2, plan_hint = time_steps(ctx, keys, [vals], aggs, args.steps, 1, est_groups=d)
    assert G == G2 == d, (G, G2)
    # untimed: copy one result out and check it through size-independent
    # properties (every key once, COUNT adds up to n, MIN <= MAX)
    r = ctx.groupby(keys, [vals], aggs, sort=False)
    assert r["n_groups"] == d
    assert int(r["aggs"][0].sum()) == n
    assert np.unique(r["keys"]).size == d
    assert np.all(r["aggs"][2] <= r["aggs"][3])
    prof = kernel_times(ctx, keys, [vals], aggs, args.steps)
    roof, kern = roofline(prof, args.steps, n, G, 12, 32, 8)
    op_alg = n * 12 + G * 32
    out = {
        "n": n, "groups": G, "ms_per_step": ms, "plan": plan, "value": n / (ms * 1e-3),
        "hinted": {"ms_per_step": round(ms_hint, 4), "rows_per_s": round(n / (ms_hint * 1e-3), 1),
                   "partition_bits": plan_hint["partition_bits"], "attempts": plan_hint["attempts"]},
        "kernels_ms_per_step": kern, "roofline": roof,
        "op_bytes": op_alg, "op_frac_of_peak": round(op_alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
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


def cpu_baseline(ctx, tga, keys, vals):
    """The restated reference algorithm on the full headline input (2^28 rows,
    27 tiles of 1e7), one run after a small warm-up."""
    import oracle as O
    kh = keys.to_numpy()
    vh = vals.to_numpy()
    n = kh.size
    th = cpu_threads()
    O.ref_run(kh[: 1 << 20], vh[: 1 << 20], with_output=False, nthreads=th, **ref_kw(1 << 20))  # warm-up
    t0 = time.perf_counter()
    res = O.ref_run(kh, vh, with_output=False, nthreads=th, **ref_kw(n))
    t = time.perf_counter() - t0
    assert res["n_groups"] == 1 << 20
    st = res["stats"]
    out = {"value": round(n / t, 1), "unit": "rows/s", "cores": th, "kind": "port",
           "sample": f"the full 2^28-row headline input ({st['tiles']} tiles of 1e7 rows, P=300, Capacity=1e7): the "
                     f"restated reference two-phase algorithm on {th} host threads, one timed run after a 2^20-row "
                     f"warm-up; {res['n_groups']} groups",
           "seconds": round(t, 3),
           "phases_s": {"phase1": round(st["t_phase1_s"], 3), "merge": round(st["t_merge_s"], 6),
                        "phase2": round(st["t_phase2_s"], 3)}}
    out.update(host_info())
    return out


def config1(ctx, tga, args):
    """2^24 rows, u32 keys uniform over 2^10, int32 values in [-2^20, 2^20),
    COUNT + SUM(int64): GPU timed, and checked bit-exact against the restated
    reference algorithm run on the CPU over the same rows (its u32 types: the
    low 32 bits of SUM, the counts, the keys), which is timed as well."""
    import oracle as O
    n, d = 1 << 24, 1 << 10
    k = ctx.generate(tga.GEN_KEYS_U32, n, a=d, seed=1, salt=3)
    v = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 20)) & M64, b=1 << 21, seed=2)
    aggs = ["count", ("sum", 0)]
    steps = max(args.steps, 20)
    ms, G, plan = time_steps(ctx, k, [v], aggs, steps, args.warmup)
    prof = kernel_times(ctx, k, [v], aggs, steps)
    roof, kern = roofline(prof, steps, n, G, 8, 20, 4)
    got = ctx.groupby(k, [v], aggs)
    kh, vh = k.to_numpy(), v.to_numpy()
    th = cpu_threads()
    O.ref_run(kh[:1 << 18], vh[:1 << 18].view(np.uint32), nthreads=th, **ref_kw(1 << 18))  # warm-up
    t0 = time.perf_counter()
    ref = O.ref_run(kh, vh.view(np.uint32), nthreads=th, **ref_kw(n))
    t_cpu = time.perf_counter() - t0
    o = np.argsort(ref["keys"], kind="stable")
    exact = O.groupby(kh, [vh], aggs)
    bit_exact = (got["n_groups"] == ref["n_groups"] == d
                 and np.array_equal(got["keys"], ref["keys"][o])
                 and np.array_equal(got["aggs"][0], ref["count"][o].astype(np.int64))
                 and np.array_equal(got["aggs"][1].astype(np.uint64) & np.uint64(0xFFFFFFFF),
                                    ref["sum"][o].astype(np.uint64) & np.uint64(0xFFFFFFFF))
                 and np.array_equal(got["aggs"][1], exact["aggs"][1]))
    k.free()
    v.free()
    return {"workload": "2^24 rows, u32 key uniform over 2^10, int32 value uniform [-2^20, 2^20), COUNT + SUM(int64)",
            "gpu_ms": round(ms, 4), "gpu_rows_per_s": round(n / (ms * 1e-3), 1),
            "gpu_frac_of_hbm_peak": round((n * 8 + G * 20) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "roofline": roof, "kernels_ms_per_step": kern, "path": plan["path"], "plan": plan,
            "cpu_rows_per_s": round(n / t_cpu, 1), "cpu_threads": th, "cpu_kind": "restated reference (u32 types)",
            "bit_exact_vs_cpu": bool(bit_exact), "groups": int(G)}


def config3(ctx, tga, args):
    """2^29 rows, int64 keys Zipf(1.1) over 2^24 distinct values, an int32
    column uniform in [-2^30, 2^30) and an fp64 column N(0, 1e3), COUNT +
    SUM(int64) + SUM(fp64), no hint.  Checks: COUNT adds up to n, the int64
    sums add up to the column total, the group count within (2^20, 2^24]."""
    n, D = 1 << 29, 1 << 24
    k = ctx.generate(tga.GEN_ZIPF_U64, n, a=D, x=1.1, seed=2930, salt=11)
    a = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & M64, b=1 << 31, seed=2931)
    f = ctx.generate(tga.GEN_NORMAL_F64, n, x=0.0, y=1e3, seed=2932)
    aggs = ["count", ("sum", 0), ("sum", 1)]
    steps = max(3, args.steps // 3)
    ms, G, plan = time_steps(ctx, k, [a, f], aggs, steps, 1)
    prof = kernel_times(ctx, k, [a, f], aggs, steps)
    roof, kern = roofline(prof, steps, n, G, 20, 32, 8)
    got = ctx.groupby(k, [a, f], aggs, sort=False)
    ah = a.to_numpy()
    ok = (int(got["aggs"][0].sum()) == n and int(got["aggs"][1].sum()) == int(ah.astype(np.int64).sum())
          and (1 << 20) < got["n_groups"] <= D)
    cpu = None
    if not args.no_cpu:
        import oracle as O
        m = 1 << 26
        kh = k.to_numpy()[:m]
        th = cpu_threads()
        t0 = time.perf_counter()
        r = O.ref_run(kh, ah[:m], with_output=False, nthreads=th, **ref_kw(m))
        t = time.perf_counter() - t0
        cpu = {"rows_per_s": round(m / t, 1), "threads": th,
               "sample": f"first 2^26 rows, key + int32 column (the restated reference carries one value column), "
                         f"{r['n_groups']} groups"}
    for x in (k, a, f):
        x.free()
    return {"workload": "2^29 rows, int64 key Zipf(1.1) over 2^24, int32 uniform [-2^30, 2^30) + fp64 N(0, 1e3), "
                        "COUNT + SUM(int64) + SUM(fp64), no hint",
            "gpu_ms": round(ms, 4), "gpu_rows_per_s": round(n / (ms * 1e-3), 1),
            "gpu_frac_of_hbm_peak": round((n * 20 + G * 32) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "roofline": roof, "kernels_ms_per_step": kern, "groups": int(G),
            "plan": {k2: plan[k2] for k2 in ("path", "partition_bits", "attempts", "est_groups")},
            "checks_ok": bool(ok), "cpu": cpu}


def host_streaming(ctx, tga, args, reps=3):
    """The headline workload from pinned host memory: the operator streams
    the columns through HBM tile by tile, copies overlapped with the
    aggregation (PCIe-inclusive rows/s, not the headline).  Beside it the
    copy of the same bytes alone (the PCIe floor) and the operator on the
    device-resident columns (the compute floor)."""
    from cpu_gpu_coprocess_groupby_agg_amd import lib
    n = 1 << args.rows_log2
    d = 1 << 20
    keys = ctx.generate(tga.GEN_KEYS_U64, n, a=d, seed=2028, salt=0x5EED)
    vals = ctx.generate(tga.GEN_VALS_I32, n, a=(-(1 << 30)) & M64, b=1 << 31, seed=2029)
    kh = ctx.pinned_empty(n, np.uint64)
    vh = ctx.pinned_empty(n, np.int32)
    ctx._check(lib().tga_memcpy_d2h(ctx.h, kh.ctypes.data, keys.ptr, n * 8))
    ctx._check(lib().tga_memcpy_d2h(ctx.h, vh.ctypes.data, vals.ptr, n * 4))
    aggs = ["count", ("sum", 0), ("min", 0), ("max", 0)]
    al = [(tga._OPS[a] if isinstance(a, str) else tga._OPS[a[0]], 0 if isinstance(a, str) else a[1]) for a in aggs]
    ag = (tga.Agg * 4)(*[tga.Agg(o, c) for o, c in al])
    kc = tga.Column(kh.ctypes.data, tga.U64, 1)
    vc = (tga.Column * 1)(tga.Column(vh.ctypes.data, tga.I32, 1))

    def call():
        r = ctypes.c_void_p()
        ctx._check(lib().tga_groupby(ctx.h, ctypes.byref(kc), vc, 1, n, ag, 4, None, ctypes.byref(r)))
        g = ctypes.c_int64()
        ctx._check(lib().tga_result_num_groups(r.value, ctypes.byref(g)))
        lib().tga_result_free(r.value)
        return g.value

    G = call()
    plan = ctx.last_plan()
    t0 = time.perf_counter()
    for _ in range(reps):
        G = call()
    t = (time.perf_counter() - t0) / reps
    assert G == d
    # the copies alone, same bytes, pinned -> HBM
    t0 = time.perf_counter()
    ctx._check(lib().tga_memcpy_h2d(ctx.h, keys.ptr, kh.ctypes.data, n * 8))
    ctx._check(lib().tga_memcpy_h2d(ctx.h, vals.ptr, vh.ctypes.data, n * 4))
    t_copy = time.perf_counter() - t0
    ms_dev, _, _ = time_steps(ctx, keys, [vals], aggs, 3, 1)
    keys.free()
    vals.free()
    return {"workload": "the headline workload in pinned host memory (2^28 rows, 3.2 GB), no hint",
            "ms": round(t * 1e3, 2), "rows_per_s_pcie_inclusive": round(n / t, 1),
            "host_to_device_gbs": round(n * 12 / t / 1e9, 1),
            "copy_alone_ms": round(t_copy * 1e3, 2), "device_resident_ms": round(ms_dev, 3),
            "overlap_efficiency": round((t_copy + ms_dev * 1e-3) / t, 3),
            "tiles": plan["streamed_tiles"], "pinned": plan["stream_pinned"], "groups": int(G)}


def coprocess(ctx, tga, args, reps=2):
    """CPU + GPU co-processing of the reference's two phases over pinned host
    buffers (its driver's set-up: u32 keys of cardinality n/10, u32 values,
    tile_len 1e7, P 300, Capacity 1e7, min/max load 6e6/8e6, 4 streams, 2^26
    rows): GPU only, the planner's split, and a forced quarter of the rows on
    16 host threads.  End-to-end time per mode (both phases), the split, and
    the result (every (key, SUM) row) compared across the modes."""
    from cpu_gpu_coprocess_groupby_agg_amd import lib
    n = 1 << 26
    card = n // 10
    dk = ctx.generate(tga.GEN_REF_U32, n, a=card, b=0, x=0.9, seed=96, salt=0)
    dv = ctx.generate(tga.GEN_VALS_I32, n, a=0, b=1 << 31, seed=97)
    k0, v0 = ctx.pinned_empty(n, np.uint32), ctx.pinned_empty(n, np.uint32)
    ctx._check(lib().tga_memcpy_d2h(ctx.h, k0.ctypes.data, dk.ptr, n * 4))
    ctx._check(lib().tga_memcpy_d2h(ctx.h, v0.ctypes.data, dv.ptr, n * 4))
    dk.free()
    dv.free()
    k, v = ctx.pinned_empty(n, np.uint32), ctx.pinned_empty(n, np.uint32)
    P, tile, cap = 300, 10_000_000, 10_000_000
    out = {"workload": "reference driver shape: 2^26 rows, u32 keys of cardinality n/10 (uniform draws), u32 values, "
                       "tile_len 1e7, P 300, Capacity 1e7, 4 streams, pinned host buffers", "rows": n,
           "cardinality": card, "host_threads": cpu_threads()}
    prints = {}
    for name, share in (("gpu_only", 0.0), ("planner", -1.0), ("cpu_quarter", 0.25)):
        ctx.set_coprocess(cpu_threads(), share)
        times = []
        for _ in range(reps):
            k[:] = k0
            v[:] = v0
            t0 = time.perf_counter()
            ps = ctypes.c_void_p()
            ctx._check(lib().tga_compat_partition(ctx.h, k.ctypes.data, v.ctypes.data, n, tile, P, 4, ctypes.byref(ps)))
            begin, num, nm = (ctypes.c_size_t * P)(), (ctypes.c_size_t * P)(), ctypes.c_size_t()
            ok, ov = ctypes.c_void_p(), ctypes.c_void_p()
            ctx._check(lib().tga_compat_intra_partition(ctx.h, ps, cap, cap * 3 // 5, cap * 4 // 5, 4, ctypes.byref(ok),
                                                        ctypes.byref(ov), begin, num, ctypes.byref(nm)))
            times.append(time.perf_counter() - t0)
            lib().tga_partset_destroy(ps)
            extent = max(begin[i] + num[i] for i in range(nm.value))
            K = np.ctypeslib.as_array(ctypes.cast(ok, ctypes.POINTER(ctypes.c_uint32)), shape=(extent,))
            V = np.ctypeslib.as_array(ctypes.cast(ov, ctypes.POINTER(ctypes.c_uint32)), shape=(extent,))
            ko = np.concatenate([K[begin[i]:begin[i] + num[i]] for i in range(nm.value)])
            vo = np.concatenate([V[begin[i]:begin[i] + num[i]] for i in range(nm.value)])
            lib().tga_host_free(ok)
            lib().tga_host_free(ov)
        s1, s2 = ctx.last_coprocess(1), ctx.last_coprocess(2)
        prints[name] = (ko.size, _fingerprint(ko, vo.astype(np.int64), vo.astype(np.int64)))
        t = min(times)
        out[name] = {"ms": round(t * 1e3, 1), "rows_per_s": round(n / t, 1), "groups": int(ko.size),
                     "phase1_tiles_cpu_gpu": [s1["cpu_units"], s1["gpu_units"]],
                     "phase2_partitions_cpu_gpu": [s2["cpu_units"], s2["gpu_units"]],
                     "cpu_row_share": round((s1["cpu_rows"] + s2["cpu_rows"]) /
                                            max(1, s1["cpu_rows"] + s1["gpu_rows"] + s2["cpu_rows"] + s2["gpu_rows"]), 4),
                     "cpu_rows_per_s_per_thread": round(max(s1["cpu_rows_per_s"], s2["cpu_rows_per_s"]), 1)}
        log(f"coprocess {name}: {out[name]}")
    ctx.set_coprocess(0, 0.0)
    out["groups_equal_cardinality"] = all(g == card for g, _ in prints.values())
    out["identical_across_modes"] = len({fp for _, fp in prints.values()}) == 1
    assert out["groups_equal_cardinality"] and out["identical_across_modes"], prints
    return out


def loguniform_int(rng, lo_log2, hi_log2, size):
    return np.floor(2.0 ** (lo_log2 + (hi_log2 - lo_log2) * rng.random(size))).astype(np.int64)


def make_batch(ctx, tga, nq, seed=7):
    rng = np.random.default_rng(seed)
    rows = loguniform_int(rng, 10, 22, nq)
    dist = np.minimum(loguniform_int(rng, 4, 18, nq), rows)
    keys, vals = [], []
    for i in range(nq):
        keys.append(ctx.generate(tga.GEN_KEYS_U32, int(rows[i]), a=int(dist[i]), seed=10_000 + i, salt=i))
        vals.append(ctx.generate(tga.GEN_VALS_I32, int(rows[i]), a=(-(1 << 20)) & M64, b=1 << 21, seed=20_000 + i))
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
        ms, G, plan = time_steps(ctx, keys, [vals], ["count", ("sum", 0)], steps, 1)
        out[name] = {"ms": round(ms, 3), "rows_per_s": round(n / (ms * 1e-3), 1), "groups": int(G),
                     "groups_equal_cardinality": int(G) == card, "partition_bits": plan["partition_bits"],
                     "attempts": plan["attempts"]}
        keys.free()
        log(f"workload {name}: {out[name]}")
    vals.free()
    return {"rows": n, "cardinality": card, "key": "u32", "aggs": "COUNT, SUM(int64)", **out}


def _mix64(z):
    """SplitMix64 finalizer on a uint64 array (wraps mod 2^64)."""
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _fingerprint(keys, cnt, sm):
    """Order-independent 64-bit hash of the multiset of (key, COUNT, SUM) rows."""
    mix64 = _mix64
    with np.errstate(over="ignore"):
        h = mix64(keys.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
                  ^ mix64(cnt.view(np.uint64) ^ mix64(sm.view(np.uint64) + np.uint64(0x632BE59BD9B4E019))))
        return int(h.sum(dtype=np.uint64))


def batch_query_result(lib, b, i, G):
    k = np.empty(G, np.uint32)
    c = np.empty(G, np.int64)
    s = np.empty(G, np.int64)
    if G:
        for idx, arr in ((-1, k), (0, c), (1, s)):
            st = lib().tga_batch_copy_to_host(b, i, idx, arr.ctypes.data, G)
            assert st == 0, st
    return k, c, s


ITEM_TYPES = ["INIT", "COUNT", "DIRECT", "SCAN", "SCATTER", "PAGG", "COMPACT"]
ITEM_PHASES = ["scatter rank", "scatter local sort", "scatter write-out", "direct table set-up", "direct rows",
               "flush or emission", "partition table set-up", "partition rows"]


def batch_item_profile(ctx, lib, qs, nq):
    """The persistent kernel's own clock profile of one run (a separate batch
    built with kernel timing on, so the timed runs carry no clock reads):
    workgroup-ms per item type working, waiting and in fences, and per phase."""
    ctx.profile(True)
    b = ctypes.c_void_p()
    try:
        ctx._check(lib().tga_batch_create(ctx.h, qs, nq, 0, ctypes.byref(b)))
        ctx._check(lib().tga_batch_run(b))
        ctx._check(lib().tga_batch_run(b))
        w, wt, fw = (ctypes.c_double * 7)(), (ctypes.c_double * 7)(), (ctypes.c_double * 7)()
        it, ph = (ctypes.c_int64 * 7)(), (ctypes.c_double * 8)()
        ctx._check(lib().tga_batch_profile(b, w, wt, it, fw, ph))
    finally:
        if b.value:
            lib().tga_batch_destroy(b)
        ctx.profile(False)
    types = {t: {"items": int(it[i]), "work_wg_ms": round(w[i], 1), "wait_wg_ms": round(wt[i], 1),
                 "fence_wg_ms": round(fw[i], 1), "us_per_item": round(w[i] / max(1, it[i]) * 1e3, 1)}
             for i, t in enumerate(ITEM_TYPES)}
    phases = {p: round(ph[i], 1) for i, p in enumerate(ITEM_PHASES)}
    log(f"batch item profile: {types} phases {phases}")
    return {"types": types, "phases_wg_ms": phases}


def batched(ctx, tga, args):
    import oracle as O
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
    check_q = list(range(0, nq, max(1, nq // 16)))
    modes = ((0, "persistent"), (1, "per_query")) if args.batch_modes == "both" else ((0, "persistent"),)
    for mode, name in modes:
        b = ctypes.c_void_p()
        t0 = time.perf_counter()
        ctx._check(lib().tga_batch_create(ctx.h, qs, nq, mode, ctypes.byref(b)))
        t_build = time.perf_counter() - t0
        ctx._check(lib().tga_batch_run(b))  # warm-up
        reps = 5 if mode == 0 else 1
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx._check(lib().tga_batch_run(b))
        dt = (time.perf_counter() - t0) / reps
        g = np.zeros(nq, np.int64)
        fp = np.zeros(nq, np.uint64)
        exact = {}
        for i in range(nq):
            x = ctypes.c_int64()
            ctx._check(lib().tga_batch_num_groups(b, i, ctypes.byref(x)))
            g[i] = x.value
            k, c, s = batch_query_result(lib, b, i, int(g[i]))
            assert int(c.sum()) == int(rows[i])
            fp[i] = _fingerprint(k, c, s)
            if i in check_q:
                o = np.argsort(k, kind="stable")
                exact[i] = (k[o], c[o], s[o])
        lib().tga_batch_destroy(b)
        res[name] = {"s_per_batch": dt, "queries_per_s": nq / dt, "build_s": t_build, "groups": g, "fp": fp,
                     "exact": exact}
        log(f"batched {name}: {dt*1e3:.2f} ms per batch, {nq/dt:.0f} q/s (build {t_build:.2f} s)")
    if args.batch_profile:
        out["item_profile"] = batch_item_profile(ctx, lib, qs, nq)
    p, q = res["persistent"], res.get("per_query")
    same_groups = q is None or bool(np.array_equal(p["groups"], q["groups"]))
    same_rows = q is None or bool(np.array_equal(p["fp"], q["fp"]))
    # a subset against the oracle, row for row
    oracle_ok = True
    for i in check_q:
        kh, vh = keys[i].to_numpy(), vals[i].to_numpy()
        ref = O.groupby(kh, [vh], ["count", ("sum", 0)])
        for mode_res in (p, q) if q is not None else (p,):
            k, c, s = mode_res["exact"][i]
            oracle_ok &= bool(np.array_equal(k, ref["keys"]) and np.array_equal(c, ref["aggs"][0])
                              and np.array_equal(s, ref["aggs"][1]))
    assert same_groups and same_rows and oracle_ok, (same_groups, same_rows, oracle_ok)
    total_groups = int(p["groups"].sum())
    alg = int(rows.sum()) * 8 + total_groups * 20
    if q is not None:
        out.update({
            "per_query_queries_per_s": round(q["queries_per_s"], 1),
            "speedup": round(q["s_per_batch"] / p["s_per_batch"], 2),
            "per_query_ms_per_batch": round(q["s_per_batch"] * 1e3, 3),
        })
    out.update({
        "persistent_queries_per_s": round(p["queries_per_s"], 1),
        "persistent_ms_per_batch": round(p["s_per_batch"] * 1e3, 3),
        "persistent_build_s": round(p["build_s"], 3),
        "total_groups": total_groups,
        "algorithmic_bytes": alg,
        "persistent_gbs": round(alg / p["s_per_batch"] / 1e9, 1),
        "persistent_frac_of_hbm_peak": round(alg / p["s_per_batch"] / 1e9 / HBM_PEAK_GBS, 4),
        "bit_identical": (same_groups and same_rows) if q is not None else None,
        "compared": (f"all {nq} queries: group counts and a 64-bit order-independent hash of every (key, COUNT, SUM) "
                     f"row, persistent vs per-query; " if q is not None else "")
                    + f"{len(check_q)} queries row for row against the oracle",
        "oracle_subset_ok": oracle_ok,
    })
    return out, (rows, dist, keys, vals, p["groups"])


def _serve_trace(ctx, tga, pool, bigk, bigv, big_d, rate, seconds, seed):
    """Poisson arrivals at `rate` for `seconds`: 95% requests drawn from the
    batch pool, 5% 2^26-row int64-key requests; returns the query array, the
    arrival times and the expected group count of every request."""
    rows, dist, keys, vals, groups = pool
    rng = np.random.default_rng(seed)
    n = int(rate * seconds)
    arr = np.cumsum(rng.exponential(1e6 / rate, n))
    big_n = 1 << 26
    aggs2 = (tga.Agg * 2)(tga.Agg(tga.COUNT, 0), tga.Agg(tga.SUM, 0))
    qs = (tga.Query * n)()
    expect = np.zeros(n, np.int64)
    keep = [aggs2]
    nlarge = 0
    for i in range(n):
        if rng.random() < 0.05:
            j = int(rng.integers(len(big_d)))
            vc = (tga.Column * 1)(bigv[j].column())
            qs[i].key = bigk[j].column()
            qs[i].n_rows = big_n
            qs[i].est_groups = big_d[j]
            expect[i] = big_d[j]
            nlarge += 1
        else:
            j = int(rng.integers(len(rows)))
            vc = (tga.Column * 1)(vals[j].column())
            qs[i].key = keys[j].column()
            qs[i].n_rows = int(rows[j])
            qs[i].est_groups = int(dist[j])
            expect[i] = groups[j]
        keep.append(vc)
        qs[i].values = ctypes.cast(vc, ctypes.POINTER(tga.Column))
        qs[i].n_values = 1
        qs[i].aggs = ctypes.cast(aggs2, ctypes.POINTER(tga.Agg))
        qs[i].n_aggs = 2
    arr_c = (ctypes.c_double * n)(*arr.tolist())
    return qs, arr_c, expect, nlarge, keep


def _serve_once(ctx, tga, qs, arr_c, expect, nlarge):
    from cpu_gpu_coprocess_groupby_agg_amd import lib
    n = len(expect)
    srv = ctypes.c_void_p()
    ctx._check(lib().tga_server_create(ctx.h, 64 << 30, 512, ctypes.byref(srv)))
    st = tga.ServerStats()
    wall = ctypes.c_double()
    got = np.zeros(n, np.int64)
    ctx._check(lib().tga_server_replay(srv, qs, arr_c, n, ctypes.byref(st), ctypes.byref(wall),
                                       got.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    lib().tga_server_destroy(srv)
    return {"requests": n, "large": nlarge, "p50_us": round(st.p50_us, 1), "p99_us": round(st.p99_us, 1),
            "throughput_qps": round(n / wall.value, 1), "launches": st.launches, "solo_runs": st.solo_runs,
            "scheduler_build_s": round(st.build_s, 3), "scheduler_finish_s": round(st.finish_s, 3),
            "peak_admitted_gb": round(st.peak_bytes / 2**30, 2),
            "group_counts_ok": bool(np.array_equal(got, expect))}


def serving(ctx, tga, args, pool):
    """Config 5: the server at the configured arrival rate beside the plain
    operator run sequentially on the same trace, then a sweep of arrival rates
    for the highest one the server sustains (p99 under 50 ms, served rate
    within 10% of the offered one)."""
    from cpu_gpu_coprocess_groupby_agg_amd import lib
    big_n = 1 << 26
    big_d = [1 << 20, 3 << 18, 1 << 19, 5 << 17]
    # closed-form columns: request group counts are known exactly
    bigk = [ctx.generate(tga.GEN_MOD_U64, big_n, a=big_d[j], salt=900 + j) for j in range(len(big_d))]
    bigv = [ctx.generate(tga.GEN_SEQ_I32, big_n, a=j) for j in range(len(big_d))]
    rate = args.serve_rate
    qs, arr_c, expect, nlarge, keep = _serve_trace(ctx, tga, pool, bigk, bigv, big_d, rate, args.serve_seconds, 11)
    served = _serve_once(ctx, tga, qs, arr_c, expect, nlarge)
    log(f"serving at {rate:.0f}/s: {served}")
    n = len(expect)
    seq = tga.ServerStats()
    wall = ctypes.c_double()
    got2 = np.zeros(n, np.int64)
    ctx._check(lib().tga_sequential_replay(ctx.h, qs, arr_c, n, ctypes.byref(seq), ctypes.byref(wall),
                                           got2.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    sequential = {"p50_us": round(seq.p50_us, 1), "p99_us": round(seq.p99_us, 1),
                  "throughput_qps": round(n / wall.value, 1), "group_counts_ok": bool(np.array_equal(got2, expect))}
    log(f"sequential: {sequential}")
    sweep = []
    ok = served["group_counts_ok"] and sequential["group_counts_ok"]
    for r in args.serve_sweep:
        if r == rate:
            res = served
        else:
            q2, a2, e2, nl2, k2 = _serve_trace(ctx, tga, pool, bigk, bigv, big_d, r, args.serve_seconds, 12)
            res = _serve_once(ctx, tga, q2, a2, e2, nl2)
            ok &= res["group_counts_ok"]
        # a server that falls behind builds a queue, so its p99 grows with
        # the trace; the served rate over a short trace also counts the last
        # requests' latency against it, hence the 10% margin
        sus = res["throughput_qps"] >= 0.9 * r and res["p99_us"] <= 50_000
        sweep.append({"offered_qps": r, "throughput_qps": res["throughput_qps"], "p50_us": res["p50_us"],
                      "p99_us": res["p99_us"], "sustained": sus})
        log(f"sweep {r:.0f}/s: {sweep[-1]}")
    for x in bigk + bigv:
        x.free()
    assert ok
    best = max([x["offered_qps"] for x in sweep if x["sustained"]], default=None)
    return {"arrival_rate_qps": rate, "budget_gb": 64, "server": served, "sequential_plain": sequential,
            "sweep": sweep, "max_sustained_qps": best,
            "sustained_means": "p99 <= 50 ms and served rate >= 90% of offered over the trace"}


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
    if args.only == "config1":
        print(json.dumps({"config1": config1(ctx, tga, args)}))
        return
    if args.only == "config3":
        args.no_cpu = True
        print(json.dumps({"config3": config3(ctx, tga, args)}))
        return
    if args.only == "coprocess":
        print(json.dumps({"coprocess": coprocess(ctx, tga, args)}))
        return
    if args.only == "stream":
        print(json.dumps({"host_streaming": host_streaming(ctx, tga, args)}))
        return
    if args.only in ("batch", "serve"):
        b, pool = batched(ctx, tga, args)
        out = {"batched": b}
        if args.only == "serve":
            out["serving"] = serving(ctx, tga, args, pool)
        print(json.dumps(out))
        return
    h, keys, vals = headline(ctx, tga, args)
    log(f"headline (no hint): {h['ms_per_step']:.3f} ms/step, {h['value']:.3e} rows/s; hinted {h['hinted']}; "
        f"kernels {h['kernels_ms_per_step']}")
    if not args.no_traffic and args.only == "all":
        traffic, detail = measure_traffic(h["roofline"]["kernel"], args)
        if traffic:
            h["roofline"]["traffic"] = round(traffic)
            h["roofline"]["traffic_detail"] = detail
            h["roofline"]["traffic_over_algorithmic"] = round(traffic / h["roofline"]["algorithmic_bytes"], 3)
        log(f"traffic of {h['roofline']['kernel']}: {traffic} {detail}")
    cpu = None
    if args.only == "headline":
        args.no_cpu = args.no_batch = args.no_serve = args.no_configs = True
        args.no_workloads = True
    if not args.no_cpu:
        cpu = cpu_baseline(ctx, tga, keys, vals)
        log(f"cpu baseline: {cpu}")
    keys.free()
    vals.free()
    extra = {}
    pool = None
    if not args.no_configs:
        extra["config1"] = config1(ctx, tga, args)
        log(f"config1: {extra['config1']}")
        extra["config3"] = config3(ctx, tga, args)
        log(f"config3: {extra['config3']}")
    if not args.no_configs:
        extra["host_streaming"] = host_streaming(ctx, tga, args)
        log(f"host streaming: {extra['host_streaming']}")
        extra["coprocess"] = coprocess(ctx, tga, args)
    if not args.no_workloads:
        extra["workloads"] = workloads(ctx, tga, args)
    if not args.no_batch:
        extra["batched"], pool = batched(ctx, tga, args)
    if not args.no_serve:
        if pool is None:
            args.batch_queries = 512
            _, pool = batched(ctx, tga, args)
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
                               "[-2^30, 2^30), COUNT/SUM(int64)/MIN/MAX, no cardinality hint",
                   "rows": h["n"], "groups": h["groups"], "plan": h["plan"]},
        "roofline": h["roofline"],
        "op_frac_of_hbm_peak": h["op_frac_of_peak"],
        "hinted": h["hinted"],
        "kernels_ms_per_step": h["kernels_ms_per_step"],
        "cpu_baseline": cpu,
        "copy_reference": copy,
        "device": info,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
