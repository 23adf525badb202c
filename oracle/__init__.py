"""CPU oracle for the tiled hash group-by -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker (or the timed CPU baseline).  The
product library never loads it.

ctypes binding of liboracle.so (oracle/tga_oracle.cpp, built by
oracle/Makefile).  See tga_oracle.cpp for what each function restates and
which reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DATAGEN = os.path.join(HERE, "_ref", "ref_datagen")

U32, I32, U64, I64, F64 = 1, 2, 3, 4, 5
COUNT, SUM, MIN, MAX = 0, 1, 2, 3
_NP = {U32: np.uint32, I32: np.int32, U64: np.uint64, I64: np.int64, F64: np.float64}
_FROM_NP = {np.dtype(v): k for k, v in _NP.items()}
_OPS = {"count": COUNT, "sum": SUM, "min": MIN, "max": MAX}


class RunStats(ctypes.Structure):
    _fields_ = [("tiles", ctypes.c_int64), ("phase1_rows", ctypes.c_int64), ("merged_parts", ctypes.c_int64),
                ("groups", ctypes.c_int64), ("t_phase1_s", ctypes.c_double), ("t_merge_s", ctypes.c_double),
                ("t_phase2_s", ctypes.c_double)]


def build():
    """Compiles liboracle.so (and the reference generator where the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/test/include"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, u64, sz, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
        P = ctypes.POINTER
        L.tga_o_splitmix.argtypes = [u64, u64]
        L.tga_o_splitmix.restype = u64
        L.tga_o_key_u32.argtypes = [u64, u64]
        L.tga_o_key_u32.restype = ctypes.c_uint32
        L.tga_o_key_u64.argtypes = [u64, u64]
        L.tga_o_key_u64.restype = u64
        L.tga_o_gen_keys_u32.argtypes = [vp, i64, u64, u64, u64]
        L.tga_o_gen_keys_u64.argtypes = [vp, i64, u64, u64, u64]
        L.tga_o_gen_vals_i32.argtypes = [vp, i64, i64, i64, u64]
        L.tga_o_groupby.argtypes = [vp, i32, i64, P(vp), P(i32), i32, P(i32), P(i32), i32, vp, P(vp), i64]
        L.tga_o_groupby.restype = i64
        L.tga_o_neumaier.argtypes = [vp, i64]
        L.tga_o_neumaier.restype = ctypes.c_double
        L.tga_o_scan_u64.argtypes = [vp, vp, i64, i32]
        L.tga_o_murmur3_u32.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.tga_o_murmur3_u32.restype = ctypes.c_uint32
        ref_args = [vp, vp, sz, sz, sz, sz, sz, sz, i32, vp, vp, vp, vp, vp, sz, P(i64), vp, vp, i64, P(i64),
                    P(RunStats)]
        L.tga_o_ref_u32.argtypes = ref_args
        L.tga_o_ref_u32.restype = i32
        L.tga_o_ref_i64.argtypes = ref_args
        L.tga_o_ref_i64.restype = i32
        L.tga_o_ref_phase1_u32.argtypes = [vp, vp, sz, sz, sz, i32, vp, i64]
        L.tga_o_ref_phase1_u32.restype = i64
        L.tga_o_merge_sizes.argtypes = [vp, i64, u64, vp, i64]
        L.tga_o_merge_sizes.restype = i64
        L.tga_o_partition_of_u32.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.tga_o_partition_of_u32.restype = ctypes.c_uint32
        _lib = L
    return _lib


# ---- generators (bit-identical with the device generators) -----------------
def gen_keys_u32(n, distinct, seed, salt=0):
    a = np.empty(n, np.uint32)
    lib().tga_o_gen_keys_u32(a.ctypes.data, n, distinct, seed, salt)
    return a


def gen_keys_u64(n, distinct, seed, salt=0):
    a = np.empty(n, np.uint64)
    lib().tga_o_gen_keys_u64(a.ctypes.data, n, distinct, seed, salt)
    return a


def gen_vals_i32(n, lo, hi, seed):
    a = np.empty(n, np.int32)
    lib().tga_o_gen_vals_i32(a.ctypes.data, n, lo, hi, seed)
    return a


# ---- exact group-by ----------------------------------------------------------
def _agg_list(aggs):
    out = []
    for a in aggs:
        if isinstance(a, str):
            a = (a,)
        op = _OPS[a[0]] if isinstance(a[0], str) else int(a[0])
        out.append((op, int(a[1]) if len(a) > 1 else 0))
    return out


def out_dtype(op, col_dtype):
    if op == COUNT:
        return I64
    if op == SUM:
        return F64 if col_dtype == F64 else I64
    return col_dtype


def groupby(keys, cols, aggs):
    """Exact GROUP BY (Neumaier for fp64 SUM), sorted by key bits."""
    keys = np.ascontiguousarray(keys)
    cols = [np.ascontiguousarray(c) for c in cols]
    n = keys.size
    al = _agg_list(aggs)
    kb = keys.itemsize
    cap = max(n, 1)
    okeys = np.empty(cap, np.uint64)
    dts = [_FROM_NP[c.dtype] for c in cols]
    outs = [np.empty(cap, _NP[out_dtype(op, dts[c] if op != COUNT else 0)]) for op, c in al]
    cp = (ctypes.c_void_p * max(1, len(cols)))(*[c.ctypes.data for c in cols])
    cd = (ctypes.c_int * max(1, len(cols)))(*dts)
    ops = (ctypes.c_int * len(al))(*[o for o, _ in al])
    acs = (ctypes.c_int * len(al))(*[c for _, c in al])
    op_ = (ctypes.c_void_p * len(al))(*[o.ctypes.data for o in outs])
    G = lib().tga_o_groupby(keys.ctypes.data, kb, n, cp, cd, len(cols), ops, acs, len(al), okeys.ctypes.data, op_, cap)
    assert G >= 0
    k = okeys[:G].astype(keys.dtype) if kb == 8 else okeys[:G].astype(np.uint32).view(keys.dtype)
    return {"keys": k, "aggs": [o[:G] for o in outs], "n_groups": int(G)}


def neumaier(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().tga_o_neumaier(x.ctypes.data, x.size)


def scan(x, exclusive=True):
    x = np.ascontiguousarray(x, dtype=np.uint64)
    out = np.empty_like(x)
    lib().tga_o_scan_u64(x.ctypes.data, out.ctypes.data, x.size, 1 if exclusive else 0)
    return out


# ---- restatement of the reference algorithm -------------------------------------
def ref_run(keys, vals, tile_len, P, capacity, min_load, max_load, nthreads=4, with_output=True):
    """Two-phase reference algorithm on host threads.

    keys u32 + vals u32 -> the reference's own types; keys u64 + vals i32 ->
    the widened CPU baseline.  Returns dict with keys/count/sum/min/max in the
    reference's per-merged-partition order, par_begin/par_count and stats;
    raises OverflowError on the reference's "hash table overflow" rule.
    """
    keys = np.ascontiguousarray(keys)
    vals = np.ascontiguousarray(vals)
    n = keys.size
    wide = keys.dtype == np.uint64
    fn = lib().tga_o_ref_i64 if wide else lib().tga_o_ref_u32
    cap = max(n, 1) if with_output else 0
    ok = np.empty(cap, keys.dtype)
    oc = np.empty(cap, np.uint64)
    osum = np.empty(cap, np.int64)
    omin = np.empty(cap, np.int64)
    omax = np.empty(cap, np.int64)
    pb = np.empty(max(P, 1), np.int64)
    pc = np.empty(max(P, 1), np.int64)
    G = ctypes.c_int64()
    NP = ctypes.c_int64()
    stats = RunStats()
    ptr = (lambda a: a.ctypes.data) if with_output else (lambda a: None)
    rc = fn(keys.ctypes.data, vals.ctypes.data, n, tile_len, P, capacity, min_load, max_load, nthreads,
            ptr(ok), ptr(oc), ptr(osum), ptr(omin), ptr(omax), cap, ctypes.byref(G), pb.ctypes.data, pc.ctypes.data,
            P, ctypes.byref(NP), ctypes.byref(stats))
    if rc == 4:
        raise OverflowError("hash table overflow")
    assert rc == 0, rc
    g = G.value
    out = {"n_groups": g, "n_par": NP.value, "par_begin": pb[:NP.value].copy(), "par_count": pc[:NP.value].copy(),
           "stats": {f: getattr(stats, f) for f, _ in RunStats._fields_}}
    if with_output:
        out.update(keys=ok[:g], count=oc[:g], sum=osum[:g], min=omin[:g], max=omax[:g])
    return out


def ref_phase1_u32(keys, vals, tile_len, P, nthreads=2):
    """Phase 1 only: mutates keys/vals in place; returns (partition, begin, len) segments."""
    n = keys.size
    cap = ((n + tile_len - 1) // tile_len) * P + 16
    seg = np.empty(cap * 3, np.int64)
    k = lib().tga_o_ref_phase1_u32(keys.ctypes.data, vals.ctypes.data, n, tile_len, P, nthreads, seg.ctypes.data, cap)
    assert k <= cap
    return seg[:3 * k].reshape(k, 3)


def merge_sizes(sizes, K):
    s = np.ascontiguousarray(sizes, dtype=np.uint32)
    out = np.empty(max(s.size, 1), np.uint64)
    m = lib().tga_o_merge_sizes(s.ctypes.data, s.size, K, out.ctypes.data, out.size)
    return out[:m]


def partition_of_u32(key, P):
    return lib().tga_o_partition_of_u32(int(key), int(P))


def ref_datagen(n, pct, dist, path):
    """Runs the reference's own generator (built from the reference checkout)."""
    subprocess.run([REF_DATAGEN, str(n), str(pct), str(dist), path], check=True, capture_output=True)
    raw = np.fromfile(path, dtype=np.uint32)
    return raw[:n].copy(), raw[n:].copy()
