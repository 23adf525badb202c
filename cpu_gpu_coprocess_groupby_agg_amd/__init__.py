"""Host-side Python mirror of the MI355X tiled hash group-by.

A ctypes binding over the C ABI in include/tga_groupby.h (libtga.so, built
in-tree by ``build.py``).  It mirrors the reference operator surface
(``groupby_agg_partition`` / ``groupby_agg_intra_partition`` of
tiled_groupby_agg_partition/include/*.cuh, reference repository) for callers
that hold numpy arrays or device buffers:

    ctx = Context()                       # one GPU, one HIP stream
    k = ctx.upload(keys); v = ctx.upload(vals)
    res = ctx.groupby(k, [v], [("count",), ("sum", 0), ("min", 0), ("max", 0)])

Device memory comes from the library itself (no torch), so exactly one HIP
runtime lives in the process.  Every call raises TgaError on failure; there is
no CPU fallback: if libtga.so is missing, importing the native layer fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtga.so")

# status codes (tga_status)
OK = 0
ERR_INVALID_ARG = 1
ERR_HIP = 2
ERR_OUT_OF_MEMORY = 3
ERR_TABLE_OVERFLOW = 4
ERR_BUDGET = 5
ERR_NOT_READY = 6
ERR_RANGE = 7
ERR_UNSUPPORTED = 8
ERR_CAPACITY = 9

# dtypes (tga_dtype)
U32, I32, U64, I64, F64 = 1, 2, 3, 4, 5
_NP = {U32: np.uint32, I32: np.int32, U64: np.uint64, I64: np.int64, F64: np.float64}
_FROM_NP = {np.dtype(np.uint32): U32, np.dtype(np.int32): I32, np.dtype(np.uint64): U64,
            np.dtype(np.int64): I64, np.dtype(np.float64): F64}

COUNT, SUM, MIN, MAX = 0, 1, 2, 3
_OPS = {"count": COUNT, "sum": SUM, "min": MIN, "max": MAX}

PATH_AUTO, PATH_DIRECT, PATH_PARTITIONED = 0, 1, 2

GEN_KEYS_U32, GEN_KEYS_U64, GEN_VALS_I32, GEN_ZIPF_U64, GEN_NORMAL_F64, GEN_REF_U32 = 1, 2, 3, 4, 5, 6
GEN_MOD_U32, GEN_MOD_U64, GEN_SEQ_I32 = 7, 8, 9


class TgaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"tga status {status}: {msg}")
        self.status = status


class Column(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dtype", ctypes.c_int32), ("location", ctypes.c_int32)]


class Agg(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("column", ctypes.c_int32)]


class Options(ctypes.Structure):
    _fields_ = [("est_groups", ctypes.c_int64), ("path", ctypes.c_int32), ("partition_bits", ctypes.c_int32),
                ("table_slots", ctypes.c_int32), ("tile_rows", ctypes.c_int32), ("workgroups", ctypes.c_int32),
                ("block_threads", ctypes.c_int32), ("chunk_rows", ctypes.c_int64),
                ("wave_fold", ctypes.c_int32), ("attempts", ctypes.c_int32), ("scatter_layout", ctypes.c_int32),
                ("write_combine", ctypes.c_int32), ("streamed_tiles", ctypes.c_int32),
                ("stream_pinned", ctypes.c_int32), ("record_words", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Query(ctypes.Structure):
    _fields_ = [("key", Column), ("values", ctypes.POINTER(Column)), ("n_values", ctypes.c_int32),
                ("n_aggs", ctypes.c_int32), ("aggs", ctypes.POINTER(Agg)), ("n_rows", ctypes.c_int64),
                ("est_groups", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("rejected", ctypes.c_int64),
                ("inflight", ctypes.c_int64), ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("max_us", ctypes.c_double), ("elapsed_s", ctypes.c_double),
                ("qps", ctypes.c_double), ("bytes_in_use", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64),
                ("budget_bytes", ctypes.c_uint64), ("launches", ctypes.c_int64), ("solo_runs", ctypes.c_int64),
                ("build_s", ctypes.c_double), ("finish_s", ctypes.c_double)]


class CoprocessStats(ctypes.Structure):
    _fields_ = [("gpu_units", ctypes.c_int64), ("cpu_units", ctypes.c_int64), ("gpu_rows", ctypes.c_int64),
                ("cpu_rows", ctypes.c_int64), ("wall_s", ctypes.c_double), ("cpu_rows_per_s", ctypes.c_double),
                ("gpu_rows_per_s", ctypes.c_double)]


_lib = None


def lib():
    """The loaded libtga.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m cpu_gpu_coprocess_groupby_agg_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, u64, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_size_t
        P = ctypes.POINTER
        sigs = {
            "tga_context_create": [i32, u64, P(vp)],
            "tga_context_destroy": [vp],
            "tga_options_default": [P(Options)],
            "tga_device_info": [vp, ctypes.c_char_p, sz, P(i32), P(i64)],
            "tga_groupby": [vp, P(Column), P(Column), i32, i64, P(Agg), i32, P(Options), P(vp)],
            "tga_groupby_into": [vp, P(Column), P(Column), i32, i64, P(Agg), i32, P(Options), vp, P(vp), i64, P(i64)],
            "tga_groupby_into_noalloc": [vp, P(Column), P(Column), i32, i64, P(Agg), i32, P(Options), vp, P(vp), i64,
                                         P(i64)],
            "tga_result_num_groups": [vp, P(i64)],
            "tga_result_column": [vp, i32, P(vp), P(i32)],
            "tga_result_copy_to_host": [vp, i32, vp, i64],
            "tga_result_free": [vp],
            "tga_last_plan": [vp, P(Options)],
            "tga_scan_exclusive_u32": [vp, vp, vp, i64],
            "tga_generate": [vp, i32, vp, i64, u64, u64, ctypes.c_double, ctypes.c_double, u64, u64],
            "tga_device_alloc": [vp, u64, P(vp)],
            "tga_host_alloc": [vp, u64, P(vp)],
            "tga_device_free": [vp, vp],
            "tga_memcpy_h2d": [vp, vp, vp, u64],
            "tga_memcpy_d2h": [vp, vp, vp, u64],
            "tga_synchronize": [vp],
            "tga_device_copy": [vp, vp, vp, u64],
            "tga_profile_enable": [vp, i32],
            "tga_profile_reset": [vp],
            "tga_profile_entry": [vp, i32, ctypes.c_char_p, sz, P(i64), P(ctypes.c_double)],
        }
        optional = {
            "tga_batch_create": [vp, P(Query), i32, i32, P(vp)],
            "tga_batch_run": [vp],
            "tga_batch_num_groups": [vp, i32, P(i64)],
            "tga_batch_copy_to_host": [vp, i32, i32, vp, i64],
            "tga_batch_destroy": [vp],
            "tga_batch_profile": [vp, P(ctypes.c_double), P(ctypes.c_double), P(i64), P(ctypes.c_double),
                                  P(ctypes.c_double)],
            "tga_server_create": [vp, u64, i32, P(vp)],
            "tga_server_submit": [vp, P(Query), P(i64)],
            "tga_server_poll": [vp, i64, P(vp)],
            "tga_server_wait": [vp, i64, P(vp)],
            "tga_server_stats_get": [vp, P(ServerStats)],
            "tga_server_destroy": [vp],
            "tga_server_replay": [vp, P(Query), P(ctypes.c_double), i64, P(ServerStats), P(ctypes.c_double), P(i64)],
            "tga_sequential_replay": [vp, P(Query), P(ctypes.c_double), i64, P(ServerStats), P(ctypes.c_double),
                                      P(i64)],
            "tga_compat_partition": [vp, vp, vp, sz, sz, sz, sz, P(vp)],
            "tga_partset_count": [vp, P(sz)],
            "tga_partset_info": [vp, sz, P(sz), P(u64)],
            "tga_partset_segment": [vp, sz, sz, P(vp), P(vp), P(ctypes.c_uint32)],
            "tga_partset_create": [vp, sz, P(vp)],
            "tga_partset_push": [vp, sz, vp, vp, ctypes.c_uint32],
            "tga_partset_destroy": [vp],
            "tga_compat_intra_partition": [vp, vp, sz, sz, sz, sz, P(vp), P(vp), P(sz), P(sz), P(sz)],
            "tga_host_free": [vp],
            "tga_set_coprocess": [vp, i32, ctypes.c_double],
            "tga_last_coprocess": [vp, i32, P(CoprocessStats)],
        }
        for name, argt in list(sigs.items()) + list(optional.items()):
            fn = getattr(L, name, None)
            if fn is None:
                if name in sigs:
                    raise ImportError(f"libtga.so lacks {name}")
                continue
            fn.argtypes = argt
            fn.restype = i32
        L.tga_last_error.argtypes = [vp]
        L.tga_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _agg_list(aggs):
    out = []
    for a in aggs:
        if isinstance(a, str):
            a = (a,)
        op = _OPS[a[0]] if isinstance(a[0], str) else int(a[0])
        col = int(a[1]) if len(a) > 1 else 0
        out.append((op, col))
    return out


def out_dtype(op, col_dtype):
    """Output dtype of one aggregate (mirrors the C ABI rules)."""
    if op == COUNT:
        return I64
    if op == SUM:
        return F64 if col_dtype == F64 else I64
    return col_dtype


class DeviceArray:
    """A device buffer allocated through the library (freed on close/GC)."""

    def __init__(self, ctx, n, dtype):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = dtype
        self.nbytes = self.n * np.dtype(_NP[dtype]).itemsize
        p = ctypes.c_void_p()
        ctx._check(lib().tga_device_alloc(ctx.h, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    def to_numpy(self):
        a = np.empty(self.n, dtype=_NP[self.dtype])
        if self.n:
            self.ctx._check(lib().tga_memcpy_d2h(self.ctx.h, a.ctypes.data, self.ptr, self.nbytes))
        return a

    def column(self):
        return Column(self.ptr, self.dtype, 0)

    def free(self):
        if self.ptr is not None and self.ctx.h:
            lib().tga_device_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device + stream + HBM arena (tga_context)."""

    def __init__(self, device=0, budget_bytes=0):
        h = ctypes.c_void_p()
        st = lib().tga_context_create(device, budget_bytes, ctypes.byref(h))
        if st != OK:
            raise TgaError(st, "tga_context_create failed (no GPU?)")
        self.h = h.value

    def close(self):
        if self.h:
            lib().tga_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != OK:
            raise TgaError(st, lib().tga_last_error(self.h).decode())

    def device_info(self):
        name = ctypes.create_string_buffer(128)
        cus = ctypes.c_int32()
        hbm = ctypes.c_int64()
        self._check(lib().tga_device_info(self.h, name, 128, ctypes.byref(cus), ctypes.byref(hbm)))
        return {"name": name.value.decode(), "cus": cus.value, "hbm_bytes": hbm.value}

    # -- memory ---------------------------------------------------------------
    def empty(self, n, dtype):
        return DeviceArray(self, n, dtype)

    def pinned_empty(self, n, dtype):
        """A numpy array in pinned host memory (tga_host_alloc), freed with
        the array.  Host columns in pinned memory are streamed to the GPU with
        the copies overlapping the aggregation."""
        import weakref
        dt = np.dtype(dtype)
        p = ctypes.c_void_p()
        self._check(lib().tga_host_alloc(self.h, max(1, int(n) * dt.itemsize), ctypes.byref(p)))
        buf = (ctypes.c_char * max(1, int(n) * dt.itemsize)).from_address(p.value)
        arr = np.frombuffer(buf, dtype=dt, count=int(n))
        weakref.finalize(buf, lib().tga_host_free, p.value)
        return arr

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        d = DeviceArray(self, arr.size, _FROM_NP[arr.dtype])
        if arr.size:
            self._check(lib().tga_memcpy_h2d(self.h, d.ptr, arr.ctypes.data, d.nbytes))
        return d

    def generate(self, kind, n, a=0, b=0, x=0.0, y=0.0, seed=1, salt=0):
        dt = {GEN_KEYS_U32: U32, GEN_KEYS_U64: U64, GEN_VALS_I32: I32, GEN_ZIPF_U64: U64, GEN_NORMAL_F64: F64,
              GEN_REF_U32: U32, GEN_MOD_U32: U32, GEN_MOD_U64: U64, GEN_SEQ_I32: I32}[kind]
        d = DeviceArray(self, n, dt)
        self._check(lib().tga_generate(self.h, kind, d.ptr, n, a & (2**64 - 1), b, x, y, seed, salt))
        return d

    def copy(self, dst, src, nbytes):
        self._check(lib().tga_device_copy(self.h, dst.ptr, src.ptr, nbytes))

    def synchronize(self):
        self._check(lib().tga_synchronize(self.h))

    # -- operator -------------------------------------------------------------
    @staticmethod
    def make_options(**kw):
        o = Options()
        lib().tga_options_default(ctypes.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def groupby(self, key, values, aggs, n_rows=None, sort=True, **opts):
        """GROUP BY key over DeviceArray or numpy columns.

        Returns {"keys": ndarray, "aggs": [ndarray, ...], "n_groups": int},
        sorted by key (unsigned order of the key bits) when sort=True.
        """
        keep = []

        def col(x):
            if isinstance(x, DeviceArray):
                return x.column(), x.n
            a = np.ascontiguousarray(x)
            keep.append(a)
            return Column(a.ctypes.data, _FROM_NP[a.dtype], 1), a.size

        kcol, n = col(key)
        if n_rows is None:
            n_rows = n
        vcols = (Column * max(1, len(values)))()
        vdt = []
        for i, v in enumerate(values):
            c, _ = col(v)
            vcols[i] = c
            vdt.append(c.dtype)
        al = _agg_list(aggs)
        ag = (Agg * len(al))(*[Agg(o, c) for o, c in al])
        o = self.make_options(**opts)
        r = ctypes.c_void_p()
        self._check(lib().tga_groupby(self.h, ctypes.byref(kcol), vcols, len(values), n_rows, ag, len(al),
                                      ctypes.byref(o), ctypes.byref(r)))
        try:
            return self._fetch(r.value, kcol.dtype, [out_dtype(op, vdt[c] if op != COUNT else 0) for op, c in al],
                               sort)
        finally:
            lib().tga_result_free(r.value)

    def groupby_resident(self, key, values, aggs, n_rows=None, **opts):
        """Runs the operator on DeviceArray columns and leaves the result in
        HBM (no copy to the host); returns (result handle, group count).
        Release the handle with free_result."""
        kcol = key.column()
        n = key.n if n_rows is None else n_rows
        vcols = (Column * max(1, len(values)))(*[v.column() for v in values])
        al = _agg_list(aggs)
        ag = (Agg * len(al))(*[Agg(o, c) for o, c in al])
        o = self.make_options(**opts)
        r = ctypes.c_void_p()
        self._check(lib().tga_groupby(self.h, ctypes.byref(kcol), vcols, len(values), n, ag, len(al),
                                      ctypes.byref(o), ctypes.byref(r)))
        g = ctypes.c_int64()
        self._check(lib().tga_result_num_groups(r.value, ctypes.byref(g)))
        return r.value, g.value

    @staticmethod
    def free_result(r):
        if r:
            lib().tga_result_free(r)

    def _fetch(self, r, key_dtype, dts, sort):
        g = ctypes.c_int64()
        self._check(lib().tga_result_num_groups(r, ctypes.byref(g)))
        G = g.value
        keys = np.empty(G, dtype=_NP[key_dtype])
        if G:
            self._check(lib().tga_result_copy_to_host(r, -1, keys.ctypes.data, G))
        outs = []
        for i, dt in enumerate(dts):
            a = np.empty(G, dtype=_NP[dt])
            if G:
                self._check(lib().tga_result_copy_to_host(r, i, a.ctypes.data, G))
            outs.append(a)
        if sort and G:
            ku = keys.view(np.uint32 if keys.itemsize == 4 else np.uint64)
            order = np.argsort(ku, kind="stable")
            keys = keys[order]
            outs = [a[order] for a in outs]
        return {"keys": keys, "aggs": outs, "n_groups": G}

    def set_coprocess(self, cpu_threads, cpu_share):
        """CPU workers for the compat phases: cpu_share 0 = GPU only, (0, 1] =
        that share of the rows, < 0 = the planner decides."""
        self._check(lib().tga_set_coprocess(self.h, int(cpu_threads), float(cpu_share)))

    def last_coprocess(self, phase):
        st = CoprocessStats()
        self._check(lib().tga_last_coprocess(self.h, int(phase), ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in CoprocessStats._fields_}

    def last_plan(self):
        o = Options()
        self._check(lib().tga_last_plan(self.h, ctypes.byref(o)))
        return {f: getattr(o, f) for f, _ in Options._fields_ if f != "reserved"}

    def scan_exclusive_u32(self, d_in, d_out):
        self._check(lib().tga_scan_exclusive_u32(self.h, d_in.ptr, d_out.ptr, d_in.n))

    # -- kernel timing ------------------------------------------------------------
    def profile(self, on=True):
        self._check(lib().tga_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        self._check(lib().tga_profile_reset(self.h))

    def profile_entries(self):
        out = {}
        i = 0
        name = ctypes.create_string_buffer(128)
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        while lib().tga_profile_entry(self.h, i, name, 128, ctypes.byref(n), ctypes.byref(ms)) == OK:
            out[name.value.decode()] = (n.value, ms.value)
            i += 1
        return out

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
