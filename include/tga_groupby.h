/* tga_groupby.h -- C ABI of the MI355X (gfx950) tiled hash group-by.
 *
 * Plain C: POD structs, int status codes, no torch or HIP types.  Every
 * function is thread-safe per context (calls on one context serialise on its
 * lock; independent contexts run concurrently); nothing prints or exits.
 *
 * Reference interfaces replaced (T/ = tiled_groupby_agg_partition/ of the
 * reference repository):
 *   groupby_agg_partition<K,V>      T/include/group_then_partition.cuh:5-12,
 *                                   T/include/lib_header.hpp:6-13
 *       -> tga_compat_partition (phase 1 over host buffers) and, for
 *          device-resident columns, tga_groupby (both phases fused)
 *   groupby_agg_intra_partition     T/include/groupby_agg_intra_partition.cuh:5-13
 *       -> tga_compat_intra_partition (phase 2) / tga_groupby
 *   class par_result                T/include/par_result.cuh:11-82
 *       -> tga_partset (opaque; segment accessors below)
 *   scan_inclusive/scan_exclusive   T/include/util.cuh:21-35 (CUB wrappers)
 *       -> tga_scan_exclusive_u32 (own decoupled look-back scan)
 * The header-only C++ layer include/tga_compat.hpp rebuilds the reference's
 * three symbols by name and parameter list on top of these entry points.
 */
#ifndef TGA_GROUPBY_H_
#define TGA_GROUPBY_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TGA_ABI_VERSION 1

typedef enum tga_status {
  TGA_OK = 0,
  TGA_ERR_INVALID_ARG = 1,
  TGA_ERR_HIP = 2,            /* a HIP runtime call failed; see tga_last_error */
  TGA_ERR_OUT_OF_MEMORY = 3,  /* device allocation failed or the context budget is exhausted */
  TGA_ERR_TABLE_OVERFLOW = 4, /* compat phase 2: the reference's "hash table overflow"
                                 (T/src/groupby_agg_intra_partition.cu:233-236, exit(2)) */
  TGA_ERR_BUDGET = 5,         /* serving: a request can never fit the memory budget */
  TGA_ERR_NOT_READY = 6,      /* serving: request still in flight */
  TGA_ERR_RANGE = 7,          /* fp64 SUM input outside the exact window (|x| >= 2^64, Inf, NaN) */
  TGA_ERR_UNSUPPORTED = 8,
  TGA_ERR_CAPACITY = 9        /* caller-provided output smaller than the group count */
} tga_status;

typedef enum tga_dtype { TGA_U32 = 1, TGA_I32 = 2, TGA_U64 = 3, TGA_I64 = 4, TGA_F64 = 5 } tga_dtype;

/* Aggregates.  Output dtypes: COUNT -> I64; SUM of an integer column -> I64
 * (exact, wraps mod 2^64; its low 32 bits equal the reference's u32
 * wrap-around sum); SUM of an F64 column -> F64, the exact sum rounded once
 * (bit-reproducible whatever the row order); MIN/MAX -> the column dtype. */
typedef enum tga_agg_op { TGA_COUNT = 0, TGA_SUM = 1, TGA_MIN = 2, TGA_MAX = 3 } tga_agg_op;

typedef enum tga_location { TGA_DEVICE = 0, TGA_HOST = 1 } tga_location;

typedef struct tga_column {
  const void *data;
  int32_t dtype;    /* tga_dtype */
  int32_t location; /* tga_location; host columns are staged through HBM */
} tga_column;

typedef struct tga_agg {
  int32_t op;     /* tga_agg_op */
  int32_t column; /* index into the value columns (ignored by COUNT) */
} tga_agg;

typedef enum tga_path { TGA_PATH_AUTO = 0, TGA_PATH_DIRECT = 1, TGA_PATH_PARTITIONED = 2 } tga_path;

/* Every field 0 = chosen by the autotuner from (rows, estimated groups). */
typedef struct tga_options {
  int64_t est_groups;   /* cardinality hint; 0 = estimate from a 16384-key sample,
                           or, when the sample bounds the count only loosely, from
                           an exact count over a 2^-10 hash slice of all keys (one
                           read of the key column) */
  int32_t path;         /* tga_path */
  int32_t partition_bits; /* log2 of the partition fanout */
  int32_t table_slots;  /* LDS hash-table slots per workgroup */
  int32_t tile_rows;    /* rows per partitioning tile */
  int32_t workgroups;   /* persistent-grid size */
  int32_t block_threads;/* threads per workgroup of the aggregation kernel */
  int64_t chunk_rows;   /* max rows per aggregation work item */
  int32_t wave_fold;    /* 0 = tuner decides, 1 = fold equal keys of a wave before the LDS atomics, 2 = never */
  int32_t attempts;     /* tga_last_plan only: executions the last call needed (finer fanout / larger tables) */
  int32_t scatter_layout; /* partitioned copy: 0 = tuner, 1 = partition-contiguous (a histogram pass first),
                             2 = per-workgroup buckets (no histogram pass) */
  int32_t write_combine;  /* bucket layout with one 4-byte value column: 0 or 1 = keep each partition's
                             unfinished cache line in LDS and write whole lines (when the LDS holds it),
                             2 = never */
  int32_t streamed_tiles; /* tga_last_plan only: host tiles the input was streamed in (host-resident
                             columns of >= 2^22 rows), 0 = device-resident or staged whole */
  int32_t stream_pinned;  /* tga_last_plan only: 1 = the streamed host columns were pinned, so the copies
                             overlapped the aggregation */
  int32_t record_words;   /* tga_last_plan only: the partitioned copy was records of the key and this many
                             4-byte value words (the value columns in order), 0 = one array per column */
  int32_t reserved[1];
} tga_options;

typedef struct tga_context tga_context;
typedef struct tga_result tga_result;

/* ---- context ------------------------------------------------------------ */
tga_status tga_context_create(int32_t device, uint64_t mem_budget_bytes, tga_context **out);
tga_status tga_context_destroy(tga_context *ctx);
const char *tga_last_error(const tga_context *ctx);
tga_status tga_options_default(tga_options *opt);
tga_status tga_device_info(tga_context *ctx, char *name, size_t name_len, int32_t *cus, int64_t *hbm_bytes);

/* ---- plain operator ------------------------------------------------------- */
/* GROUP BY key: aggregates over n_rows rows.  Blocking. */
tga_status tga_groupby(tga_context *ctx, const tga_column *key, const tga_column *values, int32_t n_values,
                       int64_t n_rows, const tga_agg *aggs, int32_t n_aggs, const tga_options *opt,
                       tga_result **out);
/* Same, into caller-owned device buffers (out_aggs[i] sized for capacity rows
 * of aggregate i's output dtype).  *n_groups receives the true count; if it
 * exceeds capacity, TGA_ERR_CAPACITY is returned and only capacity rows are
 * written. */
tga_status tga_groupby_into(tga_context *ctx, const tga_column *key, const tga_column *values, int32_t n_values,
                            int64_t n_rows, const tga_agg *aggs, int32_t n_aggs, const tga_options *opt,
                            void *out_keys, void *const *out_aggs, int64_t capacity, int64_t *n_groups);

/* Same as tga_groupby_into, allocating nothing: no hipMalloc, no arena, no
 * allocating library call.  No counterpart in the reference; it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it).  Its temporaries live in a fixed 32 MiB
 * device array that each GPU holds until the process ends, so:
 *   - key and value columns must be device-resident (TGA_DEVICE);
 *   - it runs the direct path (LDS tables flushed into one global table)
 *     and holds at most half the table's slots in groups: 2^18 groups for a
 *     64-bit key with COUNT/SUM/MIN/MAX of an int32 column (fewer with fp64
 *     SUM); a larger est_groups hint or a path of TGA_PATH_PARTITIONED is
 *     rejected, and more groups than that return TGA_ERR_OUT_OF_MEMORY;
 *   - every call on one GPU shares the array: calls on one context
 *     serialise, calls that overlap on one GPU from different contexts race.
 * Results equal tga_groupby_into's. */
tga_status tga_groupby_into_noalloc(tga_context *ctx, const tga_column *key, const tga_column *values,
                                    int32_t n_values, int64_t n_rows, const tga_agg *aggs, int32_t n_aggs,
                                    const tga_options *opt, void *out_keys, void *const *out_aggs, int64_t capacity,
                                    int64_t *n_groups);

tga_status tga_result_num_groups(const tga_result *r, int64_t *n_groups);
/* index -1 = key column, else aggregate index.  Device pointer + dtype. */
tga_status tga_result_column(const tga_result *r, int32_t index, const void **device_ptr, int32_t *dtype);
tga_status tga_result_copy_to_host(const tga_result *r, int32_t index, void *dst, int64_t capacity_rows);
tga_status tga_result_free(tga_result *r);

/* Plan the autotuner chose for the last tga_groupby on this context
 * (path, partition_bits, table_slots, tile_rows, workgroups, ...). */
tga_status tga_last_plan(tga_context *ctx, tga_options *plan);

/* ---- batched persistent path ---------------------------------------------- */
typedef struct tga_query {
  tga_column key;
  const tga_column *values;
  int32_t n_values;
  int32_t n_aggs;
  const tga_agg *aggs;
  int64_t n_rows;
  int64_t est_groups; /* 0 = estimate */
} tga_query;

typedef struct tga_batch tga_batch;

typedef enum tga_batch_mode {
  TGA_BATCH_PERSISTENT = 0, /* one persistent launch over a device work queue, captured in a hipGraph */
  TGA_BATCH_PER_QUERY = 1   /* the plain operator's kernels, launched query by query */
} tga_batch_mode;

/* Builds the work queue and device state for a batch of device-resident
 * queries (columns must stay valid until tga_batch_destroy). */
tga_status tga_batch_create(tga_context *ctx, const tga_query *queries, int32_t n_queries, int32_t mode,
                            tga_batch **out);
/* Runs the whole batch (replays its graph) and waits for it. */
tga_status tga_batch_run(tga_batch *b);
tga_status tga_batch_num_groups(const tga_batch *b, int32_t query, int64_t *n_groups);
tga_status tga_batch_copy_to_host(const tga_batch *b, int32_t query, int32_t index, void *dst, int64_t capacity_rows);
tga_status tga_batch_destroy(tga_batch *b);
/* Per item type of the persistent work queue (INIT, COUNT, DIRECT, SCAN,
 * SCATTER, PAGG, COMPACT: 7 entries each) over the last run: workgroup-ms
 * spent working, workgroup-ms spent waiting for the items they depend on, the
 * item count, and workgroup-ms spent in the memory fences that publish and
 * acquire an item's results.  phase_ms (8 entries, may be NULL) splits the
 * work: SCATTER rank, local sort, write-out; DIRECT table set-up, rows;
 * DIRECT and PAGG flush or emission; PAGG table set-up, rows.  Only for a
 * persistent batch created while kernel timing (tga_profile_enable) was on;
 * TGA_ERR_UNSUPPORTED otherwise. */
tga_status tga_batch_profile(const tga_batch *b, double *work_ms, double *wait_ms, int64_t *items,
                             double *fence_ms, double *phase_ms);

/* ---- serving loop ----------------------------------------------------------- */
typedef struct tga_server tga_server;
typedef struct tga_server_stats {
  int64_t submitted, completed, rejected, inflight;
  double p50_us, p99_us, mean_us, max_us;
  double elapsed_s, qps;
  uint64_t bytes_in_use, peak_bytes, budget_bytes;
  int64_t launches;
  int64_t solo_runs;       /* requests run alone on a solo lane */
  double build_s, finish_s; /* scheduler thread: seconds building and finishing flights */
} tga_server_stats;

tga_status tga_server_create(tga_context *ctx, uint64_t budget_bytes, int32_t max_batch, tga_server **out);
/* Copies the query descriptor; columns must stay valid until the result is taken. */
tga_status tga_server_submit(tga_server *s, const tga_query *q, int64_t *ticket);
tga_status tga_server_poll(tga_server *s, int64_t ticket, tga_result **out);  /* TGA_ERR_NOT_READY while pending */
tga_status tga_server_wait(tga_server *s, int64_t ticket, tga_result **out);
tga_status tga_server_stats_get(tga_server *s, tga_server_stats *st);
tga_status tga_server_destroy(tga_server *s);
/* Load generator: submits qs[i] at arrival_us[i] after the call starts
 * (busy wait on the calling thread), awaits and frees every result, and
 * returns the server's stats; *wall_s = first submit -> last completion;
 * groups (optional, n entries) receives each request's group count. */
tga_status tga_server_replay(tga_server *s, const tga_query *qs, const double *arrival_us, int64_t n,
                             tga_server_stats *out, double *wall_s, int64_t *groups);
/* Baseline for the same trace: the plain operator, one query at a time in
 * arrival order (latency includes the wait behind earlier queries). */
tga_status tga_sequential_replay(tga_context *ctx, const tga_query *qs, const double *arrival_us, int64_t n,
                                 tga_server_stats *out, double *wall_s, int64_t *groups);

/* ---- compat phase entry points (reference call sequence, u32 keys/values) --- */
typedef struct tga_partset tga_partset;

/* Phase 1 of the reference (T/src/group_then_partition.cu:427-528): tiles of
 * tile_len rows of the pinned host buffers are pre-aggregated (u32
 * wrap-around SUM) and hash-partitioned into P partitions; each tile's result
 * is written back over the caller's buffers at the tile offset, and every
 * partition records (offset, length) segments into those buffers. */
tga_status tga_compat_partition(tga_context *ctx, uint32_t *host_keys, uint32_t *host_vals, size_t n,
                                size_t tile_len, size_t P, size_t nstreams, tga_partset **out);
tga_status tga_partset_count(const tga_partset *ps, size_t *n_partitions);
tga_status tga_partset_info(const tga_partset *ps, size_t p, size_t *n_segments, uint64_t *rows);
/* segment j of partition p: host pointers into the caller's buffers */
tga_status tga_partset_segment(const tga_partset *ps, size_t p, size_t j, uint32_t **keys, uint32_t **vals,
                               uint32_t *len);
/* Builds a partition set from caller segment lists (what a par_result holds). */
tga_status tga_partset_create(tga_context *ctx, size_t n_partitions, tga_partset **out);
tga_status tga_partset_push(tga_partset *ps, size_t p, uint32_t *keys, uint32_t *vals, uint32_t len);
tga_status tga_partset_destroy(tga_partset *ps);

/* Phase 2 (T/src/groupby_agg_intra_partition.cu:283-411): merge_partition
 * (min-heap, K = min_load), then per merged partition: gather, chunked
 * aggregation with the reference's overflow rule, compaction.  Outputs are
 * pinned host arrays allocated here (free with tga_host_free) in the
 * reference's gapped layout: merged partition i has par_result_kv_num[i]
 * groups starting at par_kv_begin[i].  par_kv_begin / par_result_kv_num must
 * hold one entry per partition of the set (the merged count *n_merged never
 * exceeds it).  The set is then replaced by *n_merged drained partitions,
 * as the reference replaces and drains its vector (intra:294). */
tga_status tga_compat_intra_partition(tga_context *ctx, tga_partset *ps, size_t capacity, size_t min_load,
                                      size_t max_load, size_t nstreams, uint32_t **out_keys, uint32_t **out_vals,
                                      size_t *par_kv_begin, size_t *par_result_kv_num, size_t *n_merged);
tga_status tga_host_free(void *p);

/* CPU + GPU co-processing of the compat phases (the reference's open item
 * "cpu process partition only or groupby_agg_partition in phase_1",
 * README.md:16): host worker threads take phase-1 tiles (pre-aggregation and
 * hash partitioning, written back over the caller's buffers like a GPU tile)
 * and phase-2 merged partitions (aggregated into the same gapped output) while
 * the GPU takes the others; every unit is done by exactly one side, so the
 * result is the same.  cpu_threads = 0 or cpu_share = 0: GPU only (the
 * default).  cpu_share in (0, 1]: the CPU takes that share of the rows (tiles
 * from the end of the input, partitions from the smallest).  cpu_share < 0:
 * a planner gives the CPU a unit only when its measured throughput finishes it
 * before the GPU would finish the rest. */
tga_status tga_set_coprocess(tga_context *ctx, int32_t cpu_threads, double cpu_share);
typedef struct tga_coprocess_stats {
  int64_t gpu_units, cpu_units;  /* tiles (phase 1) or merged partitions (phase 2) */
  int64_t gpu_rows, cpu_rows;    /* input rows of those units */
  double wall_s;                 /* the call, end to end */
  double cpu_rows_per_s;         /* measured per CPU thread (planner input) */
  double gpu_rows_per_s;         /* measured over the GPU's units */
} tga_coprocess_stats;
/* Split of the last compat call of that phase (1 or 2) on this context. */
tga_status tga_last_coprocess(tga_context *ctx, int32_t phase, tga_coprocess_stats *st);

/* ---- primitives and utilities --------------------------------------------- */
/* Exclusive prefix sum of n u32 on the device (own decoupled look-back). */
tga_status tga_scan_exclusive_u32(tga_context *ctx, const uint32_t *d_in, uint32_t *d_out, int64_t n);

typedef enum tga_gen_kind {
  TGA_GEN_KEYS_U32 = 1,   /* a = distinct */
  TGA_GEN_KEYS_U64 = 2,   /* a = distinct */
  TGA_GEN_VALS_I32 = 3,   /* a = lo (as int64 bits), b = hi - lo */
  TGA_GEN_ZIPF_U64 = 4,   /* a = distinct, x = exponent s */
  TGA_GEN_NORMAL_F64 = 5, /* x = mean, y = sigma */
  TGA_GEN_REF_U32 = 6,    /* u32 keys of the reference's workload (test/test.cu:96-129, data_generator.hpp:7-85):
                             a = cardinality (exactly that many distinct keys), b = 0 uniform / 1 Zipf(theta < 1) /
                             2 self-similar(h) draws of the repeated keys, x = theta or h; rows shuffled */
  TGA_GEN_MOD_U32 = 7,    /* closed-form keys: fmix32((i mod a) ^ salt), a = distinct */
  TGA_GEN_MOD_U64 = 8,    /* closed-form keys: SplitMix64 finalizer of (i mod a) ^ salt, a = distinct */
  TGA_GEN_SEQ_I32 = 9     /* values a + i (low 32 bits): with the MOD keys, group k has
                             c = floor((n-1-k)/D) + 1 rows, SUM = c(a+k) + D c(c-1)/2,
                             MIN = a+k, MAX = a+k+D(c-1) */
} tga_gen_kind;
/* Fills a device buffer with seeded counter-based synthetic data. */
tga_status tga_generate(tga_context *ctx, int32_t kind, void *d_out, int64_t n, uint64_t a, uint64_t b, double x,
                        double y, uint64_t seed, uint64_t salt);

tga_status tga_device_alloc(tga_context *ctx, uint64_t bytes, void **out);
/* Pinned host memory (free with tga_host_free): host columns in it are
 * streamed with the copies overlapping the aggregation. */
tga_status tga_host_alloc(tga_context *ctx, uint64_t bytes, void **out);
tga_status tga_device_free(tga_context *ctx, void *p);
tga_status tga_memcpy_h2d(tga_context *ctx, void *dst, const void *src, uint64_t bytes);
tga_status tga_memcpy_d2h(tga_context *ctx, void *dst, const void *src, uint64_t bytes);
/* Streaming device-to-device copy kernel (16-byte aligned, bytes % 16 == 0):
 * the bandwidth and counter-scaling reference for the benchmark. */
tga_status tga_device_copy(tga_context *ctx, void *dst, const void *src, uint64_t bytes);
tga_status tga_synchronize(tga_context *ctx);

/* Per-kernel timing with HIP events on the context stream. */
tga_status tga_profile_enable(tga_context *ctx, int32_t on);
tga_status tga_profile_reset(tga_context *ctx);
/* i-th kernel name seen since the last reset, its launch count and total ms. */
tga_status tga_profile_entry(tga_context *ctx, int32_t i, char *name, size_t name_len, int64_t *launches,
                             double *total_ms);

#ifdef __cplusplus
}
#endif

#endif /* TGA_GROUPBY_H_ */
