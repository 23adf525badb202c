// tga_host.hpp -- host-side runtime shared by the C ABI translation units:
// the context (device, stream, lock), the HBM caching arena, kernel timing,
// schema construction and the autotuner.
#pragma once

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/tga_groupby.h"
#include "tga_internal.hpp"

namespace tga {

// ---------------------------------------------------------------------------
// caching arena over hipMalloc with a byte budget
// ---------------------------------------------------------------------------
class Arena {
 public:
  explicit Arena(uint64_t budget) : budget_(budget) {}
  ~Arena() { release_cached(); }
  // Returns nullptr when the budget or the device is exhausted.
  void *alloc(uint64_t bytes);
  void free(void *p);
  void release_cached();
  uint64_t in_use() const { return in_use_; }
  uint64_t peak() const { return peak_; }
  uint64_t budget() const { return budget_; }
  static uint64_t round(uint64_t bytes);

 private:
  std::mutex mu_;
  uint64_t budget_;
  uint64_t in_use_ = 0, cached_ = 0, peak_ = 0;
  std::map<uint64_t, std::vector<void *>> free_;  // size class -> blocks
  std::map<void *, uint64_t> live_;
};

// Pinned host blocks, reused.  hipHostFree waits for the whole device, so
// the blocks of finished batches are kept for the next ones and freed with
// the context.  Power-of-two size classes from 4 KiB.
class PinnedPool {
 public:
  ~PinnedPool();
  void *get(uint64_t bytes);  // nullptr when the host is out of pinned memory
  void put(void *p, uint64_t bytes);

 private:
  static uint64_t cls(uint64_t bytes);
  std::mutex mu_;
  std::map<uint64_t, std::vector<void *>> free_;
};

// ---------------------------------------------------------------------------
// per-kernel HIP-event timing on the context stream
// ---------------------------------------------------------------------------
class Profiler {
 public:
  bool on = false;
  void begin(const char *name, hipStream_t st);
  void end(hipStream_t st);
  void resolve();  // after the stream was synchronised
  void reset();
  void destroy();
  std::vector<std::pair<std::string, std::pair<int64_t, double>>> entries() const;

 private:
  struct Pending { std::string name; hipEvent_t a, b; };
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  std::map<std::string, std::pair<int64_t, double>> acc_;
  std::vector<std::string> order_;
  hipEvent_t get();
};

}  // namespace tga

struct tga_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t props{};
  int cus = 0;
  size_t lds_max = 65536;  // dynamic LDS bytes one workgroup may use
  std::recursive_mutex mu;
  tga::Arena *arena = nullptr;
  bool owns_arena = true;  // false: a lane context using its parent's arena
  tga::PinnedPool *pinned = nullptr;
  uint32_t *status_host = nullptr;  // pinned, kStatusWords: a call's status words come back in one copy
  tga::Profiler prof;
  std::string err;
  tga_options last_plan{};
  // compat phases: extra GPU lanes (own stream, arena and status words, one
  // per compat stream beyond the first) and the CPU co-processing settings
  std::vector<tga_context *> lanes;
  int32_t cpu_threads = 0;
  double cpu_share = 0.0;
  tga_coprocess_stats cop[2]{};
};

namespace tga {

// Looks up, once per device, the device address of the no-alloc entry
// point's scratch array (the device must be current).
hipError_t noalloc_scratch_init(int device);

// Sets the context's error text (printf-style) and returns `st`.
tga_status fail(tga_context *ctx, tga_status st, const char *fmt, ...);

#define TGA_HIP(ctx, call)                                                                      \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return ::tga::fail((ctx), TGA_ERR_HIP, "%s: %s (%s:%d)", #call,       \
                                             hipGetErrorString(e_), __FILE__, __LINE__);        \
  } while (0)

// RAII device buffer from the context arena.
struct DBuf {
  tga_context *ctx = nullptr;
  void *p = nullptr;
  uint64_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept { *this = std::move(o); }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) { reset(); ctx = o.ctx; p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DBuf() { reset(); }
  void reset() { if (p && ctx) ctx->arena->free(p); p = nullptr; bytes = 0; }
  bool alloc(tga_context *c, uint64_t n) {
    reset();
    ctx = c;
    bytes = n ? n : 1;
    p = c->arena->alloc(bytes);
    return p != nullptr;
  }
  template <typename T> T *as() const { return (T *)p; }
};

// A validated query: schema + key info + output dtypes.
struct QuerySpec {
  Schema schema{};
  int key_bytes = 8;
  int key_dtype = DT_U64;
  int32_t out_dtype[kMaxAccs]{};
  int nfx = 0;
};

// Builds the accumulator schema from the aggregate list; returns an error
// text (empty on success).
std::string build_spec(const tga_column *key, const tga_column *values, int32_t n_values, const tga_agg *aggs,
                       int32_t n_aggs, QuerySpec &q);

// LDS layout of a table with `slots` slots for this schema.
Layout make_layout(const Schema &s, int key_bytes, uint32_t slots);
uint32_t lds_slot_bytes(const Schema &s, int key_bytes);

// two-pass partitioning: the second pass splits each first-pass partition
// into 2^kTwoPassSubBits (2^28 rows, 2^12 partitions: second pass 1.24 ms
// with 16 sub-partitions, 1.54 ms with 4).  Records of several value words
// take the widest first pass their LDS holds and split up to 2^kTwoPassMaxSubBits.
constexpr int kTwoPassSubBits = 4;
constexpr int kTwoPassMaxSubBits = 6;

struct Plan {
  int path = TGA_PATH_DIRECT;
  int pbits = 0;
  uint32_t table_slots = 0;
  uint32_t tile_rows = 8192;
  uint32_t workgroups = 0;
  uint32_t block = 256;
  uint64_t chunk_rows = 0;
  uint64_t gslots = 0;  // global table slots
  uint64_t est = 0;     // the group estimate the plan was made for
  Layout layout{};
  int nbits = 0;
  int max_probe = 0;
  int fold = 0;  // fold equal keys of a wave before the LDS atomics
  int scatter_layout = 1;  // 1 = partition-contiguous, 2 = per-workgroup buckets
  int scatter_wc = 0;      // bucket layout of records written as whole cache lines
  int scatter_p1_bits = 0; // > 0: two passes, the first into 2^scatter_p1_bits partitions
  int max_pbits = 12;      // the largest fanout this query's layouts reach
  int rec_vd = 0;          // partitioned copy as records of rec_vd value words (0: one array per column)
  // Adaptive fanout (adapt_bits > 0): the group count is still open (the
  // sample left it so), so the first pass partitions into 2^adapt_bits, the
  // hash-slice count is taken over that pass's partition 0, and the final
  // fanout (adapt_bits up to adapt_bits + kTwoPassMaxSubBits) follows from it.
  int adapt_bits = 0;
  uint64_t sample_d = 0;   // the sample's distinct keys (a lower bound)
  // re-planning: the options the plan came from, and whether the fanout was
  // the tuner's (so a too-small estimate may raise it)
  tga_options opt{};
  bool may_replan = false;
};

// Autotuner: plan from (rows, estimated groups, schema, device), honouring
// any explicit field of `opt`.
Plan autotune(const tga_context *ctx, const QuerySpec &q, uint64_t n_rows, uint64_t est_groups,
              const tga_options *opt);

// The plan as the options tga_last_plan reports.
void plan_to_options(const Plan &p, tga_options &o);

// Cardinality estimate from a random key sample (GEE estimator,
// Charikar et al. 2000, or the equal-frequency model), replaced by an exact
// count over a 2^-10 hash slice of all keys when the sample leaves the
// count open.  Needs the context lock.
// With `deferred` non-null, an unreliable sample over at least
// kSliceMinRows rows returns the sample's estimate with *deferred = true and
// *sample_d = the sample's distinct keys, and the caller counts the slice
// itself (resolve_deferred_estimate, or over the first pass of an adaptive
// partitioned plan).
tga_status estimate_groups(tga_context *ctx, const void *d_keys, int key_bytes, uint64_t n, uint64_t *est,
                           double *top_share = nullptr, bool *deferred = nullptr, uint64_t *sample_d = nullptr);
// Distinct-count estimate of n rows from a sample of m keys (sorted in place).
// *top_share (optional): the largest share of the sample one key takes.
// *reliable (optional): false when the sample, mostly singletons, leaves
// the group count open (see estimate_groups).
uint64_t estimate_from_sample(uint64_t *v, uint64_t m, uint64_t n, double *top_share = nullptr,
                              bool *reliable = nullptr);
// The same from the sample's counts: d distinct keys, f1 of them seen once,
// the most frequent seen top times.
uint64_t estimate_from_stats(uint64_t d, uint64_t f1, uint64_t top, uint64_t m, uint64_t n,
                             double *top_share = nullptr, bool *reliable = nullptr);
// Exact distinct count of the keys in one 2^-kSliceBits slice of the
// partition hash, over all n rows (*overflow: the set filled up, a lower
// bound).  Needs the context lock.
tga_status slice_count(tga_context *ctx, const void *d_keys, int key_bytes, uint64_t n, uint64_t *count,
                       bool *overflow);
// The estimate after the hash-slice count: 2^kSliceBits times the slice's
// distinct keys, at least the sample's distinct keys sample_d; the sample's
// estimate when the slice set overflowed.
uint64_t combine_slice_estimate(uint64_t sample_est, uint64_t sample_d, uint64_t slice_count, bool slice_overflow,
                                uint64_t n);
// An unreliable sample estimate over at least this many rows is replaced by
// an exact count over a hash slice of the keys.
constexpr uint64_t kSliceMinRows = 1ULL << 22;
// After a deferred estimate: the plan made from the sample's estimate becomes
// adaptive when its partitioned copy is records in bucket layout (the slice
// is then counted over the first pass); otherwise the slice is counted over
// all keys now and the plan made again.
tga_status resolve_deferred_estimate(tga_context *ctx, const QuerySpec &q, const void *d_keys, uint64_t n,
                                     uint64_t est, uint64_t sample_d, const tga_options *opt, Plan &plan);

inline uint64_t next_pow2_u64(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Host-resident columns of at least kStreamMinRows rows are streamed tile by
// tile through HBM (run_streamed) instead of staged whole.
bool should_stream(const tga_column *key, const tga_column *values, int nv, uint64_t n);
struct PlainOut;
tga_status run_streamed(tga_context *ctx, const struct QuerySpec &q, const tga_column *key, const tga_column *values,
                        int nv, uint64_t n, const tga_options *opt, PlainOut *out, void *ext_keys,
                        void *const *ext_cols, uint64_t ext_cap, uint64_t *groups);

// Host aggregation of u32 keys with u32 wrap-around sums (every key value
// legal): open addressing over a power-of-two table kept at most half full.
class CpuAgg32 {
 public:
  void reset(uint64_t expect);
  void add(const uint32_t *k, const uint32_t *v, uint64_t n);
  uint64_t groups() const { return groups_; }
  // the groups in table order; returns their count
  uint64_t emit(uint32_t *ok, uint32_t *ov) const;

 private:
  void grow();
  void add_one(uint32_t key, uint32_t val) {
    uint64_t s = (uint64_t)(key_hash((uint64_t)key) & mask_);
    for (;;) {
      if (!used_[s]) {
        used_[s] = 1;
        keys_[s] = key;
        sums_[s] = val;
        ++groups_;
        return;
      }
      if (keys_[s] == key) {
        sums_[s] += val;
        return;
      }
      s = (s + 1) & mask_;
    }
  }
  std::vector<uint32_t> keys_, sums_;
  std::vector<uint8_t> used_;
  uint64_t mask_ = 0, groups_ = 0;
};
// The compat partition of a distinct row (upper hash bits modulo P), as the
// GPU computes it.
uint32_t compat_partition_host(uint32_t key, uint32_t P);
// A phase-1 tile on the host: rows [0, m) of keys/vals pre-aggregated and
// written back partition by partition; offs[p] = start of partition p,
// offs[P] = the group count (returned).
uint64_t cpu_compat_tile(uint32_t *keys, uint32_t *vals, uint64_t m, uint32_t P, uint32_t *offs);
// A phase-2 merged partition on the host: its m gathered rows aggregated
// with the reference's chunk rule, the groups written over them.
tga_status cpu_compat_intra(uint32_t *k, uint32_t *v, uint64_t m, uint64_t min_load, uint64_t max_load,
                            uint64_t *groups);
// The context and nstreams - 1 lane contexts on its device (created once,
// kept by the context, at most 8 in all; they allocate from the context's
// arena).  Call with ctx->mu held.
std::vector<tga_context *> lanes_of(tga_context *ctx, size_t nstreams, tga_status *st);

// Output of one plain group-by execution (arena buffers owned by caller).
struct PlainOut {
  DBuf keys;
  DBuf cols[kMaxAccs];
  uint64_t groups = 0;
  uint64_t capacity = 0;
};

// Runs the plain operator on device-resident columns.  If `ext` is non-null
// the results go to those caller buffers with capacity ext_cap.
tga_status run_plain(tga_context *ctx, const QuerySpec &q, const void *d_keys, uint64_t n_rows, const Plan &plan_in,
                     PlainOut *out, void *ext_keys, void *const *ext_cols, uint64_t ext_cap, uint64_t *groups);

hipError_t gtable_alloc(tga_context *ctx, const QuerySpec &q, uint64_t S, GTable &g, std::vector<DBuf> &bufs,
                        DBuf &flags);
tga_status aggregate_into(tga_context *ctx, const QuerySpec &q, const void *d_keys, uint64_t n, const Plan &plan,
                          const GTable &g, std::vector<DBuf> &keep);
tga_status compact_into(tga_context *ctx, const QuerySpec &q, const GTable &g, const OutDesc &od, uint32_t *hflags,
                        uint64_t *hcount);

// Timed launch helper.
template <typename F>
hipError_t timed(tga_context *ctx, const char *name, F &&f) {
  if (ctx->prof.on) ctx->prof.begin(name, ctx->stream);
  hipError_t e = f();
  if (ctx->prof.on) ctx->prof.end(ctx->stream);
  return e;
}

}  // namespace tga

struct tga_result {
  tga_context *ctx = nullptr;
  tga::PlainOut out;
  int32_t key_dtype = 0;
  int32_t naggs = 0;
  int32_t dtype[tga::kMaxAccs]{};
};
