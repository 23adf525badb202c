// tga_capi.cpp -- C ABI of the plain operator: context, arena, autotuner,
// execution of the direct and partitioned paths, results and utilities.
//
// Reference host orchestration replaced (T/ = tiled_groupby_agg_partition/):
//   groupby_agg_partition / _thread   T/src/group_then_partition.cu:308-528
//   groupby_agg_intra_partition       T/src/groupby_agg_intra_partition.cu:159-411
//   pre_device_alloc, pre_alloc_device_memory  T/src/util.cu:4-40, intra:51-86
// Columns stay resident in HBM; one host thread drives one HIP stream with no
// host round trip between kernels (the reference blocks six times per tile,
// group_then_partition.cu:97,173,221,386,409,418).
#include <algorithm>
#include <cmath>
#include <unordered_map>

#include "tga_host.hpp"

namespace tga {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
tga_status fail(tga_context *ctx, tga_status st, const char *fmt, ...) {
  if (ctx) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return st;
}

// ---------------------------------------------------------------------------
// arena
// ---------------------------------------------------------------------------
uint64_t Arena::round(uint64_t bytes) {
  if (bytes <= 4096) return 4096;
  if (bytes <= (1ULL << 21)) {  // powers of two up to 2 MiB
    uint64_t c = 4096;
    while (c < bytes) c <<= 1;
    return c;
  }
  // above: multiples of 2 MiB, rounded up by at most 1/8
  uint64_t step = 1ULL << 21;
  while (step * 8 < bytes) step <<= 1;
  return (bytes + step - 1) / step * step;
}

void *Arena::alloc(uint64_t bytes) {
  const uint64_t c = round(bytes);
  std::lock_guard<std::mutex> g(mu_);
  // a cached block of this class, or of a class at most a quarter larger
  // (above 2 MiB classes are close together: a request that varies a little
  // from call to call still finds the last call's block)
  for (auto it = free_.lower_bound(c); it != free_.end() && it->first <= c + c / 4; ++it) {
    if (it->second.empty()) continue;
    const uint64_t cc = it->first;
    void *p = it->second.back();
    it->second.pop_back();
    cached_ -= cc;
    in_use_ += cc;
    live_[p] = cc;
    peak_ = std::max(peak_, in_use_);
    return p;
  }
  if (in_use_ + c > budget_) return nullptr;
  if (in_use_ + cached_ + c > budget_) {
    // drop cached blocks until the new one fits
    for (auto &kv : free_) {
      for (void *q : kv.second) { hipFree(q); cached_ -= kv.first; }
      kv.second.clear();
      if (in_use_ + cached_ + c <= budget_) break;
    }
  }
  void *p = nullptr;
  if (hipMalloc(&p, c) != hipSuccess) {
    (void)hipGetLastError();
    for (auto &kv : free_) { for (void *q : kv.second) hipFree(q); kv.second.clear(); }
    cached_ = 0;
    if (hipMalloc(&p, c) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
  }
  in_use_ += c;
  live_[p] = c;
  peak_ = std::max(peak_, in_use_);
  return p;
}

void Arena::free(void *p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  const uint64_t c = it->second;
  live_.erase(it);
  in_use_ -= c;
  cached_ += c;
  free_[c].push_back(p);
}

void Arena::release_cached() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto &kv : free_) { for (void *q : kv.second) hipFree(q); kv.second.clear(); }
  cached_ = 0;
}

uint64_t PinnedPool::cls(uint64_t bytes) {
  uint64_t c = 4096;
  while (c < bytes) c <<= 1;
  return c;
}

void *PinnedPool::get(uint64_t bytes) {
  const uint64_t c = cls(bytes);
  {
    std::lock_guard<std::mutex> g(mu_);
    auto it = free_.find(c);
    if (it != free_.end() && !it->second.empty()) {
      void *p = it->second.back();
      it->second.pop_back();
      return p;
    }
  }
  void *p = nullptr;
  if (hipHostMalloc(&p, c) != hipSuccess) { (void)hipGetLastError(); return nullptr; }
  return p;
}

void PinnedPool::put(void *p, uint64_t bytes) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  free_[cls(bytes)].push_back(p);
}

PinnedPool::~PinnedPool() {
  for (auto &kv : free_)
    for (void *p : kv.second) hipHostFree(p);
}

// ---------------------------------------------------------------------------
// profiler
// ---------------------------------------------------------------------------
hipEvent_t Profiler::get() {
  if (!pool_.empty()) { hipEvent_t e = pool_.back(); pool_.pop_back(); return e; }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}
void Profiler::begin(const char *name, hipStream_t st) {
  Pending p{name, get(), get()};
  hipEventRecord(p.a, st);
  pending_.push_back(p);
}
void Profiler::end(hipStream_t st) {
  if (!pending_.empty()) hipEventRecord(pending_.back().b, st);
}
void Profiler::resolve() {
  for (auto &p : pending_) {
    float ms = 0;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      auto it = acc_.find(p.name);
      if (it == acc_.end()) { order_.push_back(p.name); acc_[p.name] = {1, ms}; }
      else { it->second.first += 1; it->second.second += ms; }
    }
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
}
void Profiler::reset() { resolve(); acc_.clear(); order_.clear(); }
void Profiler::destroy() {
  resolve();
  for (hipEvent_t e : pool_) hipEventDestroy(e);
  pool_.clear();
}
std::vector<std::pair<std::string, std::pair<int64_t, double>>> Profiler::entries() const {
  std::vector<std::pair<std::string, std::pair<int64_t, double>>> v;
  for (auto &n : order_) v.push_back({n, acc_.at(n)});
  return v;
}

// ---------------------------------------------------------------------------
// schema
// ---------------------------------------------------------------------------
static bool valid_dtype(int d) { return d >= DT_U32 && d <= DT_F64; }

std::string build_spec(const tga_column *key, const tga_column *values, int32_t n_values, const tga_agg *aggs,
                       int32_t n_aggs, QuerySpec &q) {
  if (!key) return "key column is null";
  if (!(key->dtype == DT_U32 || key->dtype == DT_I32 || key->dtype == DT_U64 || key->dtype == DT_I64))
    return "key dtype must be u32, i32, u64 or i64";
  if (n_values < 0 || n_values > kMaxCols) return "at most 4 value columns";
  if (n_values > 0 && !values) return "value columns are null";
  if (n_aggs < 1 || n_aggs > kMaxAccs) return "between 1 and 8 aggregates";
  if (!aggs) return "aggregate list is null";
  q = QuerySpec();
  q.key_dtype = key->dtype;
  q.key_bytes = dtype_bytes(key->dtype);
  Schema &s = q.schema;
  s.ncols = n_values;
  for (int c = 0; c < n_values; ++c) {
    if (!valid_dtype(values[c].dtype)) return "bad value dtype";
    s.cols[c] = values[c].data;
    s.col_dtype[c] = values[c].dtype;
  }
  s.nacc = n_aggs;
  int nfx = 0;
  for (int a = 0; a < n_aggs; ++a) {
    const int op = aggs[a].op;
    int kind = -1;
    int col = aggs[a].column;
    if (op == OP_COUNT) { kind = ACC_COUNT; col = 0; q.out_dtype[a] = DT_I64; }
    else {
      if (col < 0 || col >= n_values) return "aggregate column out of range";
      const int dt = s.col_dtype[col];
      if (op == OP_SUM) {
        if (dt == DT_F64) {
          if (nfx >= kMaxFx) return "at most 2 fp64 SUM aggregates per query";
          kind = ACC_SUM_F;
          s.fx_slot[a] = nfx++;
          q.out_dtype[a] = DT_F64;
        } else { kind = ACC_SUM_I; q.out_dtype[a] = DT_I64; }
      } else if (op == OP_MIN || op == OP_MAX) {
        if (dt == DT_I32) kind = op == OP_MIN ? ACC_MIN_I32 : ACC_MAX_I32;
        else kind = op == OP_MIN ? ACC_MIN_I64 : ACC_MAX_I64;
        q.out_dtype[a] = dt;
      } else {
        return "unknown aggregate op";
      }
    }
    s.acc[a].kind = kind;
    s.acc[a].col = col;
    s.acc[a].col_dtype = op == OP_COUNT ? 0 : s.col_dtype[col];
  }
  q.nfx = nfx;
  s.nfx = nfx;
  return "";
}

uint32_t lds_slot_bytes(const Schema &s, int key_bytes) {
  uint32_t b = (uint32_t)key_bytes;
  for (int a = 0; a < s.nacc; ++a) b += (uint32_t)acc_lds_bytes(s.acc[a].kind);
  return b;
}

Layout make_layout(const Schema &s, int key_bytes, uint32_t slots) {
  Layout L{};
  L.slots = slots;
  // slots + 2 entries: index `slots` holds the key equal to the empty marker,
  // index slots + 1 is a discard entry that takes the atomics of rows with no
  // slot, so the update loop needs no per-lane branches
  const uint64_t n = (uint64_t)slots + 2;
  uint64_t off = 0;
  auto place = [&](uint64_t bytes) {
    const uint64_t o = off;
    off += (bytes + 15) & ~15ULL;
    return (uint32_t)o;
  };
  // 8-byte arrays first
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 0 && key_bytes == 8) L.key_off = place(n * 8);
    if (pass == 1 && key_bytes == 4) L.key_off = place(n * 4);
    for (int a = 0; a < s.nacc; ++a) {
      const int b = acc_lds_bytes(s.acc[a].kind);
      if ((pass == 0) == (b >= 8)) L.acc_off[a] = place(n * (uint64_t)b);
    }
  }
  L.bytes = (uint32_t)off;
  return L;
}

static uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}
static int ilog2(uint64_t x) {
  int b = 0;
  while ((1ULL << b) < x) ++b;
  return b;
}

// ---------------------------------------------------------------------------
// autotuner
//
// Picks the path, the LDS table size, the partition fanout, tile and grid
// from (rows, estimated groups) and the device; every explicit option wins.
// ---------------------------------------------------------------------------
Plan autotune(const tga_context *ctx, const QuerySpec &q, uint64_t n, uint64_t est, const tga_options *opt) {
  tga_options o{};
  if (opt) o = *opt;
  Plan p;
  const uint32_t sb = lds_slot_bytes(q.schema, q.key_bytes);
  const size_t lds_max = ctx->lds_max - 1024;
  if (est < 1) est = 1;
  if (est > n && n > 0) est = n;
  p.est = est;
  // largest power-of-two table fitting one workgroup's LDS
  uint32_t tmax = 64;
  while ((uint64_t)(tmax * 2 + 2) * sb + 256 <= lds_max) tmax *= 2;

  // Records of the partitioned copy: (key, 4-byte value) for one 4-byte
  // column; otherwise the key and up to kMaxRecValueDwords value words,
  // written by the write-combining scatter only (one or two passes) and not
  // with wave folding.
  const int vd = copy_as_records(q.schema) ? 1 : o.wave_fold == 1 ? 0 : record_value_dwords(q.schema);
  int path = o.path;
  if (path == TGA_PATH_AUTO) path = (est * 2 <= tmax) ? TGA_PATH_DIRECT : TGA_PATH_PARTITIONED;
  p.path = path;

  if (path == TGA_PATH_DIRECT) {
    uint32_t T = o.table_slots > 0 ? (uint32_t)o.table_slots : (uint32_t)std::min<uint64_t>(tmax, std::max<uint64_t>(256, next_pow2(est * 2)));
    p.table_slots = T;
    // One 1024-thread workgroup per CU: every workgroup flushes its whole
    // table into the global one with device atomics, all onto the same few
    // slots when there are few groups, so the flush count (workgroups x
    // groups) is kept low.  2^24 rows over 2^10 groups: 1024 workgroups of
    // 256 threads 0.32 ms, 256 of 256 0.21 ms, 256 of 1024 0.13 ms, 64 of
    // 1024 0.24 ms.
    p.block = o.block_threads > 0 ? (uint32_t)o.block_threads : 1024;
    uint32_t W = o.workgroups > 0 ? (uint32_t)o.workgroups : (uint32_t)ctx->cus;
    // at least ~16K rows per workgroup so each flush is amortised
    const uint64_t min_rows = 16384;
    W = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(W, (n + min_rows - 1) / min_rows));
    p.workgroups = W;
    p.chunk_rows = o.chunk_rows > 0 ? (uint64_t)o.chunk_rows : std::max<uint64_t>(1, (n + W - 1) / W);
  } else {
    // the largest table one workgroup's LDS holds (any slot count: the slot
    // is a multiply-shift of the hash), filled to at most 1/2 on average:
    // the first 4-slot bucket then holds nearly every key, so almost no row
    // leaves the branch-free lookup (at 3/4 fill the aggregation runs ~4x
    // slower on the 2^28-row, 2^20-group workload).  The fanout is a power
    // of two, so the fill lands between 1/4 and 1/2.
    uint32_t tfull = (uint32_t)std::min<uint64_t>(65535, (lds_max - 512) / sb - 2);
    tfull &= ~63u;
    uint32_t T = o.table_slots > 0 ? (uint32_t)o.table_slots : std::max<uint32_t>(tfull, 256);
    p.table_slots = T;
    const bool recs = o.scatter_layout != 1 && o.write_combine != 2 && vd > 0;
    int pbits = o.partition_bits;
    if (pbits <= 0) {
      const uint64_t P = next_pow2((est * 2 + (uint64_t)T - 1) / (uint64_t)T);
      pbits = std::max(1, ilog2(P));
      // Tables a quarter full where the one-pass write-combining scatter
      // still reaches that fanout (2^28 rows of u32 keys over 2^20 groups:
      // aggregation 1.00 ms at 2^8 partitions, 40% full, 0.63 ms at 2^9).
      // Past it, the second pass costs more than the fuller tables (2^21
      // groups: 2.25 ms at 2^9, 2.79 ms at 2^10), and among two-pass fanouts
      // a quarter is not reliably faster than a half.
      if (recs) {
        const int b4 = std::max(1, ilog2(next_pow2((est * 4 + (uint64_t)T - 1) / (uint64_t)T)));
        if (b4 > pbits && scatter_wc_tile(q.key_bytes, vd, b4, lds_max)) pbits = b4;
        // At least a partition per CU: the aggregation gives each partition
        // one workgroup, so the fanout the groups need alone can leave most
        // CUs idle (2^24 rows over 2^16 groups: 2^6 partitions, the
        // aggregation on 64 of 256 CUs).  Only while every partition keeps
        // 16K rows or more and one write-combining pass reaches the fanout.
        int bw = std::min(ilog2(next_pow2((uint64_t)ctx->cus)), ilog2(std::max<uint64_t>(1, n >> 14)));
        while (bw > pbits && !scatter_wc_tile(q.key_bytes, vd, bw, lds_max)) --bw;
        pbits = std::max(pbits, bw);
      }
    }
    // 2^12 partitions at most, 2^13 where two passes of records reach them
    // (2^28 rows over 26.8M groups: 2^12 partitions fill their tables to
    // 65%, and the aggregation takes 2.7 ms against 1.2 at 2^13), 2^14 for
    // wider records (larger slots: smaller tables)
    if (vd == 1) p.max_pbits = recs && scatter_wc_tile(q.key_bytes, 1, 13 - kTwoPassSubBits, lds_max) ? 13 : 12;
    else p.max_pbits = recs && scatter_wc_tile(q.key_bytes, vd, 14 - kTwoPassMaxSubBits, lds_max) ? 14 : 12;
    pbits = std::min(pbits, p.max_pbits);
    p.pbits = pbits;
    uint32_t tile = o.tile_rows > 0 ? (uint32_t)o.tile_rows : 8192;
    // the scatter tile must fit the LDS with the 2^pbits histogram
    while (tile > 1024 && scatter_lds_bytes(q.schema, q.key_bytes, pbits, tile) > lds_max) tile /= 2;
    p.tile_rows = tile;
    p.block = o.block_threads > 0 ? (uint32_t)o.block_threads : 1024;
    const uint64_t lds = make_layout(q.schema, q.key_bytes, T).bytes;
    const uint32_t per_cu = (uint32_t)std::max<uint64_t>(1, ctx->lds_max / std::max<uint64_t>(lds, 1));
    p.workgroups = o.workgroups > 0 ? (uint32_t)o.workgroups : (uint32_t)ctx->cus * std::min<uint32_t>(per_cu, 2);
    const uint64_t P = 1ULL << pbits;
    p.chunk_rows = o.chunk_rows > 0 ? (uint64_t)o.chunk_rows : std::max<uint64_t>(65536, 2 * ((n + P - 1) / P));
  }
  p.table_slots = std::max<uint32_t>(4, p.table_slots & ~3u);  // whole 4-slot buckets
  p.layout = make_layout(q.schema, q.key_bytes, p.table_slots);
  p.nbits = ilog2((uint64_t)p.table_slots + 1);
  p.max_probe = (int)std::min<uint32_t>(p.table_slots / 4, 64);  // in buckets
  // Wave folding (one ballot per slot bit, then a serial fold into the
  // lowest lane) is opt-in: on MI355X the LDS serialises same-address atomics
  // of one wave faster than the fold costs, even when every row has the same
  // key (2^26 rows, one key: 0.52 ms without, 6.9 ms with; Zipf(1.1) keys:
  // 2x slower with).
  p.fold = o.wave_fold == 1 ? 1 : 0;
  // Per-workgroup buckets by default: no histogram pass over the keys.  The
  // partition-contiguous copy (histogram + column scan first) makes the
  // aggregation ~12% faster (no partially filled buckets) but costs more in
  // the extra pass (2^28 rows, 2^20 groups on MI355X: 4.24 vs 4.68 ms).
  p.scatter_layout = o.scatter_layout == 1 ? 1 : 2;
  // Write combining (bucket layout, one 4-byte value column): each
  // workgroup keeps every partition's unfinished cache line in LDS, so the
  // scatter writes whole lines only.  Its carry buffer (128 B per partition)
  // takes LDS from the tile, which the tuner shrinks to 6144 or 4096 rows.
  if (p.path == TGA_PATH_PARTITIONED && p.scatter_layout == 2 && o.write_combine != 2 && vd > 0) {
    const uint32_t wt = scatter_wc_tile(q.key_bytes, vd, p.pbits, lds_max);
    if (wt && (o.tile_rows <= 0 || (uint32_t)o.tile_rows == wt)) {
      p.tile_rows = wt;
      p.scatter_wc = 1;
    } else if (!wt && vd == 1 && p.pbits - kTwoPassSubBits >= 1) {
      // A fanout whose carry buffer does not fit: two passes, written
      // whole-line, the first into 2^(pbits - 4) partitions, the second
      // splitting each by the next 4 hash bits.  Past 2^9 partitions the
      // one-pass scatter's tiles hold under 16 rows per partition and nearly
      // every line it writes is partial (2^28 rows of 8-byte records at 2^12
      // partitions: 7.7 ms in one pass, 1.0 + 1.2 ms in two).
      const int b1 = p.pbits - kTwoPassSubBits;
      const uint32_t t1 = scatter_wc_tile(q.key_bytes, 1, b1, lds_max);
      if (t1) {
        p.tile_rows = t1;
        p.scatter_wc = 1;
        p.scatter_p1_bits = b1;
      }
    } else if (!wt && vd > 1) {
      // wider records: the widest first pass whose tile fits, then at most
      // 2^kTwoPassMaxSubBits ways
      for (int b1 = std::min(p.pbits - 1, 9); b1 >= 1 && p.pbits - b1 <= kTwoPassMaxSubBits; --b1) {
        const uint32_t t1 = scatter_wc_tile(q.key_bytes, vd, b1, lds_max);
        if (t1 && scatter_wc_tile(q.key_bytes, vd, p.pbits - b1, lds_max)) {
          p.tile_rows = t1;
          p.scatter_wc = 1;
          p.scatter_p1_bits = b1;
          break;
        }
      }
    }
  }
  if (p.path == TGA_PATH_PARTITIONED) p.rec_vd = vd == 1 ? 1 : p.scatter_wc ? vd : 0;
  p.opt = o;
  p.may_replan = p.path == TGA_PATH_PARTITIONED && o.partition_bits <= 0 && p.pbits < p.max_pbits;
  // global table at load <= 1/2 of the estimate, at most 2 n slots
  uint64_t gs = next_pow2(std::max<uint64_t>(1024, est * 2));
  gs = std::min<uint64_t>(gs, next_pow2(std::max<uint64_t>(1024, 2 * n)));
  p.gslots = gs;
  return p;
}

void plan_to_options(const Plan &p, tga_options &o) {
  o = tga_options{};
  o.path = p.path;
  o.partition_bits = p.pbits;
  o.table_slots = (int32_t)p.table_slots;
  o.tile_rows = (int32_t)p.tile_rows;
  o.workgroups = (int32_t)p.workgroups;
  o.block_threads = (int32_t)p.block;
  o.wave_fold = p.fold ? 1 : 2;
  o.chunk_rows = (int64_t)p.chunk_rows;
  o.est_groups = (int64_t)(p.gslots / 2);
  o.scatter_layout = p.scatter_layout;
  o.write_combine = p.scatter_wc ? 1 : 2;
  o.record_words = p.rec_vd;
}

// ---------------------------------------------------------------------------
// cardinality estimate
// ---------------------------------------------------------------------------
uint64_t estimate_from_sample(uint64_t *v, uint64_t m, uint64_t n, double *top_share, bool *reliable) {
  if (top_share) *top_share = 0.0;
  if (m == 0 || n == 0) return 1;
  std::sort(v, v + m);
  uint64_t f1 = 0, d = 0, top = 0;
  for (uint64_t i = 0; i < m;) {
    uint64_t j = i;
    while (j < m && v[j] == v[i]) ++j;
    if (j - i == 1) ++f1;
    top = std::max<uint64_t>(top, j - i);
    ++d;
    i = j;
  }
  return estimate_from_stats(d, f1, top, m, n, top_share, reliable);
}

uint64_t estimate_from_stats(uint64_t d, uint64_t f1, uint64_t top, uint64_t m, uint64_t n, double *top_share,
                             bool *reliable) {
  if (top_share) *top_share = 0.0;
  if (reliable) *reliable = true;
  if (m == 0 || n == 0) return 1;
  if (top_share) *top_share = (double)top / (double)m;
  if (m >= n) return d;  // the sample is the whole input
  // GEE (Charikar et al. 2000): sqrt(n/m) f1 + (d - f1); good under skew
  const double gee = std::sqrt((double)n / (double)m) * (double)f1 + (double)(d - f1);
  // equal-frequency model: a key of D equally frequent ones is a singleton
  // of the sample with probability ~exp(-m/D), so f1/m = exp(-m/D)
  double uni = (double)n;
  if (f1 < m) uni = -(double)m / std::log((double)f1 / (double)m + 1e-300);
  if (f1 == 0) uni = (double)d;
  uni = std::min(uni, (double)n);
  // Without heavy hitters in the sample (its largest count is what sampling
  // D equally frequent keys gives), the equal-frequency model is the right
  // one: there GEE returns ~sqrt(n m) whatever D is (2x too high for 2^20
  // keys in 2^28 rows).  With heavy hitters (skewed keys) take the larger of
  // the two; an underestimate is caught by the overflow replan anyway.
  const double mean = (double)m / std::max(uni, 1.0);
  const bool flat = (double)top <= 8.0 + 4.0 * mean;
  const double e = flat ? uni : std::max(gee, uni);
  // The sample bounds D only by d + (f1/m) n: every row of the mass it has
  // not seen may be a key of its own, and no statistic of the sample tells
  // how far that tail reaches.  Zipf(0.9) over 26.8M and over 1.6M keys give
  // nearly the same sample (GEE ~1.6M for both); self-similar(0.9) draws put
  // 90% of the rows on 10% of the keys, and their sample is flat and
  // matches the equal-frequency model in f1 and f2 as well as 2^20 uniform
  // keys do, at 8x too few keys.  The estimate is taken as it is only when
  // that bound is within 2x of it.
  const double hi = (double)d + (double)f1 / (double)m * (double)n;
  if (reliable) *reliable = hi <= 2.0 * std::max<double>((double)d, e);
  return (uint64_t)std::min<double>((double)n, std::max<double>((double)d, e));
}

tga_status estimate_groups(tga_context *ctx, const void *d_keys, int key_bytes, uint64_t n, uint64_t *est,
                           double *top_share, bool *deferred, uint64_t *sample_d_out) {
  if (top_share) *top_share = 0.0;
  if (deferred) *deferred = false;
  if (n == 0) { *est = 1; return TGA_OK; }
  // the sample is counted on the GPU; only its counts come back
  const uint64_t m = std::min<uint64_t>(n, 16384);
  DBuf buf;
  if (!buf.alloc(ctx, m * 8 + 64)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "sample buffer");
  uint64_t *smp = buf.as<uint64_t>() + 8;
  TGA_HIP(ctx, launch_sample_stats(key_bytes, d_keys, n, (uint32_t)m, smp, buf.as<uint64_t>(), ctx->stream));
  // read back through the context's pinned status block (a pageable
  // destination costs a staged copy on this wait)
  uint64_t h[4];
  TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host, buf.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(h, ctx->status_host, sizeof h);
  bool reliable = true;
  uint64_t sample_d = h[0];  // distinct keys seen: a lower bound on the group count
  if (!h[3]) {
    *est = estimate_from_stats(h[0], h[1], h[2], m, n, top_share, &reliable);
  } else {
    // keys that crowd one quarter of the counting table: count the sample
    // on the host instead
    std::vector<uint64_t> v(m);
    TGA_HIP(ctx, hipMemcpyAsync(v.data(), smp, m * 8, hipMemcpyDeviceToHost, ctx->stream));
    TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *est = estimate_from_sample(v.data(), m, n, top_share, &reliable);
    sample_d = 0;
    for (uint64_t i = 0; i < m; ++i) sample_d += (i == 0 || v[i] != v[i - 1]);  // v is sorted now
  }
  if (reliable || n < kSliceMinRows) return TGA_OK;
  if (deferred) {
    *deferred = true;
    if (sample_d_out) *sample_d_out = sample_d;
    return TGA_OK;
  }
  // Otherwise count the distinct keys of one 2^-10 slice of the partition
  // hash exactly, over all rows (one read of the keys): whatever the
  // frequencies, 2^10 times that count is within ~(D/2^10)^-1/2 of D.
  // 2^28 rows of Zipf(0.9) keys over 26.8M groups: the sample says 1.6M,
  // and the plan for it is abandoned and run again (7.5 ms instead of 5.3).
  uint64_t count = 0;
  bool overflow = false;
  if (tga_status e = slice_count(ctx, d_keys, key_bytes, n, &count, &overflow)) return e;
  *est = combine_slice_estimate(*est, sample_d, count, overflow, n);
  return TGA_OK;
}

// The widest first-pass fanout of an adaptive plan: the write-combining
// scatter's largest (at most 2^9) for this record width, where the second
// pass may still reach the plan's maximum fanout.
static int adapt_first_bits(const tga_context *ctx, const QuerySpec &q, const Plan &p) {
  const size_t lds_max = ctx->lds_max - 1024;
  for (int b = 9; b >= 1; --b)
    if (scatter_wc_tile(q.key_bytes, p.rec_vd, b, lds_max) && b + kTwoPassMaxSubBits >= p.max_pbits) return b;
  return 0;
}

tga_status resolve_deferred_estimate(tga_context *ctx, const QuerySpec &q, const void *d_keys, uint64_t n,
                                     uint64_t est, uint64_t sample_d, const tga_options *opt, Plan &plan) {
  const bool forced = opt && (opt->partition_bits > 0 || opt->tile_rows > 0 || opt->path != TGA_PATH_AUTO);
  if (!forced && plan.path == TGA_PATH_PARTITIONED && plan.scatter_layout == 2 && plan.scatter_wc &&
      plan.rec_vd > 0 && n < (1ULL << 31)) {
    const int b1 = adapt_first_bits(ctx, q, plan);
    if (b1 > 0 && b1 <= kSliceBits) {
      plan.adapt_bits = b1;
      plan.sample_d = sample_d;
      return TGA_OK;
    }
  }
  uint64_t c = 0;
  bool ovf = false;
  if (tga_status e = slice_count(ctx, d_keys, q.key_bytes, n, &c, &ovf)) return e;
  plan = autotune(ctx, q, n, combine_slice_estimate(est, sample_d, c, ovf, n), opt);
  return TGA_OK;
}

tga_status slice_count(tga_context *ctx, const void *d_keys, int key_bytes, uint64_t n, uint64_t *count,
                       bool *overflow) {
  const uint32_t slots = slice_set_slots(n);
  DBuf set, buf;
  if (!set.alloc(ctx, (uint64_t)slots * key_bytes) || !buf.alloc(ctx, 64))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "slice set");
  TGA_HIP(ctx, timed(ctx, "slice_distinct", [&] {
    return launch_slice_distinct(key_bytes, d_keys, n, set.p, slots, buf.as<uint64_t>(), (unsigned)ctx->cus * 2,
                                 ctx->stream);
  }));
  uint64_t c[2];
  TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host, buf.p, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(c, ctx->status_host, sizeof c);
  *count = c[0];
  *overflow = c[1] != 0;
  return TGA_OK;
}

uint64_t combine_slice_estimate(uint64_t sample_est, uint64_t sample_d, uint64_t slice_count, bool slice_overflow,
                                uint64_t n) {
  // An overflowed slice set counted only part of its slice: keep the
  // sample's estimate.  Otherwise the scaled slice count replaces it, but
  // never below the distinct keys the sample saw (a hot key with a short
  // tail may leave the slice empty).
  if (slice_overflow) return std::max<uint64_t>(1, std::max(sample_est, sample_d));
  const uint64_t scaled = std::min<uint64_t>(n, slice_count << kSliceBits);
  return std::max<uint64_t>(1, std::max(scaled, std::min<uint64_t>(n, sample_d)));
}

// ---------------------------------------------------------------------------
// plain operator execution
// ---------------------------------------------------------------------------
hipError_t gtable_alloc(tga_context *ctx, const QuerySpec &q, uint64_t S, GTable &g, std::vector<DBuf> &bufs,
                               DBuf &flags) {
  bufs.clear();
  bufs.resize(q.schema.nacc + 1);
  const uint64_t n = S + 1;
  if (!bufs[0].alloc(ctx, n * q.key_bytes)) return hipErrorOutOfMemory;
  g.keys = bufs[0].p;
  for (int a = 0; a < q.schema.nacc; ++a) {
    if (!bufs[a + 1].alloc(ctx, n * (uint64_t)acc_global_bytes(q.schema.acc[a].kind))) return hipErrorOutOfMemory;
    g.acc[a] = bufs[a + 1].p;
  }
  if (!flags.alloc(ctx, 64)) return hipErrorOutOfMemory;
  g.flags = flags.as<uint32_t>();
  g.S = S;
  return hipSuccess;
}

static uint64_t out_elem_bytes(int dt) { return (uint64_t)dtype_bytes(dt); }

// Persistent scatter grid: as many workgroups as the LDS lets run at once,
// with at least 8 tiles each.  The plain bucket scatter uses three quarters
// of that: its scattered writes saturate each XCD's write path with fewer
// workgroups, and every workgroup leaves one partly filled bucket per
// partition for the aggregation (2^28 rows, 2^20 groups on MI355X: scatter
// 2.46 ms with 192 or 256 workgroups, aggregation 1.52 vs 1.62 ms).  The
// write-combining scatter is bound by its LDS work, not by its writes, and
// uses every CU (1.76 ms with 192 workgroups, 1.43 ms with 256).
static unsigned scatter_grid(const tga_context *ctx, const QuerySpec &q, const Plan &plan, uint64_t m) {
  const size_t lds = plan.scatter_wc ? scatter_wc_lds_bytes(q.key_bytes, plan.rec_vd, plan.pbits, plan.tile_rows)
                                     : scatter_lds_bytes(q.schema, q.key_bytes, plan.pbits, plan.tile_rows);
  const size_t per_cu = std::max<size_t>(1, std::min<size_t>(4, ctx->lds_max / std::max<size_t>(lds, 1)));
  const uint32_t ntiles = (uint32_t)((m + plan.tile_rows - 1) / plan.tile_rows);
  size_t max_grid = ctx->cus * per_cu;
  if (plan.scatter_layout == 2 && !plan.scatter_wc) max_grid = std::max<size_t>(8, max_grid * 3 / 4);
  return bucket_grid(ntiles, (unsigned)max_grid);
}

// Plan buffer layout of the partitioned path: pstart[P+1] u64 | cstart[P+1]
// u32 | ticket u32 | split_rows u64 (8-byte aligned) | global-table slots
// u64.  gs_cap > 0: the plan kernel sizes the global table (AggPlan::gslots).
static AggPlan part_plan(const DBuf &buf, uint32_t P, uint64_t gs_cap = 0, uint32_t gs_per_split = 0) {
  char *b = buf.as<char>();
  AggPlan ap{};
  ap.pstart = (uint64_t *)b;
  ap.cstart = (uint32_t *)(b + (uint64_t)(P + 1) * 8);
  ap.P = P;
  ap.ticket = ap.cstart + (P + 1);
  ap.split_rows = (uint64_t *)(b + (((uint64_t)(P + 1) * 12 + 4 + 7) & ~7ULL));
  if (gs_cap) {
    ap.gslots = ap.split_rows + 1;
    ap.gs_cap = gs_cap;
    ap.gs_per_split = gs_per_split;
  }
  return ap;
}

// Buffers of one partitioning pass; they must outlive the work queued on
// the stream.
struct PartBufs {
  DBuf heads, meta, pkeys, planbuf;  // bucket layout: headers/lists, counters
  DBuf cnt, offs, total, pbase, desc;  // partition-contiguous layout
  DBuf pcols[kMaxCols];
  BucketBufs bb{};
  bool buckets = false;
  int rec = 0;  // > 0: pkeys holds records of `rec` value words (Plan::rec_vd), no pcols
  // two-pass partitioning: the first pass's buffers, and its bucket pool
  // (checked against its capacity once the call is done)
  DBuf heads1, meta1, pkeys1, plan1, items1;
  uint32_t *pool1 = nullptr;
  uint32_t cap1 = 0;
  BucketBufs bb1{};
  AggPlan ap1{};
  // adaptive plans: the first pass is done (pass1_final: it is the final
  // copy, its lists and plan built for the aggregation with gs_cap1)
  bool pass1_done = false, pass1_final = false;
  uint64_t gs_cap1 = 0;
  AggPlan ap{};
  Schema sp{};  // the schema over the partitioned copy
  // > 0: the plan kernel sizes the global table, up to gs_cap slots
  uint64_t gs_cap = 0;
  uint32_t gs_per_split = 0;
  void keep(std::vector<DBuf> &k) {
    for (DBuf *b : {&heads, &meta, &pkeys, &planbuf, &cnt, &offs, &total, &pbase, &desc, &heads1, &meta1, &pkeys1,
                    &plan1, &items1})
      if (b->p) k.push_back(std::move(*b));
    for (auto &c : pcols)
      if (c.p) k.push_back(std::move(c));
  }
};

// (b) + (c) for rows [0, m) of keys / s.cols: the partitioned copy and the
// work-item plan of k_agg.  Partition-contiguous layout: per-tile histogram,
// column scan, scatter to the scanned offsets.  Bucket layout: count-free
// scatter into buckets, then the per-partition bucket lists.
// Partly filled buckets are paired for k_agg only for 8-byte keys: its
// 4-byte-key kernels read one bucket per block (2^28 rows of u32 keys over
// 26.8M groups ran 2-4% slower with the paired form compiled in).
static bool pairable(int key_bytes) { return key_bytes == 8; }

// Bucket headers, lists and counters of a bucket pool of `cap` buckets over
// P partitions (counters zeroed on the stream).
static tga_status bucket_bufs(tga_context *ctx, uint64_t cap, uint32_t P, DBuf &heads, DBuf &meta_buf,
                              BucketBufs &bb) {
  if (cap >= (1ULL << 32)) return fail(ctx, TGA_ERR_UNSUPPORTED, "too many buckets (%llu)", (unsigned long long)cap);
  // meta: pool u32 (+ pad) | pbk [P] u32 | bcur [P] u32 | bend [P] u32 (+ pad) | prow [P] u64
  const uint64_t head_words = 1 + ((uint64_t)P * 12 + 7) / 8;
  const uint64_t meta_words = head_words + P;
  if (!heads.alloc(ctx, cap * 24) || !meta_buf.alloc(ctx, meta_words * 8))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
  bb.blist = heads.as<uint64_t>();
  bb.blist2 = bb.blist + cap;
  bb.bpart = (uint32_t *)(bb.blist2 + cap);
  bb.bcnt = bb.bpart + cap;
  uint64_t *meta = meta_buf.as<uint64_t>();
  bb.pool = (uint32_t *)meta;
  bb.pbk = (uint32_t *)(meta + 1);
  bb.bcur = bb.pbk + P;
  bb.bend = bb.bcur + P;
  bb.prow = meta + head_words;
  bb.cap = (uint32_t)cap;
  TGA_HIP(ctx, launch_fill_u64(meta, meta_words, 0, ctx->stream));
  return TGA_OK;
}

// Two-pass partitioning into 2^plan.pbits partitions (records, bucket
// layout): the first pass into 2^plan.scatter_p1_bits partitions in the
// buffers *1 of pb, then each of its partitions split again into the final
// pool of pb.bb.  Its list entries are cut into work items of 1024 entries
// (2^18 rows) below 2^8 first-pass partitions, 512 from there (2^28 rows:
// at 2^6 first-pass partitions 1.15 ms against 1.19 with 512; at 2^9, 1.23
// against 1.27 ms, and 1.53 against 1.71 ms under Zipf(0.9) keys, whose hot
// partitions spread over more items).
static uint32_t pass2_entries(uint32_t P1) { return P1 >= 256 ? 512 : 1024; }

// First pass of a bucket-layout partitioning of records into 2^b1
// partitions (write-combining scatter with plan.tile_rows-row tiles), and its
// bucket lists and work plan (items of chmax rows; gs_cap > 0: the plan
// kernel also sizes the global table) into the *1 buffers of pb.
static tga_status first_pass(tga_context *ctx, const QuerySpec &q, const void *keys, const Schema &s, uint64_t m,
                             const Plan &plan, int b1, unsigned grid, uint64_t chmax, uint64_t gs_cap,
                             uint32_t gs_per_split, PartBufs &pb) {
  hipStream_t st = ctx->stream;
  const uint32_t P1 = 1u << b1;
  const uint64_t row_bytes = (uint64_t)q.key_bytes + 4 * (uint64_t)plan.rec_vd;
  const uint64_t cap1 = bucket_cap(m, grid, b1, scatter_block(plan.tile_rows, 1));
  if (tga_status e = bucket_bufs(ctx, cap1, P1, pb.heads1, pb.meta1, pb.bb1)) return e;
  if (!pb.pkeys1.alloc(ctx, (cap1 << kBucketShift) * row_bytes) || !pb.plan1.alloc(ctx, (uint64_t)(P1 + 1) * 12 + 64))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
  pb.pool1 = pb.bb1.pool;
  pb.cap1 = (uint32_t)cap1;
  PartOut po{};
  TGA_HIP(ctx, timed(ctx, "part_scatter", [&] {
    return launch_part_scatter(q.key_bytes, keys, s, m, b1, plan.tile_rows, pb.pkeys1.p, po, &pb.bb1, nullptr, nullptr,
                               grid, st, 1);
  }));
  pb.ap1 = part_plan(pb.plan1, P1, gs_cap, gs_per_split);
  pb.ap1.blist = pb.bb1.blist;
  TGA_HIP(ctx, timed(ctx, "bucket_lists", [&] { return launch_bucket_lists(pb.bb1, P1, chmax, pb.ap1, st, false, false); }));
  return TGA_OK;
}

static tga_status partition_two_pass(tga_context *ctx, const QuerySpec &q, const void *keys, const Schema &s,
                                     uint64_t m, const Plan &plan, unsigned grid, PartBufs &pb, uint64_t &slots) {
  hipStream_t st = ctx->stream;
  const int b1 = plan.scatter_p1_bits, b2 = plan.pbits - b1;
  const uint32_t P1 = 1u << b1, P = 1u << plan.pbits;
  const int vd = plan.rec_vd;
  const uint64_t row_bytes = (uint64_t)q.key_bytes + 4 * (uint64_t)vd;
  // pass 1 (already done by an adaptive plan)
  if (!pb.pass1_done)
    if (tga_status e = first_pass(ctx, q, keys, s, m, plan, b1, grid, ~0ULL >> 1, 0, 0, pb)) return e;
  const uint64_t cap1 = pb.cap1;
  const AggPlan &ap1 = pb.ap1;
  if (!pb.items1.alloc(ctx, (uint64_t)(P1 + 2) * 4)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
  const uint32_t per_item = pass2_entries(P1);
  uint32_t *istart = pb.items1.as<uint32_t>(), *ticket = istart + P1 + 1;
  TGA_HIP(ctx, timed(ctx, "bucket_lists", [&] { return launch_pass2_plan(ap1.pstart, P1, per_item, istart, ticket, st); }));
  // pass 2: at most P1 + cap1 / per_item work items, each leaving up to
  // 2^b2 buckets open, and a reserve per wave
  const unsigned grid2 = (unsigned)(ctx->cus + 7) / 8 * 8;  // every CU, a multiple of the XCD count
  const uint64_t items_max = P1 + cap1 / per_item + 1;
  const uint64_t cap = (m + kBucketRows - 1) / kBucketRows + items_max * (1ULL << b2) + (uint64_t)grid2 * 1024 + 1;
  if (tga_status e = bucket_bufs(ctx, cap, P, pb.heads, pb.meta, pb.bb)) return e;
  slots = cap << kBucketShift;
  if (!pb.pkeys.alloc(ctx, slots * row_bytes)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
  TGA_HIP(ctx, timed(ctx, "part_scatter2", [&] {
    return launch_part_scatter2(q.key_bytes, vd, pb.pkeys1.p, pb.bb1.blist, ap1.pstart, istart, ticket, P1,
                                plan.pbits, per_item, b2, pb.pkeys.p, pb.bb, grid2, st);
  }));
  return TGA_OK;
}

static tga_status partition_rows(tga_context *ctx, const QuerySpec &q, const void *keys, const Schema &s, uint64_t m,
                                 const Plan &plan, PartBufs &pb) {
  hipStream_t st = ctx->stream;
  const uint32_t P = 1u << plan.pbits;
  const unsigned grid = scatter_grid(ctx, q, plan, m);
  pb.buckets = plan.scatter_layout == 2;
  pb.rec = plan.rec_vd;
  uint64_t slots = m;
  if (pb.pass1_final) {
    // an adaptive plan's first pass is the partitioned copy
    pb.pkeys = std::move(pb.pkeys1);
    pb.heads = std::move(pb.heads1);
    pb.meta = std::move(pb.meta1);
    pb.planbuf = std::move(pb.plan1);
    pb.bb = pb.bb1;
    pb.ap = pb.ap1;
    pb.ap.blist2 = pb.bb.blist2;
    const bool pair = pairable(q.key_bytes);
    TGA_HIP(ctx, timed(ctx, "bucket_pairs", [&] { return launch_bucket_pairs(pb.bb, P, pb.ap, st, pair); }));
    pb.pool1 = nullptr;
    pb.cap1 = 0;
    pb.sp = s;
    for (int c = 0; c < kMaxCols; ++c) pb.sp.cols[c] = nullptr;  // the values travel in the records
    return TGA_OK;
  }
  if (pb.buckets && plan.scatter_p1_bits > 0) {
    if (!pb.rec) return fail(ctx, TGA_ERR_INVALID_ARG, "two-pass partitioning needs a record copy");
    if (tga_status e = partition_two_pass(ctx, q, keys, s, m, plan, grid, pb, slots)) return e;
    if (!pb.planbuf.alloc(ctx, (uint64_t)(P + 1) * 12 + 64)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
    pb.sp = s;
    for (int c = 0; c < kMaxCols; ++c) pb.sp.cols[c] = nullptr;  // the values travel in the records
    pb.ap = part_plan(pb.planbuf, P, pb.gs_cap, pb.gs_per_split);
    pb.ap.blist = pb.bb.blist;
    pb.ap.blist2 = pb.bb.blist2;
    const bool pair = pairable(q.key_bytes);
    TGA_HIP(ctx, timed(ctx, "bucket_lists",
                       [&] { return launch_bucket_lists(pb.bb, P, plan.chunk_rows, pb.ap, st, true, pair); }));
    return TGA_OK;
  }
  if (pb.buckets) {
    const uint64_t cap = bucket_cap(m, grid, plan.pbits, scatter_block(plan.tile_rows, plan.scatter_wc));
    if (tga_status e = bucket_bufs(ctx, cap, P, pb.heads, pb.meta, pb.bb)) return e;
    slots = cap << kBucketShift;
  }
  const uint64_t row_bytes = (uint64_t)q.key_bytes + 4 * (uint64_t)pb.rec;
  if (!pb.pkeys.alloc(ctx, std::max<uint64_t>(slots, 1) * row_bytes) ||
      !pb.planbuf.alloc(ctx, (uint64_t)(P + 1) * 12 + 64))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
  PartOut po{};
  pb.sp = s;
  if (pb.rec)
    for (int c = 0; c < kMaxCols; ++c) pb.sp.cols[c] = nullptr;  // the values travel in the records
  for (int c = 0; c < s.ncols && !pb.rec; ++c) {
    if (!pb.pcols[c].alloc(ctx, std::max<uint64_t>(slots, 1) * (uint64_t)dtype_bytes(s.col_dtype[c])))
      return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
    po.cols[c] = pb.pcols[c].p;
    pb.sp.cols[c] = pb.pcols[c].p;
  }
  pb.ap = part_plan(pb.planbuf, P, pb.gs_cap, pb.gs_per_split);
  if (pb.buckets) {
    BucketBufs &bb = pb.bb;
    TGA_HIP(ctx, timed(ctx, "part_scatter", [&] {
      return launch_part_scatter(q.key_bytes, keys, s, m, plan.pbits, plan.tile_rows, pb.pkeys.p, po, &bb, nullptr,
                                 nullptr, grid, st, plan.scatter_wc);
    }));
    pb.ap.blist = bb.blist;
    pb.ap.blist2 = bb.blist2;
    const bool pair = pairable(q.key_bytes);
    TGA_HIP(ctx, timed(ctx, "bucket_lists",
                       [&] { return launch_bucket_lists(bb, P, plan.chunk_rows, pb.ap, st, true, pair); }));
    return TGA_OK;
  }
  const uint32_t ntiles = (uint32_t)((m + plan.tile_rows - 1) / plan.tile_rows);
  const uint64_t cells = (uint64_t)P * ntiles;
  const uint64_t dwords = colscan_desc_words(ntiles, P);
  if (!pb.cnt.alloc(ctx, cells * 4) || !pb.offs.alloc(ctx, cells * 4) || !pb.total.alloc(ctx, (uint64_t)P * 4) ||
      !pb.pbase.alloc(ctx, (uint64_t)(P + 1) * 4) || !pb.desc.alloc(ctx, (dwords + 1) * 8))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
  TGA_HIP(ctx, launch_fill_u64(pb.desc.as<uint64_t>(), dwords + 1, 0, st));
  TGA_HIP(ctx, timed(ctx, "part_count", [&] {
    return launch_part_count(q.key_bytes, keys, m, plan.pbits, plan.tile_rows, ntiles, pb.cnt.as<uint32_t>(), st);
  }));
  TGA_HIP(ctx, timed(ctx, "scan", [&] {
    return launch_colscan(pb.cnt.as<uint32_t>(), ntiles, P, pb.desc.as<uint64_t>(),
                          (uint32_t *)(pb.desc.as<uint64_t>() + dwords), pb.offs.as<uint32_t>(),
                          pb.total.as<uint32_t>(), pb.pbase.as<uint32_t>(), st);
  }));
  TGA_HIP(ctx, timed(ctx, "part_scatter", [&] {
    return launch_part_scatter(q.key_bytes, keys, s, m, plan.pbits, plan.tile_rows, pb.pkeys.p, po, nullptr,
                               pb.offs.as<uint32_t>(), pb.pbase.as<uint32_t>(), grid, st);
  }));
  TGA_HIP(ctx, timed(ctx, "plan_chunks", [&] {
    return launch_plan_chunks(pb.pbase.as<uint32_t>(), P, m, plan.chunk_rows, pb.ap, st);
  }));
  return TGA_OK;
}

// Aggregates rows [0, n) of device-resident columns into the global table g
// (direct or partitioned path).  Temporaries are appended to `keep` and must
// outlive the work queued on the context stream.
tga_status aggregate_into(tga_context *ctx, const QuerySpec &q, const void *d_keys, uint64_t n, const Plan &plan,
                          const GTable &g, std::vector<DBuf> &keep) {
  const Schema &s = q.schema;
  hipStream_t st = ctx->stream;
  DBuf planbuf;  // pstart[P+1] u64 | cstart[P+1] u32 | ticket
  if (n > 0 && plan.path == TGA_PATH_DIRECT) {
    if (!planbuf.alloc(ctx, 64)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "plan");
    AggPlan ap{planbuf.as<uint64_t>(), (uint32_t *)(planbuf.as<char>() + 16), 1, (uint32_t *)(planbuf.as<char>() + 32),
               (uint64_t *)(planbuf.as<char>() + 40)};
    const uint64_t nch = (n + plan.chunk_rows - 1) / plan.chunk_rows;
    TGA_HIP(ctx, launch_plan_direct(ap, n, (uint32_t)nch, st));
    const unsigned grid = (unsigned)std::min<uint64_t>(plan.workgroups, nch);
    TGA_HIP(ctx, timed(ctx, "agg_direct", [&] {
      return launch_agg(q.key_bytes, d_keys, s, ap, plan.layout, g, plan.nbits, plan.max_probe, grid, plan.block, st,
                        0, nullptr, nullptr, plan.fold);
    }));
  } else if (n > 0) {
    // partitioned path, in super-chunks of < 2^31 rows (u32 offsets)
    const uint64_t kSuper = 1ULL << 31;
    for (uint64_t r0 = 0; r0 < n; r0 += kSuper) {
      const uint64_t m = std::min<uint64_t>(kSuper, n - r0);
      Schema sc = s;
      for (int c = 0; c < s.ncols; ++c) sc.cols[c] = (const char *)s.cols[c] + r0 * (uint64_t)dtype_bytes(s.col_dtype[c]);
      PartBufs pb;
      {
        tga_status e = partition_rows(ctx, q, (const char *)d_keys + r0 * q.key_bytes, sc, m, plan, pb);
        if (e) return e;
      }
      TGA_HIP(ctx, timed(ctx, "agg_partitions", [&] {
        return launch_agg(q.key_bytes, pb.pkeys.p, pb.sp, pb.ap, plan.layout, g, plan.nbits, plan.max_probe,
                          plan.workgroups, plan.block, st, 0, nullptr, nullptr, plan.fold, nullptr, pb.rec);
      }));
      pb.keep(keep);
      if (r0 + kSuper < n) {  // reuse the buffers of this super-chunk
        TGA_HIP(ctx, hipStreamSynchronize(st));
        keep.clear();
      }
    }
  }

  if (planbuf.p) keep.push_back(std::move(planbuf));
  return TGA_OK;
}

// Compacts the occupied slots of g into od (slot order), then waits for the
// stream and returns the status words and the group count.
tga_status compact_into(tga_context *ctx, const QuerySpec &q, const GTable &g, const OutDesc &od, uint32_t *hflags,
                        uint64_t *hcount) {
  hipStream_t st = ctx->stream;
  const uint64_t ctiles = compact_tiles(g.S + 1);
  DBuf cdesc;  // look-back descriptors | ticket | count | gathered status words
  if (!cdesc.alloc(ctx, (ctiles + 2) * 8 + kStatusWords * 4)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "compaction");
  uint32_t *cticket = (uint32_t *)(cdesc.as<uint64_t>() + ctiles);
  uint64_t *count = cdesc.as<uint64_t>() + ctiles + 1;
  TGA_HIP(ctx, launch_fill_u64(cdesc.as<uint64_t>(), ctiles + 2, 0, st));
  TGA_HIP(ctx, timed(ctx, "compact", [&] {
    return launch_compact(q.key_bytes, g, q.schema, od, cdesc.as<uint64_t>(), cticket, count, st);
  }));
  // flags and group count gathered into one pinned copy
  WordGather wg{};
  wg.src[0] = g.flags;
  wg.n[0] = kFlagWords;
  wg.src[1] = (const uint32_t *)count;
  wg.n[1] = 2;
  wg.count = 2;
  uint32_t *dst = (uint32_t *)(cdesc.as<uint64_t>() + ctiles + 2);
  TGA_HIP(ctx, launch_gather_words(wg, dst, st));
  TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host, dst, (kFlagWords + 2) * 4, hipMemcpyDeviceToHost, st));
  TGA_HIP(ctx, hipStreamSynchronize(st));
  std::memcpy(hflags, ctx->status_host, sizeof(uint32_t) * kFlagWords);
  std::memcpy(hcount, ctx->status_host + kFlagWords, 8);
  return TGA_OK;
}

// Partitioned path with direct emission (n < 2^31): partitions that one work
// item holds whole write final rows straight to the output; only split
// partitions and LDS spills go through the global table, sized for them.
// One partitioned execution.  *overflow: the global table or the output was
// too small (retry with larger ones; *reduced: the table had fewer than
// plan.gslots slots, so a full-size one may do).  *replan_est > 0: partitions processed
// whole overflowed their LDS table and were abandoned; the value estimates
// the group count for a finer partitioning (retry with it).
// Adaptive plan (plan.adapt_bits > 0): the first pass into 2^b1 partitions
// with its bucket lists and work plan as a one-pass plan at 2^b1 would have
// them, then the exact distinct count of the keys in the 2^-kSliceBits hash
// slice, which lies inside its partition 0 (only that partition's records
// are read: 1/2^b1 of the rows, where the count over all keys reads every
// key), then the plan for the resulting estimate: the first pass is the
// partitioned copy if that plan's fanout is at most 2^b1, else a second pass
// splits it up to 2^kTwoPassMaxSubBits ways.  plan.adapt_bits is 0 after.
static tga_status adapt_first_pass(tga_context *ctx, const QuerySpec &q, const void *keys, uint64_t n, Plan &plan,
                                   bool full_table, PartBufs &pb) {
  hipStream_t st = ctx->stream;
  const int b1 = plan.adapt_bits, vd = plan.rec_vd;
  const uint32_t P1 = 1u << b1;
  const size_t lds_max = ctx->lds_max - 1024;
  Plan p1 = plan;  // the first pass as a one-pass plan at 2^b1
  p1.pbits = b1;
  p1.scatter_p1_bits = 0;
  p1.scatter_wc = 1;
  p1.tile_rows = scatter_wc_tile(q.key_bytes, vd, b1, lds_max);
  p1.chunk_rows = plan.opt.chunk_rows > 0 ? (uint64_t)plan.opt.chunk_rows
                                          : std::max<uint64_t>(65536, 2 * ((n + P1 - 1) / P1));
  if (!p1.tile_rows) return fail(ctx, TGA_ERR_INVALID_ARG, "adaptive plan without a first-pass tile");
  const unsigned grid = scatter_grid(ctx, q, p1, n);
  const uint64_t T = plan.table_slots;
  pb.gs_cap1 = full_table ? 0 : next_pow2(std::max<uint64_t>(4096, T * P1));
  if (tga_status e = first_pass(ctx, q, keys, q.schema, n, p1, b1, grid, p1.chunk_rows, pb.gs_cap1,
                                (uint32_t)(2 * T), pb))
    return e;
  // the hash-slice count over partition 0
  const uint32_t slots = slice_set_slots(n);
  DBuf set, cbuf;
  if (!set.alloc(ctx, (uint64_t)slots * q.key_bytes) || !cbuf.alloc(ctx, 64))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "slice set");
  TGA_HIP(ctx, timed(ctx, "slice_records", [&] {
    return launch_slice_records(q.key_bytes, pb.pkeys1.p, (uint32_t)(q.key_bytes / 4 + vd), pb.ap1.blist,
                                pb.ap1.pstart, set.p, slots, cbuf.as<uint64_t>(), (unsigned)ctx->cus, st);
  }));
  uint64_t c[2];
  TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host, cbuf.p, sizeof c, hipMemcpyDeviceToHost, st));
  TGA_HIP(ctx, hipStreamSynchronize(st));
  std::memcpy(c, ctx->status_host, sizeof c);
  const uint64_t est = combine_slice_estimate(plan.est, plan.sample_d, c[0], c[1] != 0, n);
  Plan np = autotune(ctx, q, n, est, &plan.opt);
  int bits = np.path == TGA_PATH_PARTITIONED ? np.pbits : b1;
  bits = std::max(b1, std::min(bits, b1 + kTwoPassMaxSubBits));
  if (np.path != TGA_PATH_PARTITIONED || np.rec_vd != vd || np.table_slots != plan.table_slots) np = plan;
  np.est = est;
  np.gslots = std::min<uint64_t>(next_pow2(std::max<uint64_t>(1024, est * 2)),
                                 next_pow2(std::max<uint64_t>(1024, 2 * n)));
  np.pbits = bits;
  np.max_pbits = std::max(np.max_pbits, bits);
  np.scatter_wc = 1;
  np.tile_rows = p1.tile_rows;
  np.scatter_p1_bits = bits > b1 ? b1 : 0;
  np.chunk_rows = plan.opt.chunk_rows > 0 ? (uint64_t)plan.opt.chunk_rows
                                          : std::max<uint64_t>(65536, 2 * ((n + (1ULL << bits) - 1) >> bits));
  np.may_replan = bits < np.max_pbits;
  np.adapt_bits = 0;
  plan = np;
  pb.pass1_done = true;
  pb.pass1_final = bits == b1;
  return TGA_OK;
}

static tga_status run_partitioned(tga_context *ctx, const QuerySpec &q, const void *d_keys, uint64_t n, Plan &plan,
                                  bool full_table, PlainOut *out, void *ext_keys, void *const *ext_cols,
                                  uint64_t ext_cap, uint64_t *groups, bool *overflow, uint64_t *replan_est,
                                  bool *reduced) {
  *replan_est = 0;
  hipStream_t st = ctx->stream;
  const Schema &s = q.schema;
  PartBufs pb;
  if (plan.adapt_bits > 0)
    if (tga_status e = adapt_first_pass(ctx, q, d_keys, n, plan, full_table, pb)) return e;
  const uint32_t P = 1u << plan.pbits;
  DBuf counters, ovf;
  // at most one abandoned record per work item: partitions plus split chunks
  const uint32_t ovf_cap = P + (uint32_t)std::min<uint64_t>(1u << 24, n / std::max<uint64_t>(plan.chunk_rows, 1) + 1);
  if (!counters.alloc(ctx, 64) || !ovf.alloc(ctx, (uint64_t)(ovf_cap + 1) * 8))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "partition buffers");
  uint64_t *emit_count = counters.as<uint64_t>();
  const bool may_abandon = plan.may_replan;
  // both are first used by the aggregation: the global table's init launch
  // clears them
  WordInit wi{};
  wi.p[0] = emit_count;
  wi.n[0] = 2;
  if (may_abandon) {
    wi.p[1] = ovf.as<uint64_t>();
    wi.v[1] = (uint64_t)ovf_cap << 32;
    wi.n[1] = 1;
  }
  // Global table: split partitions and LDS spills only.  A partition is
  // split past chunk_rows (twice its expected rows, so only a key holding
  // more than half a partition's share of the rows splits one), and it holds
  // about T/2 groups.  The plan kernel sizes the table from the split
  // partitions it counts, 2T slots each (at most 2 slots a row), so the host
  // never waits for the partitioning; the allocation is that size's bound:
  // fewer than P/2 partitions can exceed 2 n/P rows.  A table that turns out
  // too small (spills from partitions holding more groups than their LDS
  // table) raises the overflow flag, and the call is retried with the
  // full-size one.
  const uint64_t T = plan.table_slots;
  uint64_t S = plan.gslots;
  if (!full_table) {
    S = std::min(S, next_pow2(std::max<uint64_t>(4096, T * P)));
    // a first pass kept as the copy had its plan kernel size the table up
    // to gs_cap1
    if (pb.pass1_final) S = std::max(S, pb.gs_cap1);
    pb.gs_cap = S;
    pb.gs_per_split = (uint32_t)(2 * T);
  }
  {
    tga_status e = partition_rows(ctx, q, d_keys, s, n, plan, pb);
    if (e) return e;
  }
  const AggPlan &ap = pb.ap;
  GTable g{};
  std::vector<DBuf> gbufs;
  DBuf flags;
  if (gtable_alloc(ctx, q, S, g, gbufs, flags) != hipSuccess)
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "global table of %llu slots", (unsigned long long)S);
  g.S_dev = ap.gslots;
  TGA_HIP(ctx, timed(ctx, "gtable_init", [&] { return launch_gtable_init(q.key_bytes, g, s, st, wi); }));
  // outputs
  const uint64_t cap = ext_keys ? ext_cap : plan.gslots + 1;
  PlainOut tmp;
  OutDesc od{};
  od.capacity = cap;
  if (ext_keys) {
    od.keys = ext_keys;
    for (int a = 0; a < s.nacc; ++a) od.cols[a] = ext_cols[a];
  } else {
    if (!tmp.keys.alloc(ctx, std::max<uint64_t>(cap, 1) * q.key_bytes)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "output");
    od.keys = tmp.keys.p;
    for (int a = 0; a < s.nacc; ++a) {
      if (!tmp.cols[a].alloc(ctx, std::max<uint64_t>(cap, 1) * (uint64_t)dtype_bytes(q.out_dtype[a])))
        return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "output");
      od.cols[a] = tmp.cols[a].p;
    }
  }
  TGA_HIP(ctx, timed(ctx, "agg_partitions", [&] {
    return launch_agg(q.key_bytes, pb.pkeys.p, pb.sp, ap, plan.layout, g, plan.nbits, plan.max_probe, plan.workgroups,
                      plan.block, st, 1, &od, emit_count, plan.fold, may_abandon ? ovf.as<uint32_t>() : nullptr,
                      pb.rec);
  }));
  // the global table's groups follow the emitted ones
  const uint64_t ctiles = compact_tiles(g.S + 1);
  DBuf cdesc;
  if (!cdesc.alloc(ctx, (ctiles + 2) * 8 + 64)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "compaction");
  TGA_HIP(ctx, launch_fill_u64(cdesc.as<uint64_t>(), ctiles + 2, 0, st));
  uint64_t *count = cdesc.as<uint64_t>() + ctiles + 1;
  TGA_HIP(ctx, timed(ctx, "compact", [&] {
    return launch_compact(q.key_bytes, g, s, od, cdesc.as<uint64_t>(), (uint32_t *)(cdesc.as<uint64_t>() + ctiles),
                          count, st, emit_count);
  }));
  // flags | group count | abandoned items | buckets taken, in one copy
  uint32_t hflags[kFlagWords];
  uint64_t hcount = 0;
  uint32_t novf = 0, used = 0, used1 = 0;
  uint64_t S_used = S;
  {
    WordGather wg{};
    wg.src[0] = g.flags;
    wg.n[0] = kFlagWords;
    wg.src[1] = (const uint32_t *)count;
    wg.n[1] = 2;
    wg.src[2] = may_abandon ? ovf.as<uint32_t>() : (const uint32_t *)count;
    wg.n[2] = 1;
    wg.src[3] = pb.buckets ? pb.bb.pool : (const uint32_t *)count;
    wg.n[3] = 1;
    wg.src[4] = pb.pool1 ? pb.pool1 : (const uint32_t *)count;
    wg.n[4] = 1;
    wg.src[5] = ap.gslots ? (const uint32_t *)ap.gslots : (const uint32_t *)count;
    wg.n[5] = 2;
    wg.count = 6;
    DBuf sbuf;
    if (!sbuf.alloc(ctx, kStatusWords * 4)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "status words");
    uint32_t *dst = sbuf.as<uint32_t>();
    TGA_HIP(ctx, launch_gather_words(wg, dst, st));
    TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host, dst, (kFlagWords + 7) * 4, hipMemcpyDeviceToHost, st));
    TGA_HIP(ctx, hipStreamSynchronize(st));
    const uint32_t *h = ctx->status_host;
    std::memcpy(hflags, h, sizeof hflags);
    std::memcpy(&hcount, h + kFlagWords, 8);
    if (may_abandon) novf = h[kFlagWords + 2];
    if (pb.buckets) used = h[kFlagWords + 3];
    if (pb.pool1) used1 = h[kFlagWords + 4];
    if (ap.gslots) std::memcpy(&S_used, h + kFlagWords + 5, 8);
  }
  *reduced = S_used < plan.gslots;
  if (ctx->prof.on) ctx->prof.resolve();
  // the scatter never takes more buckets than bucket_cap(); a larger count
  // would mean dropped rows
  if (pb.buckets && used > pb.bb.cap)
    return fail(ctx, TGA_ERR_HIP, "bucket pool overflow (%u > %u)", used, pb.bb.cap);
  if (pb.pool1 && used1 > pb.cap1)
    return fail(ctx, TGA_ERR_HIP, "first-pass bucket pool overflow (%u > %u)", used1, pb.cap1);
  if (novf) {
    // Abandoned work items: a table of T slots was full after r of the
    // item's R rows, so the item holds about T * R / r groups; the finished
    // ones are counted exactly (hcount).
    novf = std::min(novf, ovf_cap);
    std::vector<uint32_t> list(2 * (size_t)novf);
    TGA_HIP(ctx, hipMemcpyAsync(list.data(), ovf.as<uint32_t>() + 2, list.size() * 4, hipMemcpyDeviceToHost, st));
    TGA_HIP(ctx, hipStreamSynchronize(st));
    double est = (double)hcount;
    for (uint32_t i = 0; i < novf; ++i) {
      const double r = (double)std::max<uint32_t>(1, list[2 * i]), R = (double)list[2 * i + 1];
      est += (double)plan.table_slots * std::max(1.0, R / r);
    }
    *replan_est = (uint64_t)std::min<double>((double)n, est);
    return TGA_OK;
  }
  *overflow = hflags[FLAG_OVERFLOW] != 0 || (!ext_keys && hcount > cap);
  if (*overflow) {
    // either the spill table or the output was too small: caller retries
    if (ext_keys && hcount > ext_cap) {
      *overflow = false;
      *groups = hcount;
      return fail(ctx, TGA_ERR_CAPACITY, "%llu groups > capacity %llu", (unsigned long long)hcount,
                  (unsigned long long)ext_cap);
    }
    return TGA_OK;
  }
  if (hflags[FLAG_RANGE]) return fail(ctx, TGA_ERR_RANGE, "fp64 SUM input outside [2^-96, 2^64) or not finite");
  *groups = hcount;
  if (out) {
    out->keys = std::move(tmp.keys);
    for (int a = 0; a < s.nacc; ++a) out->cols[a] = std::move(tmp.cols[a]);
    out->groups = hcount;
    out->capacity = cap;
  }
  return TGA_OK;
}

tga_status run_plain(tga_context *ctx, const QuerySpec &q, const void *d_keys, uint64_t n, const Plan &plan_in,
                     PlainOut *out, void *ext_keys, void *const *ext_cols, uint64_t ext_cap, uint64_t *groups) {
  Plan plan = plan_in;
  hipStream_t st = ctx->stream;
  const Schema &s = q.schema;
  if (plan.path == TGA_PATH_PARTITIONED && n > 0 && n < (1ULL << 31)) {
    // Partitions expected to fill their LDS tables past 3/4 (a fanout capped
    // below the estimate's) spill rows to the global table, which the
    // split-partition sizing does not count: the full-size table from the
    // start, instead of a run that overflows a reduced one.
    bool grown = plan.est * 4 > ((uint64_t)plan.table_slots << plan.pbits) * 3;
    for (int attempt = 0; attempt < 8; ++attempt) {
      bool overflow = false, reduced = false;
      uint64_t replan = 0;
      tga_status st2 = run_partitioned(ctx, q, d_keys, n, plan, grown, out, ext_keys, ext_cols, ext_cap, groups,
                                       &overflow, &replan, &reduced);
      if (st2) return st2;
      if (replan) {
        // the estimate was too low for the fanout: plan again from the
        // group count the abandoned partitions imply, at least doubling P
        const int old_bits = plan.pbits;
        Plan np = autotune(ctx, q, n, std::max<uint64_t>(replan, 1), &plan.opt);
        if (np.path != TGA_PATH_PARTITIONED) np = plan;
        if (np.pbits <= old_bits) {
          tga_options o2 = plan.opt;
          o2.partition_bits = std::min(plan.max_pbits, old_bits + 1);
          o2.est_groups = (int64_t)replan;
          np = autotune(ctx, q, n, std::max<uint64_t>(replan, 1), &o2);
          np.may_replan = np.pbits < np.max_pbits;
          np.opt = plan.opt;
        }
        plan = np;
        continue;
      }
      if (!overflow) {
        ctx->last_plan = tga_options{};
        plan_to_options(plan, ctx->last_plan);
        ctx->last_plan.attempts = attempt + 1;
        return TGA_OK;
      }
      // a reduced split/spill table that overflowed: the full-size one first;
      // otherwise the estimate was too low: larger tables, run again
      if (!(reduced && !grown)) plan.gslots *= 4;
      grown = true;
    }
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "global table kept overflowing");
  }
  bool sliced = false;
  for (int attempt = 0; attempt < 6; ++attempt) {
    GTable g{};
    std::vector<DBuf> gbufs;
    DBuf flags;
    if (gtable_alloc(ctx, q, plan.gslots, g, gbufs, flags) != hipSuccess)
      return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "global table of %llu slots", (unsigned long long)plan.gslots);
    TGA_HIP(ctx, timed(ctx, "gtable_init", [&] { return launch_gtable_init(q.key_bytes, g, s, st); }));

    std::vector<DBuf> keep;  // temporaries alive until the stream is synchronised
    {
      tga_status a = aggregate_into(ctx, q, d_keys, n, plan, g, keep);
      if (a) return a;
    }

    // (d) compaction
    const uint64_t cap = ext_keys ? ext_cap : plan.gslots + 1;
    PlainOut tmp;
    OutDesc od{};
    od.capacity = cap;
    if (ext_keys) {
      od.keys = ext_keys;
      for (int a = 0; a < s.nacc; ++a) od.cols[a] = ext_cols[a];
    } else {
      if (!tmp.keys.alloc(ctx, std::max<uint64_t>(cap, 1) * q.key_bytes)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "output");
      od.keys = tmp.keys.p;
      for (int a = 0; a < s.nacc; ++a) {
        if (!tmp.cols[a].alloc(ctx, std::max<uint64_t>(cap, 1) * out_elem_bytes(q.out_dtype[a])))
          return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "output");
        od.cols[a] = tmp.cols[a].p;
      }
    }
    uint32_t hflags[kFlagWords];
    uint64_t hcount = 0;
    {
      tga_status c = compact_into(ctx, q, g, od, hflags, &hcount);
      if (c) return c;
    }
    if (ctx->prof.on) ctx->prof.resolve();
    if (ext_keys && hcount > ext_cap) {
      *groups = hcount;
      return fail(ctx, TGA_ERR_CAPACITY, "%llu groups > capacity %llu", (unsigned long long)hcount,
                  (unsigned long long)ext_cap);
    }
    if (hflags[FLAG_OVERFLOW]) {
      // The estimate was too low: the table held more keys than half its
      // slots.  Plan again for at least 4x as many groups -- over 2^22 rows
      // from the exact count of a hash slice of the keys, which no row order
      // or key frequency misleads -- possibly on the partitioned path.
      uint64_t est2 = plan.gslots * 2;
      if (!sliced && n >= kSliceMinRows) {
        uint64_t c = 0;
        bool ovf = false;
        if (tga_status e = slice_count(ctx, d_keys, q.key_bytes, n, &c, &ovf)) return e;
        est2 = std::max(est2, combine_slice_estimate(est2, 0, c, ovf, n));
        sliced = true;
      }
      const uint64_t grown = plan.gslots * 4;
      Plan np = autotune(ctx, q, n, std::min<uint64_t>(est2, n), &plan.opt);
      if (np.path == TGA_PATH_PARTITIONED)
        return run_plain(ctx, q, d_keys, n, np, out, ext_keys, ext_cols, ext_cap, groups);
      plan = np;
      plan.gslots = std::max(plan.gslots, grown);
      continue;
    }
    if (hflags[FLAG_RANGE]) return fail(ctx, TGA_ERR_RANGE, "fp64 SUM input outside [2^-96, 2^64) or not finite");
    *groups = hcount;
    if (out) {
      out->keys = std::move(tmp.keys);
      for (int a = 0; a < s.nacc; ++a) out->cols[a] = std::move(tmp.cols[a]);
      out->groups = hcount;
      out->capacity = cap;
    }
    ctx->last_plan = tga_options{};
    plan_to_options(plan, ctx->last_plan);
    ctx->last_plan.attempts = attempt + 1;
    if (ext_keys && hcount > ext_cap) return fail(ctx, TGA_ERR_CAPACITY, "%llu groups > capacity %llu",
                                                 (unsigned long long)hcount, (unsigned long long)ext_cap);
    return TGA_OK;
  }
  return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "global table kept overflowing");
}

}  // namespace tga

using namespace tga;

// ---------------------------------------------------------------------------
// staging of host columns
// ---------------------------------------------------------------------------
namespace {
struct Staged {
  tga_column key{};
  tga_column vals[kMaxCols]{};
  DBuf bufs[kMaxCols + 1];
};

tga_status stage(tga_context *ctx, const tga_column *key, const tga_column *values, int32_t nv, int64_t n, Staged &o) {
  o.key = *key;
  for (int c = 0; c < nv; ++c) o.vals[c] = values[c];
  auto one = [&](tga_column &col, DBuf &b) -> tga_status {
    if (col.location != TGA_HOST || n == 0) return TGA_OK;
    const uint64_t bytes = (uint64_t)n * dtype_bytes(col.dtype);
    if (!b.alloc(ctx, bytes)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "staging");
    TGA_HIP(ctx, hipMemcpyAsync(b.p, col.data, bytes, hipMemcpyHostToDevice, ctx->stream));
    col.data = b.p;
    col.location = TGA_DEVICE;
    return TGA_OK;
  };
  tga_status st = one(o.key, o.bufs[0]);
  if (st) return st;
  for (int c = 0; c < nv; ++c)
    if ((st = one(o.vals[c], o.bufs[c + 1]))) return st;
  return TGA_OK;
}

tga_status validate(tga_context *ctx, const tga_column *key, const tga_column *values, int32_t n_values,
                    int64_t n_rows, const tga_agg *aggs, int32_t n_aggs, QuerySpec &q) {
  if (n_rows < 0) return fail(ctx, TGA_ERR_INVALID_ARG, "negative row count");
  std::string e = build_spec(key, values, n_values, aggs, n_aggs, q);
  if (!e.empty()) return fail(ctx, TGA_ERR_INVALID_ARG, "%s", e.c_str());
  if (n_rows > 0 && !key->data) return fail(ctx, TGA_ERR_INVALID_ARG, "key data is null");
  for (int c = 0; c < n_values; ++c)
    if (n_rows > 0 && !values[c].data) return fail(ctx, TGA_ERR_INVALID_ARG, "value column %d is null", c);
  return TGA_OK;
}

// Stages host columns whole, estimates the group count (hint, else a device
// sample) and plans; q comes from validate().
tga_status prepare(tga_context *ctx, const tga_column *key, const tga_column *values, int32_t n_values, int64_t n_rows,
                   const tga_options *opt, Staged &sg, QuerySpec &q, Plan &plan) {
  tga_status st = stage(ctx, key, values, n_values, n_rows, sg);
  if (st) return st;
  for (int c = 0; c < n_values; ++c) q.schema.cols[c] = sg.vals[c].data;
  uint64_t est = opt && opt->est_groups > 0 ? (uint64_t)opt->est_groups : 0;
  bool deferred = false;
  uint64_t sample_d = 0;
  if (!est) {
    st = estimate_groups(ctx, sg.key.data, q.key_bytes, (uint64_t)n_rows, &est, nullptr, &deferred, &sample_d);
    if (st) return st;
  }
  plan = autotune(ctx, q, (uint64_t)n_rows, est, opt);
  if (deferred) return resolve_deferred_estimate(ctx, q, sg.key.data, (uint64_t)n_rows, est, sample_d, opt, plan);
  return TGA_OK;
}
}  // namespace

#define TGA_API extern "C" __attribute__((visibility("default")))

TGA_API tga_status tga_options_default(tga_options *opt) {
  if (!opt) return TGA_ERR_INVALID_ARG;
  *opt = tga_options{};
  return TGA_OK;
}

TGA_API tga_status tga_context_create(int32_t device, uint64_t mem_budget_bytes, tga_context **out) {
  if (!out) return TGA_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TGA_ERR_HIP;
  if (device < 0 || device >= n) return TGA_ERR_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return TGA_ERR_HIP;
  auto *c = new tga_context();
  c->device = device;
  if (hipGetDeviceProperties(&c->props, device) != hipSuccess) { delete c; return TGA_ERR_HIP; }
  c->cus = c->props.multiProcessorCount;
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->lds_max = (size_t)lds;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return TGA_ERR_HIP; }
  uint64_t budget = mem_budget_bytes;
  if (!budget) {
    size_t fr = 0, tot = 0;
    hipMemGetInfo(&fr, &tot);
    budget = (uint64_t)(fr * 0.9);
  }
  c->arena = new Arena(budget);
  c->pinned = new PinnedPool();
  // the no-alloc entry point's scratch address on this device (once per
  // device; a failure only disables that entry point)
  (void)noalloc_scratch_init(device);
  if (hipHostMalloc((void **)&c->status_host, kStatusWords * 4, hipHostMallocDefault) != hipSuccess) {
    delete c->arena;
    hipStreamDestroy(c->stream);
    delete c;
    return TGA_ERR_HIP;
  }
  *out = c;
  return TGA_OK;
}

TGA_API tga_status tga_context_destroy(tga_context *ctx) {
  if (!ctx) return TGA_ERR_INVALID_ARG;
  for (tga_context *l : ctx->lanes) tga_context_destroy(l);
  ctx->lanes.clear();
  {
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    ctx->prof.destroy();
    if (ctx->owns_arena) delete ctx->arena;
    ctx->arena = nullptr;
    delete ctx->pinned;
    ctx->pinned = nullptr;
    hipHostFree(ctx->status_host);
    ctx->status_host = nullptr;
    hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return TGA_OK;
}

TGA_API const char *tga_last_error(const tga_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

TGA_API tga_status tga_device_info(tga_context *ctx, char *name, size_t name_len, int32_t *cus, int64_t *hbm) {
  if (!ctx) return TGA_ERR_INVALID_ARG;
  if (name && name_len) { std::snprintf(name, name_len, "%s", ctx->props.gcnArchName); }
  if (cus) *cus = ctx->cus;
  if (hbm) *hbm = (int64_t)ctx->props.totalGlobalMem;
  return TGA_OK;
}

TGA_API tga_status tga_groupby(tga_context *ctx, const tga_column *key, const tga_column *values, int32_t n_values,
                               int64_t n_rows, const tga_agg *aggs, int32_t n_aggs, const tga_options *opt,
                               tga_result **out) {
  if (!ctx || !out) return TGA_ERR_INVALID_ARG;
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  Staged sg;
  QuerySpec q;
  Plan plan;
  tga_status st = validate(ctx, key, values, n_values, n_rows, aggs, n_aggs, q);
  if (st) return st;
  const bool streamed = should_stream(key, values, n_values, (uint64_t)n_rows);
  if (!streamed && (st = prepare(ctx, key, values, n_values, n_rows, opt, sg, q, plan))) return st;
  auto *r = new tga_result();
  r->ctx = ctx;
  r->key_dtype = q.key_dtype;
  r->naggs = n_aggs;
  for (int a = 0; a < n_aggs; ++a) r->dtype[a] = q.out_dtype[a];
  uint64_t G = 0;
  st = streamed ? run_streamed(ctx, q, key, values, n_values, (uint64_t)n_rows, opt, &r->out, nullptr, nullptr, 0, &G)
                : run_plain(ctx, q, sg.key.data, (uint64_t)n_rows, plan, &r->out, nullptr, nullptr, 0, &G);
  if (st) { delete r; return st; }
  *out = r;
  return TGA_OK;
}

TGA_API tga_status tga_groupby_into(tga_context *ctx, const tga_column *key, const tga_column *values, int32_t n_values,
                                    int64_t n_rows, const tga_agg *aggs, int32_t n_aggs, const tga_options *opt,
                                    void *out_keys, void *const *out_aggs, int64_t capacity, int64_t *n_groups) {
  if (!ctx || !out_keys || !out_aggs || capacity < 0 || !n_groups) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  for (int a = 0; a < n_aggs && a < kMaxAccs; ++a)
    if (!out_aggs[a]) return fail(ctx, TGA_ERR_INVALID_ARG, "output column %d is null", a);
  Staged sg;
  QuerySpec q;
  Plan plan;
  tga_status st = validate(ctx, key, values, n_values, n_rows, aggs, n_aggs, q);
  if (st) return st;
  uint64_t G = 0;
  if (should_stream(key, values, n_values, (uint64_t)n_rows)) {
    st = run_streamed(ctx, q, key, values, n_values, (uint64_t)n_rows, opt, nullptr, out_keys, out_aggs,
                      (uint64_t)capacity, &G);
  } else {
    if ((st = prepare(ctx, key, values, n_values, n_rows, opt, sg, q, plan))) return st;
    st = run_plain(ctx, q, sg.key.data, (uint64_t)n_rows, plan, nullptr, out_keys, out_aggs, (uint64_t)capacity, &G);
  }
  *n_groups = (int64_t)G;
  return st;
}

TGA_API tga_status tga_last_plan(tga_context *ctx, tga_options *plan) {
  if (!ctx || !plan) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  *plan = ctx->last_plan;
  return TGA_OK;
}

TGA_API tga_status tga_result_num_groups(const tga_result *r, int64_t *n) {
  if (!r || !n) return TGA_ERR_INVALID_ARG;
  *n = (int64_t)r->out.groups;
  return TGA_OK;
}

TGA_API tga_status tga_result_column(const tga_result *r, int32_t index, const void **ptr, int32_t *dtype) {
  if (!r || index < -1 || index >= r->naggs) return TGA_ERR_INVALID_ARG;
  if (index < 0) {
    if (ptr) *ptr = r->out.keys.p;
    if (dtype) *dtype = r->key_dtype;
  } else {
    if (ptr) *ptr = r->out.cols[index].p;
    if (dtype) *dtype = r->dtype[index];
  }
  return TGA_OK;
}

TGA_API tga_status tga_result_copy_to_host(const tga_result *r, int32_t index, void *dst, int64_t cap) {
  if (!r || !dst || index < -1 || index >= r->naggs || cap < 0) return TGA_ERR_INVALID_ARG;
  const int dt = index < 0 ? r->key_dtype : r->dtype[index];
  const void *src = index < 0 ? r->out.keys.p : r->out.cols[index].p;
  const uint64_t rows = std::min<uint64_t>((uint64_t)cap, r->out.groups);
  if (rows == 0) return TGA_OK;
  std::lock_guard<std::recursive_mutex> g(r->ctx->mu);
  hipSetDevice(r->ctx->device);
  TGA_HIP(r->ctx, hipMemcpyAsync(dst, src, rows * dtype_bytes(dt), hipMemcpyDeviceToHost, r->ctx->stream));
  TGA_HIP(r->ctx, hipStreamSynchronize(r->ctx->stream));
  return TGA_OK;
}

TGA_API tga_status tga_result_free(tga_result *r) {
  if (!r) return TGA_ERR_INVALID_ARG;
  tga_context *ctx = r->ctx;
  {
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    delete r;
  }
  return TGA_OK;
}

TGA_API tga_status tga_scan_exclusive_u32(tga_context *ctx, const uint32_t *d_in, uint32_t *d_out, int64_t n) {
  if (!ctx || n < 0 || (n > 0 && (!d_in || !d_out))) return TGA_ERR_INVALID_ARG;
  if (n == 0) return TGA_OK;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  const uint64_t tiles = scan_tiles((uint64_t)n);
  DBuf desc;
  if (!desc.alloc(ctx, (tiles + 1) * 8 + 64)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "scan descriptors");
  TGA_HIP(ctx, launch_fill_u64(desc.as<uint64_t>(), tiles + 1, 0, ctx->stream));
  TGA_HIP(ctx, timed(ctx, "scan", [&] {
    return launch_scan_u32(d_in, d_out, (uint64_t)n, desc.as<uint64_t>(), (uint32_t *)(desc.as<uint64_t>() + tiles),
                           ctx->stream);
  }));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->prof.on) ctx->prof.resolve();
  return TGA_OK;
}

TGA_API tga_status tga_generate(tga_context *ctx, int32_t kind, void *d_out, int64_t n, uint64_t a, uint64_t b,
                                double x, double y, uint64_t seed, uint64_t salt) {
  if (!ctx || n < 0 || (n > 0 && !d_out)) return TGA_ERR_INVALID_ARG;
  if ((kind == TGA_GEN_KEYS_U32 || kind == TGA_GEN_KEYS_U64 || kind == TGA_GEN_ZIPF_U64 ||
       kind == TGA_GEN_MOD_U32 || kind == TGA_GEN_MOD_U64) && a == 0)
    return fail(ctx, TGA_ERR_INVALID_ARG, "distinct count must be positive");
  if (kind == TGA_GEN_VALS_I32 && b == 0) return fail(ctx, TGA_ERR_INVALID_ARG, "empty value range");
  if (n == 0) return TGA_OK;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  TGA_HIP(ctx, launch_gen(kind, d_out, (uint64_t)n, a, b, x, y, seed, salt, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return TGA_OK;
}

TGA_API tga_status tga_host_alloc(tga_context *ctx, uint64_t bytes, void **out) {
  if (!ctx || !out) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  void *p = nullptr;
  TGA_HIP(ctx, hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault));
  *out = p;
  return TGA_OK;
}

TGA_API tga_status tga_device_alloc(tga_context *ctx, uint64_t bytes, void **out) {
  if (!ctx || !out) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  void *p = nullptr;
  TGA_HIP(ctx, hipMalloc(&p, bytes ? bytes : 1));
  *out = p;
  return TGA_OK;
}

TGA_API tga_status tga_device_free(tga_context *ctx, void *p) {
  if (!ctx) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  TGA_HIP(ctx, hipFree(p));
  return TGA_OK;
}

TGA_API tga_status tga_memcpy_h2d(tga_context *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  TGA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return TGA_OK;
}

TGA_API tga_status tga_memcpy_d2h(tga_context *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  TGA_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return TGA_OK;
}

TGA_API tga_status tga_device_copy(tga_context *ctx, void *dst, const void *src, uint64_t bytes) {
  if (!ctx || (bytes && (!dst || !src)) || (bytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
    return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  hipSetDevice(ctx->device);
  TGA_HIP(ctx, timed(ctx, "copy16", [&] { return launch_copy16(src, dst, bytes / 16, ctx->stream); }));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->prof.on) ctx->prof.resolve();
  return TGA_OK;
}

TGA_API tga_status tga_synchronize(tga_context *ctx) {
  if (!ctx) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return TGA_OK;
}

TGA_API tga_status tga_profile_enable(tga_context *ctx, int32_t on) {
  if (!ctx) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  ctx->prof.on = on != 0;
  return TGA_OK;
}

TGA_API tga_status tga_profile_reset(tga_context *ctx) {
  if (!ctx) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  ctx->prof.reset();
  return TGA_OK;
}

TGA_API tga_status tga_profile_entry(tga_context *ctx, int32_t i, char *name, size_t name_len, int64_t *launches,
                                     double *total_ms) {
  if (!ctx || i < 0) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> g(ctx->mu);
  ctx->prof.resolve();
  auto v = ctx->prof.entries();
  if ((size_t)i >= v.size()) return TGA_ERR_INVALID_ARG;
  if (name && name_len) std::snprintf(name, name_len, "%s", v[(size_t)i].first.c_str());
  if (launches) *launches = v[(size_t)i].second.first;
  if (total_ms) *total_ms = v[(size_t)i].second.second;
  return TGA_OK;
}
