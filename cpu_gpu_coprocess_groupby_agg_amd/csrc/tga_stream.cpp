// tga_stream.cpp -- host-resident input streamed through HBM tile by tile,
// the copies overlapped with the aggregation of the previous tile.
//
// The reference streams its pinned host arrays to the GPU one tile at a time
// (T/src/group_then_partition.cu:353-354, 413-418, 466-499: a tile's rows are
// copied to the device, pre-aggregated and partitioned there).  Here every
// tile's rows go straight into one global hash table that lives in HBM for
// the whole call (the plain operator's direct or partitioned path per tile,
// its LDS tables flushed into that table), and the table is compacted once at
// the end: integer aggregates and the exact fp64 limbs stay exact across
// tiles, and the input may be far larger than the device budget -- only two
// tiles and the table are resident.
//
// Copies run on a second stream into two staging sets: while tile i is
// aggregated on the context stream, tile i+1 is copied; the context stream
// waits for that copy's event before its own work on the tile, and a staging
// set is refilled only after the tile it held was aggregated (the host waits
// for each tile's aggregation before issuing the copy into its set again).
// Pageable host memory is copied synchronously by the runtime, so only
// pinned columns (hipHostMalloc / tga_host_alloc) overlap.
#include <algorithm>
#include <thread>

#include "tga_host.hpp"

namespace tga {

// A host column is streamed when the call has at least this many rows; below
// it the columns are staged whole.
constexpr uint64_t kStreamMinRows = 1ULL << 22;

bool should_stream(const tga_column *key, const tga_column *values, int nv, uint64_t n) {
  if (n < kStreamMinRows) return false;
  if (key->location == TGA_HOST) return true;
  for (int c = 0; c < nv; ++c)
    if (values[c].location == TGA_HOST) return true;
  return false;
}

namespace {

bool is_pinned(const void *p) {
  hipPointerAttribute_t a{};
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

// Group estimate of host-resident keys: 16384 keys at random rows; when that
// sample leaves the count open, the exact distinct count of the keys in a
// 2^-10 slice of their hash, over all rows, on host threads.
uint64_t host_estimate(const void *keys, int key_bytes, uint64_t n) {
  auto key_at = [&](uint64_t r) -> uint64_t {
    return key_bytes == 4 ? (uint64_t)((const uint32_t *)keys)[r] : ((const uint64_t *)keys)[r];
  };
  const uint64_t m = std::min<uint64_t>(n, 16384);
  std::vector<uint64_t> v(m);
  for (uint64_t i = 0; i < m; ++i) v[i] = key_at(sample_row(i, m, n));
  bool reliable = true;
  uint64_t est = estimate_from_sample(v.data(), m, n, nullptr, &reliable);
  uint64_t d = 0;
  for (uint64_t i = 0; i < m; ++i) d += (i == 0 || v[i] != v[i - 1]);
  if (reliable || n < kSliceMinRows) return est;
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::vector<uint64_t>> part(nt);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const uint64_t b = n * t / nt, e = n * (t + 1) / nt;
      auto &out = part[t];
      for (uint64_t r = b; r < e; ++r) {
        const uint64_t k = key_at(r);
        if ((key_hash(k) >> (64 - kSliceBits)) == 0) out.push_back(k);
      }
      std::sort(out.begin(), out.end());
      out.erase(std::unique(out.begin(), out.end()), out.end());
    });
  for (auto &x : th) x.join();
  std::vector<uint64_t> all;
  for (auto &p : part) all.insert(all.end(), p.begin(), p.end());
  std::sort(all.begin(), all.end());
  const uint64_t c = (uint64_t)(std::unique(all.begin(), all.end()) - all.begin());
  return combine_slice_estimate(est, d, c, false, n);
}

// The tuner's plan for `rows` rows, except that the bucket layout's pool --
// beside the rows a reserve of open buckets per scatter workgroup (up to 4
// per CU) and partition -- is traded for the partition-contiguous copy (a
// histogram pass first, exactly the rows) when that reserve would take more
// than a quarter of a small budget.
Plan lean_plan(tga_context *ctx, const QuerySpec &q, uint64_t rows, uint64_t est, const tga_options *opt) {
  Plan plan = autotune(ctx, q, rows, est, opt);
  if (plan.path == TGA_PATH_PARTITIONED && plan.scatter_layout == 2 && !(opt && opt->scatter_layout)) {
    uint64_t rb = (uint64_t)q.key_bytes;
    for (int c = 0; c < q.schema.ncols; ++c) rb += (uint64_t)dtype_bytes(q.schema.col_dtype[c]);
    const uint64_t cap = rows / kBucketRows + (uint64_t)ctx->cus * 4 * ((1ULL << plan.pbits) + 1024);
    if (cap * kBucketRows * rb > ctx->arena->budget() / 4) {
      tga_options o2{};
      if (opt) o2 = *opt;
      o2.scatter_layout = 1;
      plan = autotune(ctx, q, rows, est, &o2);
    }
  }
  return plan;
}

}  // namespace

tga_status run_streamed(tga_context *ctx, const QuerySpec &q, const tga_column *key, const tga_column *values,
                        int nv, uint64_t n, const tga_options *opt, PlainOut *out, void *ext_keys,
                        void *const *ext_cols, uint64_t ext_cap, uint64_t *groups) {
  hipStream_t st = ctx->stream;
  const Schema &s0 = q.schema;
  // host columns: key (-1) and values; device value columns are read in place
  struct Src { const char *p; uint32_t bytes; bool host; };
  Src src[kMaxCols + 1];
  src[0] = {(const char *)key->data, (uint32_t)q.key_bytes, key->location == TGA_HOST};
  uint64_t host_row_bytes = src[0].host ? (uint64_t)q.key_bytes : 0;
  bool pinned = true;
  for (int c = 0; c < nv; ++c) {
    src[c + 1] = {(const char *)values[c].data, (uint32_t)dtype_bytes(values[c].dtype), values[c].location == TGA_HOST};
    if (src[c + 1].host) host_row_bytes += src[c + 1].bytes;
  }
  for (int c = 0; c <= nv; ++c)
    if (src[c].host) pinned = pinned && is_pinned(src[c].p);
  // tiles: the two staging sets and one tile's partitioned copy within an
  // eighth of the budget (the group tables and partials take their own
  // share), 2^16 .. 2^24 rows
  uint64_t tile = ctx->arena->budget() / 16 / std::max<uint64_t>(1, host_row_bytes + q.key_bytes + 8);
  tile = std::max<uint64_t>(1ULL << 16, std::min<uint64_t>(1ULL << 24, tile)) & ~4095ULL;
  tile = std::min<uint64_t>(tile, n);
  const uint64_t ntiles = (n + tile - 1) / tile;
  // The group estimate: the caller's hint; for integer aggregates without
  // one, estimates made on the device as the rows arrive (the first tile's
  // keys for the tiles' plan, the partial groups' keys for each merge), so
  // nothing reads the host keys ahead of the copies; otherwise one over all
  // rows, from the host keys.
  bool partials = s0.nfx == 0 && s0.nacc <= kMaxCols;
  uint64_t est = opt && opt->est_groups > 0 ? (uint64_t)opt->est_groups : 0;
  const bool deferred = !est && partials;
  if (!est && !deferred) {
    if (src[0].host) est = host_estimate(src[0].p, q.key_bytes, n);
    else if (tga_status e = estimate_groups(ctx, src[0].p, q.key_bytes, n, &est)) return e;
  }
  est = std::max<uint64_t>(1, std::min<uint64_t>(est, n));
  Plan plan = lean_plan(ctx, q, tile, std::min<uint64_t>(est, tile), opt);
  // staging sets
  DBuf stage[2][kMaxCols + 1];
  for (int b = 0; b < 2; ++b)
    for (int c = 0; c <= nv; ++c)
      if (src[c].host && !stage[b][c].alloc(ctx, tile * src[c].bytes))
        return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "streaming buffers of %llu rows", (unsigned long long)tile);
  hipStream_t cs = nullptr;
  hipEvent_t copied[2] = {nullptr, nullptr};
  struct Cleanup {
    hipStream_t *cs;
    hipEvent_t *ev;
    ~Cleanup() {
      if (*cs) { hipStreamSynchronize(*cs); hipStreamDestroy(*cs); }
      for (int i = 0; i < 2; ++i)
        if (ev[i]) hipEventDestroy(ev[i]);
    }
  } cleanup{&cs, copied};
  TGA_HIP(ctx, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) TGA_HIP(ctx, hipEventCreateWithFlags(&copied[i], hipEventDisableTiming));
  auto issue_copy = [&](uint64_t t) -> tga_status {
    const int b = (int)(t & 1);
    const uint64_t r0 = t * tile, m = std::min<uint64_t>(tile, n - r0);
    for (int c = 0; c <= nv; ++c)
      if (src[c].host)
        TGA_HIP(ctx, hipMemcpyAsync(stage[b][c].p, src[c].p + r0 * src[c].bytes, m * src[c].bytes,
                                    hipMemcpyHostToDevice, cs));
    TGA_HIP(ctx, hipEventRecord(copied[b], cs));
    return TGA_OK;
  };
  auto tile_cols = [&](uint64_t t, QuerySpec &qt, const void *&tkeys, uint64_t &m) {
    const int b = (int)(t & 1);
    const uint64_t r0 = t * tile;
    m = std::min<uint64_t>(tile, n - r0);
    qt = q;
    for (int c = 0; c < nv; ++c)
      qt.schema.cols[c] = src[c + 1].host ? stage[b][c + 1].p : (const void *)(src[c + 1].p + r0 * src[c + 1].bytes);
    tkeys = src[0].host ? stage[b][0].p : (const void *)(src[0].p + r0 * q.key_bytes);
  };
  auto report = [&](uint64_t est_slots, int attempts) {
    plan_to_options(plan, ctx->last_plan);  // the plan of one tile
    ctx->last_plan.est_groups = (int64_t)est_slots;
    ctx->last_plan.attempts = attempts;
    ctx->last_plan.streamed_tiles = (int32_t)ntiles;
    ctx->last_plan.stream_pinned = pinned ? 1 : 0;
  };

  // Without a hint, the first tile's keys (on the device) give the tiles'
  // plan.  Where a tile hardly reduces (its groups above an eighth of its
  // rows), per-tile partials would cost more than they save: one global
  // table instead, sized by an estimate over all rows.
  bool staged0 = false;
  if (deferred) {
    if (tga_status e = issue_copy(0)) return e;
    staged0 = true;
    TGA_HIP(ctx, hipEventSynchronize(copied[0]));
    QuerySpec q0;
    const void *k0;
    uint64_t m0;
    tile_cols(0, q0, k0, m0);
    uint64_t et = 0;
    if (tga_status e = estimate_groups(ctx, k0, q.key_bytes, m0, &et)) return e;
    est = std::max<uint64_t>(1, et);
  }
  if (partials && std::min<uint64_t>(est, tile) * 8 > tile) {
    partials = false;
    if (deferred) {
      if (src[0].host) est = host_estimate(src[0].p, q.key_bytes, n);
      else if (tga_status e = estimate_groups(ctx, src[0].p, q.key_bytes, n, &est)) return e;
      est = std::max<uint64_t>(1, std::min<uint64_t>(est, n));
    }
  }
  if (deferred) plan = lean_plan(ctx, q, tile, std::min<uint64_t>(est, tile), opt);

  // Integer aggregates (at most as many as value columns a call takes): each
  // tile runs the whole operator into a buffer of partial groups -- whole
  // partitions written straight out of LDS, no global-table atomics -- and
  // the partials are merged by the same operator (COUNT and SUM summed, MIN
  // and MAX of the partials) whenever the buffer fills, and once at the end.
  if (partials) {
    QuerySpec qm;
    {
      tga_column mk{nullptr, q.key_dtype, TGA_DEVICE};
      tga_column mv[kMaxCols];
      tga_agg ma[kMaxCols];
      for (int a = 0; a < s0.nacc; ++a) {
        mv[a] = tga_column{nullptr, q.out_dtype[a], TGA_DEVICE};
        const int kind = s0.acc[a].kind;
        ma[a].op = (kind == ACC_MIN_I32 || kind == ACC_MIN_I64) ? OP_MIN
                   : (kind == ACC_MAX_I32 || kind == ACC_MAX_I64) ? OP_MAX : OP_SUM;
        ma[a].column = a;
      }
      const std::string e = build_spec(&mk, mv, s0.nacc, ma, s0.nacc, qm);
      if (!e.empty()) return fail(ctx, TGA_ERR_INVALID_ARG, "stream merge: %s", e.c_str());
    }
    uint64_t out_row = (uint64_t)q.key_bytes;
    for (int a = 0; a < s0.nacc; ++a) out_row += (uint64_t)dtype_bytes(q.out_dtype[a]);
    uint64_t pcap = 0;
    struct Part { DBuf keys, cols[kMaxCols]; };
    Part part[2];
    auto alloc_part = [&](Part &pt, uint64_t rows) -> bool {
      if (!pt.keys.alloc(ctx, rows * q.key_bytes)) return false;
      for (int a = 0; a < s0.nacc; ++a)
        if (!pt.cols[a].alloc(ctx, rows * (uint64_t)dtype_bytes(q.out_dtype[a]))) return false;
      return true;
    };
    int cur = 0;
    uint64_t off = 0;  // partial rows held in part[cur]
    // merges part[cur][0, off) into dst (ext buffers of capacity cap, or out)
    auto merge = [&](void *dk, void *const *dc, uint64_t cap, PlainOut *mo, uint64_t *mg) -> tga_status {
      QuerySpec qmm = qm;
      for (int a = 0; a < s0.nacc; ++a) qmm.schema.cols[a] = part[cur].cols[a].p;
      uint64_t em = std::min<uint64_t>(est, std::max<uint64_t>(off, 1));
      if (deferred)
        if (tga_status e = estimate_groups(ctx, part[cur].keys.p, q.key_bytes, off, &em)) return e;
      const Plan mp = lean_plan(ctx, qmm, off, em, nullptr);
      return run_plain(ctx, qmm, part[cur].keys.p, off, mp, mo, dk, dc, cap, mg);
    };
    auto ptrs = [&](Part &pt, uint64_t at, void **cols) {
      for (int a = 0; a < s0.nacc; ++a) cols[a] = (char *)pt.cols[a].p + at * dtype_bytes(q.out_dtype[a]);
      return (void *)((char *)pt.keys.p + at * q.key_bytes);
    };
    if (!staged0)
      if (tga_status e = issue_copy(0)) return e;
    for (uint64_t t = 0; t < ntiles; ++t) {
      TGA_HIP(ctx, hipStreamWaitEvent(st, copied[t & 1], 0));
      if (t + 1 < ntiles)
        if (tga_status e = issue_copy(t + 1)) return e;
      QuerySpec qt;
      const void *tkeys;
      uint64_t m;
      tile_cols(t, qt, tkeys, m);
      if (t == 0) {
        const uint64_t per_tile = std::min<uint64_t>(est, tile);
        pcap = std::max<uint64_t>(4 * per_tile, 1ULL << 16);
        pcap = std::min<uint64_t>(pcap, std::max<uint64_t>(2 * per_tile, ctx->arena->budget() / 16 / out_row));
        pcap = std::min<uint64_t>(pcap, n);
        if (!alloc_part(part[0], pcap) || !alloc_part(part[1], pcap))
          return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "streaming partials of %llu rows", (unsigned long long)pcap);
      }
      for (int tries = 0;; ++tries) {
        void *pc[kMaxCols];
        void *pk = ptrs(part[cur], off, pc);
        uint64_t gt = 0;
        const tga_status e = run_plain(ctx, qt, tkeys, m, plan, nullptr, pk, pc, pcap - off, &gt);
        if (e == TGA_OK) { off += gt; break; }
        if (e != TGA_ERR_CAPACITY || tries > 4) return e;
        if (off > 0) {
          // the partials so far, merged into the other buffer
          void *mc[kMaxCols];
          void *mk = ptrs(part[cur ^ 1], 0, mc);
          uint64_t gm = 0;
          if (tga_status e2 = merge(mk, mc, pcap, nullptr, &gm)) return e2;
          cur ^= 1;
          off = gm;
          if (off + std::min<uint64_t>(gt, m) <= pcap) continue;
        }
        // one tile's groups do not fit beside the merged ones: larger buffers
        const uint64_t ncap = std::min<uint64_t>(n, std::max<uint64_t>(2 * pcap, off + gt));
        Part grown;
        if (!alloc_part(grown, ncap)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "streaming partials");
        if (off) {
          TGA_HIP(ctx, hipMemcpyAsync(grown.keys.p, part[cur].keys.p, off * q.key_bytes, hipMemcpyDeviceToDevice, st));
          for (int a = 0; a < s0.nacc; ++a)
            TGA_HIP(ctx, hipMemcpyAsync(grown.cols[a].p, part[cur].cols[a].p, off * dtype_bytes(q.out_dtype[a]),
                                        hipMemcpyDeviceToDevice, st));
          TGA_HIP(ctx, hipStreamSynchronize(st));
        }
        part[cur] = std::move(grown);
        part[cur ^ 1] = Part{};
        if (!alloc_part(part[cur ^ 1], ncap)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "streaming partials");
        pcap = ncap;
      }
      if (ctx->prof.on) ctx->prof.resolve();
    }
    // the staging sets and the spare partial buffer are done with: their
    // memory goes to the final merge
    TGA_HIP(ctx, hipStreamSynchronize(cs));
    for (auto &sb : stage)
      for (auto &b : sb) b.reset();
    part[cur ^ 1] = Part{};
    const tga_status e = merge(ext_keys, ext_cols, ext_cap, out, groups);
    if (e && e != TGA_ERR_CAPACITY) return e;
    report(next_pow2_u64(std::max<uint64_t>(1024, 2 * est)) / 2, 1);
    return e;
  }

  // Exact fp64 SUM (or more aggregates than a merge takes): the tiles
  // aggregate into one global table, compacted at the end.
  uint64_t S = plan.gslots;
  S = std::max<uint64_t>(S, next_pow2_u64(std::max<uint64_t>(1024, 2 * est)));
  for (int attempt = 0; attempt < 5; ++attempt) {
    GTable g{};
    std::vector<DBuf> gbufs;
    DBuf flags;
    if (gtable_alloc(ctx, q, S, g, gbufs, flags) != hipSuccess)
      return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "global table of %llu slots", (unsigned long long)S);
    TGA_HIP(ctx, timed(ctx, "gtable_init", [&] { return launch_gtable_init(q.key_bytes, g, s0, st); }));
    if (tga_status e = issue_copy(0)) return e;
    for (uint64_t t = 0; t < ntiles; ++t) {
      const int b = (int)(t & 1);
      const uint64_t r0 = t * tile, m = std::min<uint64_t>(tile, n - r0);
      TGA_HIP(ctx, hipStreamWaitEvent(st, copied[b], 0));
      // the other set's tile was aggregated (synchronised below): refill it
      if (t + 1 < ntiles)
        if (tga_status e = issue_copy(t + 1)) return e;
      QuerySpec qt = q;
      for (int c = 0; c < nv; ++c)
        qt.schema.cols[c] = src[c + 1].host ? stage[b][c + 1].p : (const void *)(src[c + 1].p + r0 * src[c + 1].bytes);
      const void *tkeys = src[0].host ? stage[b][0].p : (const void *)(src[0].p + r0 * q.key_bytes);
      std::vector<DBuf> keep;
      if (tga_status e = aggregate_into(ctx, qt, tkeys, m, plan, g, keep)) return e;
      TGA_HIP(ctx, hipStreamSynchronize(st));
      if (ctx->prof.on) ctx->prof.resolve();
    }
    // compaction into the outputs
    const uint64_t cap = ext_keys ? ext_cap : S + 1;
    PlainOut tmp;
    OutDesc od{};
    od.capacity = cap;
    if (ext_keys) {
      od.keys = ext_keys;
      for (int a = 0; a < s0.nacc; ++a) od.cols[a] = ext_cols[a];
    } else {
      if (!tmp.keys.alloc(ctx, std::max<uint64_t>(cap, 1) * q.key_bytes)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "output");
      od.keys = tmp.keys.p;
      for (int a = 0; a < s0.nacc; ++a) {
        if (!tmp.cols[a].alloc(ctx, std::max<uint64_t>(cap, 1) * (uint64_t)dtype_bytes(q.out_dtype[a])))
          return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "output");
        od.cols[a] = tmp.cols[a].p;
      }
    }
    uint32_t hflags[kFlagWords];
    uint64_t hcount = 0;
    if (tga_status e = compact_into(ctx, q, g, od, hflags, &hcount)) return e;
    if (ctx->prof.on) ctx->prof.resolve();
    if (ext_keys && hcount > ext_cap) {
      *groups = hcount;
      return fail(ctx, TGA_ERR_CAPACITY, "%llu groups > capacity %llu", (unsigned long long)hcount,
                  (unsigned long long)ext_cap);
    }
    if (hflags[FLAG_OVERFLOW]) {
      S *= 4;  // the estimate was too low: a larger table, the tiles again
      continue;
    }
    if (hflags[FLAG_RANGE]) return fail(ctx, TGA_ERR_RANGE, "fp64 SUM input outside [2^-96, 2^64) or not finite");
    *groups = hcount;
    if (out) {
      out->keys = std::move(tmp.keys);
      for (int a = 0; a < s0.nacc; ++a) out->cols[a] = std::move(tmp.cols[a]);
      out->groups = hcount;
      out->capacity = cap;
    }
    report(S / 2, attempt + 1);
    return TGA_OK;
  }
  return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "global table kept overflowing");
}

}  // namespace tga
