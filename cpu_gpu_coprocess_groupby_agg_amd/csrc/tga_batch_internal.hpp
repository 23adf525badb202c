// tga_batch_internal.hpp -- batched persistent execution shared by the batch
// C ABI and the serving loop.
#pragma once

#include <array>
#include <vector>

#include "tga_host.hpp"

namespace tga {

enum ItemType : uint32_t { IT_INIT = 0, IT_COUNT = 1, IT_DIRECT = 2, IT_SCAN = 3, IT_SCATTER = 4, IT_PAGG = 5,
                           IT_COMPACT = 6, kItemTypes = 7 };

struct BQuery {
  Schema s;
  GTable g;
  Layout L;
  AccProg ap;
  OutDesc od;
  const void *keys;
  uint64_t n;
  int32_t key_bytes;
  int32_t path;
  int32_t pbits;
  int32_t nbits;
  int32_t max_probe;
  uint32_t ntiles;     // partition tiles
  uint32_t tile_rows;
  uint32_t pad0;
  uint32_t *cnt;       // [P][ntiles]
  uint32_t *offs;      // [P][ntiles]
  uint64_t *pstart;    // [P+1]
  void *pkeys;
  void *pcols[kMaxCols];
  uint64_t *count_out; // groups written
  uint32_t *done;      // [kItemTypes] completed items per type
  uint32_t need[kItemTypes];
};

struct BItem {
  uint32_t q;
  uint32_t type;
  uint64_t a, b;
};

struct BatchExec {
  tga_context *ctx = nullptr;
  int32_t nq = 0;
  std::vector<QuerySpec> spec;
  std::vector<uint64_t> rows;
  std::vector<const void *> keys;
  std::vector<uint64_t> est;
  std::vector<BQuery> hq;
  DBuf dq, ditems, ddone, dcounts, dticket, dflags, dprof;
  std::vector<DBuf> bufs;                          // tables and partition buffers
  std::vector<DBuf> okeys;                         // per query output keys
  std::vector<std::array<DBuf, kMaxAccs>> ocols;   // per query output columns
  std::vector<uint64_t> cap;
  // 1: okeys/ocols of the query are the buffers the batch's kernels write
  // (referenced by the device query records and the captured graph); 0: they
  // hold a rerun's result and nothing on the device points at them
  std::vector<uint8_t> graph_out;
  uint32_t nitems = 0;
  size_t lds = 0;
  unsigned grid = 0;
  const void *kernel = nullptr;  // k_batch variant for this batch
  uint64_t *h_counts = nullptr;  // pinned: group count per query
  uint32_t *h_flags = nullptr;   // pinned: status words per query
  void *h_stage = nullptr;       // pinned: query records + work items, when uploaded without a wait
  uint64_t stage_bytes = 0;
  ~BatchExec();
};

// Validates the queries and computes the cardinality estimates.
tga_status batch_prepare(tga_context *ctx, const tga_query *qs, int32_t n, BatchExec &b);
// Allocates tables/outputs and uploads the work queue: on ctx->stream, waited
// for, or (upload != null) from pinned staging on `upload`, not waited for --
// the caller then enqueues the batch on that same stream.
tga_status batch_build(BatchExec &b, const tga_query *qs, hipStream_t upload = nullptr);
// Device bytes batch_build takes from the arena for one query (admission).
uint64_t batch_query_bytes(const tga_context *ctx, const QuerySpec &q, uint64_t n, uint64_t est);
// Queues reset + persistent kernel + status copies on `st`.
tga_status batch_enqueue(BatchExec &b, hipStream_t st);
// After `st` finished: reruns overflowed queries with the plain operator on
// ctx->stream (results replace that query's outputs).
tga_status batch_finalize(BatchExec &b, std::vector<uint64_t> &groups);

}  // namespace tga
