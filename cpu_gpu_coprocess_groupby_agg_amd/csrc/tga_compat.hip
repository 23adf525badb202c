// tga_compat.hip -- the reference's two phase entry points over host buffers
// (u32 keys, u32 values, u32 wrap-around SUM), rebuilt on the gfx950 kernels,
// with the tiles (phase 1) and merged partitions (phase 2) spread over GPU
// lanes and, when asked, host worker threads.
//
//   tga_compat_partition        <- groupby_agg_partition<K,V>
//                                  T/src/group_then_partition.cu:427-528
//   tga_compat_intra_partition  <- groupby_agg_intra_partition
//                                  T/src/groupby_agg_intra_partition.cu:283-411
//   tga_partset_*               <- class par_result, T/include/par_result.cuh:11-82
//
// GPU lanes: the reference runs one host thread per CUDA stream, each taking
// tiles (gtp:466-499) or merged partitions (intra:344-383) in turn.  Here
// `nstreams` lanes (the context and nstreams - 1 lane contexts kept by it,
// each with its own HIP stream, arena and status words) run on their own
// threads, so one lane's copies overlap another's kernels.
//
// Phase 1 per tile (on a lane): host->device copy, the plain operator's
// pre-aggregation, then a hash partition of the tile's distinct rows into P
// partitions -- histogram, the project's decoupled look-back scan of the P
// counts, scatter -- and a device->host copy over the caller's buffers at the
// tile offset, exactly where the reference writes its tile result
// (gtp:413-418).  Phase 2 per merged partition: gather (intra:89-107), the
// reference's chunk loop with its overflow rule (intra:213-253) on the device
// table, and the compaction (intra:259-276).
//
// CPU co-processing (README.md:16, "cpu process partition only or
// groupby_agg_partition in phase_1"): host threads take tiles from the end of
// the input and merged partitions from the smallest, as tga_set_coprocess
// asks, doing the same unit of work with the same semantics (tga_cpu.cpp).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <queue>
#include <thread>

#include "tga_device.hpp"
#include "tga_host.hpp"

namespace tga {

// Partition of a distinct result row: upper hash bits modulo P (P need not be
// a power of two, as in the reference's default P = 300).
__device__ __forceinline__ uint32_t compat_partition(uint32_t key, uint32_t P) {
  return (uint32_t)(key_hash((uint64_t)key) >> 32) % P;
}

__global__ void k_compat_hist(const uint32_t *keys, uint64_t n, uint32_t P, uint32_t *cnt) {
  extern __shared__ uint32_t h[];
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) h[p] = 0;
  __syncthreads();
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&h[compat_partition(keys[i], P)], 1u);
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x)
    if (h[p]) atomicAdd(&cnt[p], h[p]);
}

// cursor[p] starts at the exclusive offset of partition p; the SUM leaves as
// its low 32 bits (the reference's u32 atomicAdd wraps, kernel.cu:48,152).
__global__ void k_compat_scatter(const uint32_t *keys, const uint64_t *sums, uint64_t n, uint32_t P, uint32_t *cursor,
                                 uint32_t *okeys, uint32_t *ovals) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t k = keys[i];
    const uint32_t d = atomicAdd(&cursor[compat_partition(k, P)], 1u);
    okeys[d] = k;
    ovals[d] = (uint32_t)sums[i];
  }
}

static unsigned grid_of(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(2048, (n + 255) / 256)); }

}  // namespace tga

using namespace tga;

struct tga_partset {
  tga_context *ctx = nullptr;
  struct Seg { uint32_t *k; uint32_t *v; uint32_t len; };
  struct Part { std::vector<Seg> seg; uint64_t size = 0; };
  std::vector<Part> parts;
};

#define TGA_API extern "C" __attribute__((visibility("default")))

namespace tga {

// The context and nstreams - 1 lane contexts on its device (created once,
// kept by the context, at most 8 in all).  A lane has its own stream and
// status words and allocates from the context's arena: one budget and one
// block cache per device, so a block one lane frees serves the next
// allocation of any other.
std::vector<tga_context *> lanes_of(tga_context *ctx, size_t nstreams, tga_status *st) {
  const size_t L = std::max<size_t>(1, std::min<size_t>(nstreams, 8));
  *st = TGA_OK;
  while (ctx->lanes.size() + 1 < L) {
    tga_context *l = nullptr;
    const tga_status e = tga_context_create(ctx->device, 1, &l);
    if (e) { *st = fail(ctx, e, "lane context"); break; }
    delete l->arena;
    l->arena = ctx->arena;
    l->owns_arena = false;
    ctx->lanes.push_back(l);
  }
  std::vector<tga_context *> v{ctx};
  for (size_t i = 0; i + 1 < L && i < ctx->lanes.size(); ++i) v.push_back(ctx->lanes[i]);
  return v;
}

}  // namespace tga

namespace {

using clk = std::chrono::steady_clock;

// The reference query: SUM over a u32 column (zero extended), u32 keys.
QuerySpec compat_spec() {
  tga_column key{nullptr, DT_U32, TGA_DEVICE};
  tga_column val{nullptr, DT_U32, TGA_DEVICE};
  tga_agg agg{OP_SUM, 0};
  QuerySpec q;
  build_spec(&key, &val, 1, &agg, 1, q);
  return q;
}


// Units [0, n) claimed from both ends: GPU lanes from the front, host
// workers from the back.  The host takes a unit only while its planned share
// allows (forced share) or while it finishes the unit before the GPU would
// finish the rest (planner).
struct Claims {
  std::mutex mu;
  int64_t lo = 0, hi = 0;
  int64_t cpu_target = -1;      // forced share: host units at most; -1 = planner
  int64_t cpu_rows = 0, gpu_rows = 0, cpu_units = 0, gpu_units = 0;
  double cpu_rate = 0;          // rows/s of one host thread
  double gpu_busy_s = 0;        // summed over lanes
  int gpu_lanes = 1;
  std::atomic<int> err{0};
  std::string err_text;

  int64_t take_front() {
    std::lock_guard<std::mutex> g(mu);
    if (err.load() || lo >= hi) return -1;
    return lo++;
  }
  // rows(i): rows of unit i
  template <typename Rows>
  int64_t take_back(Rows rows) {
    std::lock_guard<std::mutex> g(mu);
    if (err.load() || lo >= hi) return -1;
    const int64_t i = hi - 1;
    if (cpu_target >= 0) {
      if (cpu_units >= cpu_target) return -1;
    } else {
      // planner: this unit on one host thread, against the GPU lanes on
      // every unit left but this one, at the rate they have shown (an
      // optimistic 4e9 rows/s per lane before any unit finished)
      uint64_t rest = 0;
      for (int64_t j = lo; j < i; ++j) rest += rows(j);
      const double gpu_rate = gpu_units ? (double)gpu_rows / std::max(gpu_busy_s, 1e-9) * gpu_lanes
                                        : 4e9 * gpu_lanes;
      if (cpu_rate <= 0 || (double)rows(i) / cpu_rate >= (double)rest / gpu_rate) return -1;
    }
    --hi;
    ++cpu_units;
    cpu_rows += (int64_t)rows(i);
    return i;
  }
  void fail_with(tga_status st, const std::string &text) {
    std::lock_guard<std::mutex> g(mu);
    if (!err.load()) {
      err = (int)st;
      err_text = text;
    }
  }
};

// Rows per second of one host thread on u32 aggregation, from up to 2^16
// rows of the given keys/values (copied, so the input is untouched).
double calibrate_cpu(const uint32_t *k, const uint32_t *v, uint64_t m) {
  m = std::min<uint64_t>(m, 1 << 16);
  if (m == 0) return 0;
  std::vector<uint32_t> kk(k, k + m), vv(v, v + m);
  CpuAgg32 t;
  const auto a = clk::now();
  t.reset(std::min<uint64_t>(m, 1 << 12));
  t.add(kk.data(), vv.data(), m);
  const double s = std::chrono::duration<double>(clk::now() - a).count();
  return s > 0 ? (double)m / s : 0;
}

void record(tga_context *ctx, int phase, const Claims &c, double wall) {
  tga_coprocess_stats &s = ctx->cop[phase - 1];
  s.gpu_units = c.gpu_units;
  s.cpu_units = c.cpu_units;
  s.gpu_rows = c.gpu_rows;
  s.cpu_rows = c.cpu_rows;
  s.wall_s = wall;
  s.cpu_rows_per_s = c.cpu_rate;
  s.gpu_rows_per_s = c.gpu_busy_s > 0 ? (double)c.gpu_rows / c.gpu_busy_s * c.gpu_lanes : 0;
}

}  // namespace

TGA_API tga_status tga_set_coprocess(tga_context *ctx, int32_t cpu_threads, double cpu_share) {
  if (!ctx || cpu_threads < 0 || cpu_share > 1.0) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  ctx->cpu_threads = cpu_threads;
  ctx->cpu_share = cpu_share;
  return TGA_OK;
}

TGA_API tga_status tga_last_coprocess(tga_context *ctx, int32_t phase, tga_coprocess_stats *st) {
  if (!ctx || !st || phase < 1 || phase > 2) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  *st = ctx->cop[phase - 1];
  return TGA_OK;
}

TGA_API tga_status tga_partset_create(tga_context *ctx, size_t n, tga_partset **out) {
  if (!ctx || !out || n == 0) return TGA_ERR_INVALID_ARG;
  auto *ps = new tga_partset();
  ps->ctx = ctx;
  ps->parts.resize(n);
  *out = ps;
  return TGA_OK;
}

TGA_API tga_status tga_partset_push(tga_partset *ps, size_t p, uint32_t *keys, uint32_t *vals, uint32_t len) {
  if (!ps || p >= ps->parts.size() || (len && (!keys || !vals))) return TGA_ERR_INVALID_ARG;
  ps->parts[p].seg.push_back({keys, vals, len});
  ps->parts[p].size += len;
  return TGA_OK;
}

TGA_API tga_status tga_partset_count(const tga_partset *ps, size_t *n) {
  if (!ps || !n) return TGA_ERR_INVALID_ARG;
  *n = ps->parts.size();
  return TGA_OK;
}

TGA_API tga_status tga_partset_info(const tga_partset *ps, size_t p, size_t *nseg, uint64_t *rows) {
  if (!ps || p >= ps->parts.size()) return TGA_ERR_INVALID_ARG;
  if (nseg) *nseg = ps->parts[p].seg.size();
  if (rows) *rows = ps->parts[p].size;
  return TGA_OK;
}

TGA_API tga_status tga_partset_segment(const tga_partset *ps, size_t p, size_t j, uint32_t **keys, uint32_t **vals,
                                       uint32_t *len) {
  if (!ps || p >= ps->parts.size() || j >= ps->parts[p].seg.size()) return TGA_ERR_INVALID_ARG;
  const auto &s = ps->parts[p].seg[j];
  if (keys) *keys = s.k;
  if (vals) *vals = s.v;
  if (len) *len = s.len;
  return TGA_OK;
}

TGA_API tga_status tga_partset_destroy(tga_partset *ps) {
  if (!ps) return TGA_ERR_INVALID_ARG;
  delete ps;
  return TGA_OK;
}

TGA_API tga_status tga_host_free(void *p) {
  if (p && hipHostFree(p) != hipSuccess) return TGA_ERR_HIP;
  return TGA_OK;
}

// ---------------------------------------------------------------------------
// phase 1
// ---------------------------------------------------------------------------
namespace {

// One GPU lane of phase 1: its device buffers, reused for every tile it takes.
struct P1Lane {
  tga_context *L = nullptr;
  DBuf dk, dv, pk, pv, cnt, offs, desc;
};

tga_status p1_lane_init(P1Lane &w, size_t tl, size_t P) {
  tga_context *L = w.L;
  if (!w.dk.alloc(L, tl * 4) || !w.dv.alloc(L, tl * 4) || !w.pk.alloc(L, tl * 4) || !w.pv.alloc(L, tl * 4) ||
      !w.cnt.alloc(L, (P + 1) * 4) || !w.offs.alloc(L, (P + 1) * 4) || !w.desc.alloc(L, (scan_tiles(P + 1) + 2) * 8))
    return fail(L, TGA_ERR_OUT_OF_MEMORY, "compat partition buffers");
  if (P * 4 > 65536)
    TGA_HIP(L, hipFuncSetAttribute((const void *)&k_compat_hist, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(P * 4)));
  return TGA_OK;
}

// Tile [b, b + m) of the caller's buffers on GPU lane w: offs[0..P] of its
// partitions in the tile result (written back at the tile offset).
tga_status p1_gpu_tile(P1Lane &w, uint32_t *hk, uint32_t *hv, size_t b, size_t m, size_t P, const QuerySpec &q,
                       uint32_t *offs) {
  tga_context *L = w.L;
  hipStream_t st = L->stream;
  TGA_HIP(L, hipMemcpyAsync(w.dk.p, hk + b, m * 4, hipMemcpyHostToDevice, st));
  TGA_HIP(L, hipMemcpyAsync(w.dv.p, hv + b, m * 4, hipMemcpyHostToDevice, st));
  QuerySpec qt = q;
  qt.schema.cols[0] = w.dv.p;
  uint64_t est = 0;
  if (tga_status e = estimate_groups(L, w.dk.p, 4, m, &est)) return e;
  const Plan plan = autotune(L, qt, m, est, nullptr);
  PlainOut po;
  uint64_t G = 0;
  if (tga_status e = run_plain(L, qt, w.dk.p, m, plan, &po, nullptr, nullptr, 0, &G)) return e;
  // hash-partition the G distinct rows: histogram, the look-back scan of the
  // P + 1 counts (the last, zero, becomes the total), scatter
  const uint64_t tiles = scan_tiles(P + 1);
  TGA_HIP(L, launch_fill_u32(w.cnt.as<uint32_t>(), P + 1, 0, st));
  TGA_HIP(L, launch_fill_u64(w.desc.as<uint64_t>(), tiles + 1, 0, st));
  hipLaunchKernelGGL(k_compat_hist, dim3(grid_of(G)), dim3(256), P * 4, st, po.keys.as<uint32_t>(), G, (uint32_t)P,
                     w.cnt.as<uint32_t>());
  TGA_HIP(L, hipGetLastError());
  TGA_HIP(L, launch_scan_u32(w.cnt.as<uint32_t>(), w.offs.as<uint32_t>(), P + 1, w.desc.as<uint64_t>(),
                             (uint32_t *)(w.desc.as<uint64_t>() + tiles), st));
  TGA_HIP(L, hipMemcpyAsync(offs, w.offs.p, (P + 1) * 4, hipMemcpyDeviceToHost, st));
  hipLaunchKernelGGL(k_compat_scatter, dim3(grid_of(G)), dim3(256), 0, st, po.keys.as<uint32_t>(),
                     po.cols[0].as<uint64_t>(), G, (uint32_t)P, w.offs.as<uint32_t>(), w.pk.as<uint32_t>(),
                     w.pv.as<uint32_t>());
  TGA_HIP(L, hipGetLastError());
  // the tile result overwrites the caller's buffers at the tile offset
  TGA_HIP(L, hipMemcpyAsync(hk + b, w.pk.p, G * 4, hipMemcpyDeviceToHost, st));
  TGA_HIP(L, hipMemcpyAsync(hv + b, w.pv.p, G * 4, hipMemcpyDeviceToHost, st));
  TGA_HIP(L, hipStreamSynchronize(st));
  if (offs[P] != G)
    return fail(L, TGA_ERR_HIP, "compat partition: %u rows scattered of %llu", offs[P], (unsigned long long)G);
  return TGA_OK;
}

}  // namespace

TGA_API tga_status tga_compat_partition(tga_context *ctx, uint32_t *host_keys, uint32_t *host_vals, size_t n,
                                        size_t tile_len, size_t P, size_t nstreams, tga_partset **out) {
  if (!ctx || !out) return TGA_ERR_INVALID_ARG;
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  if (tile_len == 0 || P == 0 || P > 65536 || nstreams == 0 || (n && (!host_keys || !host_vals)))
    return fail(ctx, TGA_ERR_INVALID_ARG, "compat partition: bad tile_len/P/nstreams or null buffers");
  if (tile_len >= (1ULL << 31)) return fail(ctx, TGA_ERR_INVALID_ARG, "tile_len must be < 2^31");
  // the partition histogram is one LDS word per partition
  if (P * 4 > ctx->lds_max)
    return fail(ctx, TGA_ERR_INVALID_ARG, "P = %zu partitions exceed the LDS histogram (at most %zu)", P,
                ctx->lds_max / 4);
  hipSetDevice(ctx->device);
  const auto t0 = clk::now();
  const QuerySpec q = compat_spec();
  const size_t tl = std::min(tile_len, std::max<size_t>(n, 1));
  const int64_t tiles = (int64_t)((n + tile_len - 1) / tile_len);
  tga_status lst = TGA_OK;
  const std::vector<tga_context *> gl = lanes_of(ctx, std::min<size_t>(nstreams, std::max<int64_t>(tiles, 1)), &lst);
  if (lst) return lst;
  std::vector<std::vector<uint32_t>> toffs((size_t)tiles);
  auto rows_of = [&](int64_t t) -> uint64_t { return std::min(n, (size_t)(t + 1) * tile_len) - (size_t)t * tile_len; };
  Claims cl;
  cl.hi = tiles;
  cl.gpu_lanes = (int)gl.size();
  const int cpu_threads = ctx->cpu_share != 0.0 ? ctx->cpu_threads : 0;
  if (cpu_threads > 0 && tiles > 0) {
    if (ctx->cpu_share > 0) {
      cl.cpu_target = (int64_t)std::llround(ctx->cpu_share * (double)tiles);
    } else {
      const size_t b = (size_t)(tiles - 1) * tile_len;
      cl.cpu_rate = calibrate_cpu(host_keys + b, host_vals + b, rows_of(tiles - 1));
    }
  }
  std::vector<P1Lane> lanes(gl.size());
  for (size_t i = 0; i < gl.size(); ++i) {
    lanes[i].L = gl[i];
    std::lock_guard<std::recursive_mutex> lg(gl[i]->mu);
    hipSetDevice(gl[i]->device);
    if (tga_status e = p1_lane_init(lanes[i], tl, P)) return i ? fail(ctx, e, "%s", gl[i]->err.c_str()) : e;
  }
  auto gpu_worker = [&](P1Lane &w) {
    hipSetDevice(w.L->device);
    std::unique_lock<std::recursive_mutex> lg(w.L->mu, std::defer_lock);
    if (w.L != ctx) lg.lock();
    for (;;) {
      const int64_t t = cl.take_front();
      if (t < 0) break;
      std::vector<uint32_t> offs(P + 1);
      const auto a = clk::now();
      const tga_status e = p1_gpu_tile(w, host_keys, host_vals, (size_t)t * tile_len, rows_of(t), P, q, offs.data());
      if (e) { cl.fail_with(e, w.L->err); break; }
      const double s = std::chrono::duration<double>(clk::now() - a).count();
      std::lock_guard<std::mutex> g(cl.mu);
      toffs[(size_t)t] = std::move(offs);
      ++cl.gpu_units;
      cl.gpu_rows += (int64_t)rows_of(t);
      cl.gpu_busy_s += s;
    }
  };
  auto cpu_worker = [&]() {
    for (;;) {
      const int64_t t = cl.take_back(rows_of);
      if (t < 0) break;
      std::vector<uint32_t> offs(P + 1);
      cpu_compat_tile(host_keys + (size_t)t * tile_len, host_vals + (size_t)t * tile_len, rows_of(t), (uint32_t)P,
                      offs.data());
      std::lock_guard<std::mutex> g(cl.mu);
      toffs[(size_t)t] = std::move(offs);
    }
  };
  {
    std::vector<std::thread> th;
    for (size_t i = 1; i < lanes.size(); ++i) th.emplace_back(gpu_worker, std::ref(lanes[i]));
    for (int i = 0; i < cpu_threads; ++i) th.emplace_back(cpu_worker);
    gpu_worker(lanes[0]);  // the calling thread drives the context's own stream
    for (auto &x : th) x.join();
  }
  if (cl.err.load()) return fail(ctx, (tga_status)cl.err.load(), "%s", cl.err_text.c_str());
  // update_par_result_task (par_result.cu:3-13), tile by tile in input order
  auto *ps = new tga_partset();
  ps->ctx = ctx;
  ps->parts.resize(P);
  for (int64_t t = 0; t < tiles; ++t) {
    const std::vector<uint32_t> &o = toffs[(size_t)t];
    const size_t b = (size_t)t * tile_len;
    for (size_t p = 0; p < P; ++p) {
      const uint32_t len = o[p + 1] - o[p];
      ps->parts[p].seg.push_back({host_keys + b + o[p], host_vals + b + o[p], len});
      ps->parts[p].size += len;
    }
  }
  record(ctx, 1, cl, std::chrono::duration<double>(clk::now() - t0).count());
  *out = ps;
  return TGA_OK;
}

// ---------------------------------------------------------------------------
// phase 2
// ---------------------------------------------------------------------------
namespace {
struct MergedPart {
  std::vector<tga_partset::Seg> seg;
  uint64_t size = 0;
};

// merge_partition (intra:13-48) with the reference's std::priority_queue and
// comparator, so equal sizes break ties the same way.
std::vector<MergedPart> merge_partition(const std::vector<tga_partset::Part> &in, uint64_t K) {
  auto cmp = [](const MergedPart &l, const MergedPart &r) { return l.size > r.size; };
  std::vector<MergedPart> init;
  for (const auto &p : in) init.push_back({p.seg, p.size});
  std::priority_queue<MergedPart, std::vector<MergedPart>, decltype(cmp)> H(init.begin(), init.end(), cmp);
  for (;;) {
    if (H.size() == 1) break;
    MergedPart tmp = H.top();
    H.pop();
    MergedPart top = H.top();
    if (tmp.size + top.size > K) { H.push(tmp); break; }
    do {
      tmp.seg.insert(tmp.seg.end(), top.seg.begin(), top.seg.end());
      tmp.size += top.size;
      H.pop();
      if (H.size() == 0) break;
      top = H.top();
    } while (top.size + tmp.size <= K);
    H.push(tmp);
  }
  std::vector<MergedPart> rv;
  while (!H.empty()) { rv.push_back(H.top()); H.pop(); }
  return rv;
}

// One GPU lane of phase 2: a device table of at most max_load groups and the
// chunk and result buffers, reused for every partition it takes.
struct P2Lane {
  tga_context *L = nullptr;
  GTable g{};
  std::vector<DBuf> gbufs;
  DBuf flags, dk, dv, rk, rs, gcount;
  std::vector<uint64_t> hs;
};

// par_result_in_continous_mem (intra:89-107): the segments one after another
// into the output slice
void gather(const MergedPart &m, uint32_t *ok, uint32_t *ov, size_t at) {
  size_t w = at;
  for (const auto &s : m.seg) {
    std::memcpy(ok + w, s.k, (size_t)s.len * 4);
    std::memcpy(ov + w, s.v, (size_t)s.len * 4);
    w += s.len;
  }
}

tga_status p2_gpu_part(P2Lane &w, const QuerySpec &q, uint32_t *ok, uint32_t *ov, size_t begin, uint64_t size,
                       uint64_t S, size_t min_load, size_t max_load, uint64_t *groups) {
  tga_context *L = w.L;
  hipStream_t st = L->stream;
  TGA_HIP(L, launch_gtable_init(4, w.g, q.schema, st));
  uint64_t rest = size, already = 0, tail = begin + size;
  while (rest > 0) {
    const uint64_t next = max_load - std::min<uint64_t>(already, max_load);
    uint64_t take;
    if (rest <= next) take = rest;
    else {
      if (already > min_load || next == 0) return fail(L, TGA_ERR_TABLE_OVERFLOW, "hash table overflow");
      take = next;
    }
    // pop2device: the chunk comes from the back (par_result.cuh:48-74)
    TGA_HIP(L, hipMemcpyAsync(w.dk.p, ok + tail - take, take * 4, hipMemcpyHostToDevice, st));
    TGA_HIP(L, hipMemcpyAsync(w.dv.p, ov + tail - take, take * 4, hipMemcpyHostToDevice, st));
    QuerySpec qc = q;
    qc.schema.cols[0] = w.dv.p;
    tga_options o{};
    o.path = TGA_PATH_AUTO;
    Plan plan = autotune(L, qc, take, std::min<uint64_t>(take, max_load), &o);
    plan.gslots = S;
    std::vector<DBuf> keep;
    if (tga_status s2 = aggregate_into(L, qc, w.dk.p, take, plan, w.g, keep)) return s2;
    uint32_t hf[kFlagWords];
    TGA_HIP(L, launch_fill_u64(w.gcount.as<uint64_t>(), 1, 0, st));
    TGA_HIP(L, launch_count_groups(4, w.g, w.gcount.as<uint64_t>(), st));
    TGA_HIP(L, hipMemcpyAsync(hf, w.g.flags, sizeof hf, hipMemcpyDeviceToHost, st));
    TGA_HIP(L, hipMemcpyAsync(&already, w.gcount.p, 8, hipMemcpyDeviceToHost, st));
    TGA_HIP(L, hipStreamSynchronize(st));
    if (hf[FLAG_OVERFLOW]) return fail(L, TGA_ERR_TABLE_OVERFLOW, "hash table overflow");
    tail -= take;
    rest -= take;
  }
  OutDesc od{};
  od.keys = w.rk.p;
  od.cols[0] = w.rs.p;
  od.capacity = S + 1;
  uint32_t hf[kFlagWords];
  uint64_t G = 0;
  if (tga_status s2 = compact_into(L, q, w.g, od, hf, &G)) return s2;
  w.hs.resize(G);
  TGA_HIP(L, hipMemcpyAsync(ok + begin, w.rk.p, G * 4, hipMemcpyDeviceToHost, st));
  TGA_HIP(L, hipMemcpyAsync(w.hs.data(), w.rs.p, G * 8, hipMemcpyDeviceToHost, st));
  TGA_HIP(L, hipStreamSynchronize(st));
  for (uint64_t j = 0; j < G; ++j) ov[begin + j] = (uint32_t)w.hs[j];
  *groups = G;
  return TGA_OK;
}
}  // namespace

TGA_API tga_status tga_compat_intra_partition(tga_context *ctx, tga_partset *ps, size_t capacity, size_t min_load,
                                              size_t max_load, size_t nstreams, uint32_t **out_keys,
                                              uint32_t **out_vals, size_t *par_kv_begin, size_t *par_result_kv_num,
                                              size_t *n_merged) {
  if (!ctx || !ps || !out_keys || !out_vals || !par_kv_begin || !par_result_kv_num || !n_merged)
    return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  if (capacity == 0 || max_load == 0 || min_load > max_load || max_load > capacity || nstreams == 0)
    return fail(ctx, TGA_ERR_INVALID_ARG, "compat intra: need 0 <= min_load <= max_load <= capacity, capacity > 0");
  hipSetDevice(ctx->device);
  const auto t0 = clk::now();
  std::vector<MergedPart> mp = merge_partition(ps->parts, min_load);
  const size_t np = mp.size();
  std::vector<size_t> begin(np);
  size_t total = 0;
  for (size_t i = 0; i < np; ++i) { begin[i] = total; total += mp[i].size; }
  uint32_t *ok = nullptr, *ov = nullptr;
  if (hipHostMalloc((void **)&ok, std::max<size_t>(total, 1) * 4) != hipSuccess ||
      hipHostMalloc((void **)&ov, std::max<size_t>(total, 1) * 4) != hipSuccess) {
    if (ok) hipHostFree(ok);
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "pinned outputs");
  }
  auto fail_free = [&](tga_status s) { hipHostFree(ok); hipHostFree(ov); return s; };
  const QuerySpec q = compat_spec();
  // device table: at most max_load groups ever live in it (intra:215-236)
  uint64_t S = 1024;
  while (S < 2 * (uint64_t)max_load) S <<= 1;
  tga_status lst = TGA_OK;
  const std::vector<tga_context *> gl = lanes_of(ctx, std::min(nstreams, std::max<size_t>(np, 1)), &lst);
  if (lst) return fail_free(lst);
  std::vector<P2Lane> lanes(gl.size());
  for (size_t i = 0; i < gl.size(); ++i) {
    P2Lane &w = lanes[i];
    w.L = gl[i];
    std::lock_guard<std::recursive_mutex> lg(w.L->mu);
    hipSetDevice(w.L->device);
    if (gtable_alloc(w.L, q, S, w.g, w.gbufs, w.flags) != hipSuccess || !w.dk.alloc(w.L, (uint64_t)max_load * 4) ||
        !w.dv.alloc(w.L, (uint64_t)max_load * 4) || !w.rk.alloc(w.L, (S + 1) * 4) || !w.rs.alloc(w.L, (S + 1) * 8) ||
        !w.gcount.alloc(w.L, 8))
      return fail_free(fail(ctx, TGA_ERR_OUT_OF_MEMORY, "compat intra buffers"));
  }
  // merge_partition returns the partitions smallest first: the GPU takes
  // them from the largest, host workers from the smallest
  auto rows_of = [&](int64_t u) -> uint64_t { return mp[(size_t)(np - 1 - (size_t)u)].size; };
  Claims cl;
  cl.hi = (int64_t)np;
  cl.gpu_lanes = (int)gl.size();
  const int cpu_threads = ctx->cpu_share != 0.0 ? ctx->cpu_threads : 0;
  if (cpu_threads > 0 && np > 0) {
    if (ctx->cpu_share > 0) {
      // the smallest partitions, up to the share of the rows
      uint64_t acc = 0;
      int64_t k = 0;
      while (k < (int64_t)np && (double)(acc + mp[(size_t)k].size) <= ctx->cpu_share * (double)total + 0.5) {
        acc += mp[(size_t)k].size;
        ++k;
      }
      cl.cpu_target = k;
    } else {
      std::vector<uint32_t> sk, sv;
      for (const auto &s : mp[0].seg) {
        if (sk.size() >= (1u << 16)) break;
        sk.insert(sk.end(), s.k, s.k + s.len);
        sv.insert(sv.end(), s.v, s.v + s.len);
      }
      cl.cpu_rate = calibrate_cpu(sk.data(), sv.data(), sk.size());
    }
  }
  auto gpu_worker = [&](P2Lane &w) {
    hipSetDevice(w.L->device);
    std::unique_lock<std::recursive_mutex> lg(w.L->mu, std::defer_lock);
    if (w.L != ctx) lg.lock();
    for (;;) {
      const int64_t u = cl.take_front();
      if (u < 0) break;
      const size_t i = np - 1 - (size_t)u;
      const auto a = clk::now();
      gather(mp[i], ok, ov, begin[i]);
      uint64_t G = 0;
      const tga_status e = p2_gpu_part(w, q, ok, ov, begin[i], mp[i].size, S, min_load, max_load, &G);
      if (e) { cl.fail_with(e, w.L->err); break; }
      const double s = std::chrono::duration<double>(clk::now() - a).count();
      std::lock_guard<std::mutex> g(cl.mu);
      par_kv_begin[i] = begin[i];
      par_result_kv_num[i] = G;
      ++cl.gpu_units;
      cl.gpu_rows += (int64_t)mp[i].size;
      cl.gpu_busy_s += s;
    }
  };
  auto cpu_worker = [&]() {
    for (;;) {
      const int64_t u = cl.take_back(rows_of);
      if (u < 0) break;
      const size_t i = np - 1 - (size_t)u;
      gather(mp[i], ok, ov, begin[i]);
      uint64_t G = 0;
      const tga_status e = cpu_compat_intra(ok + begin[i], ov + begin[i], mp[i].size, min_load, max_load, &G);
      if (e) { cl.fail_with(e, "hash table overflow"); break; }
      std::lock_guard<std::mutex> g(cl.mu);
      par_kv_begin[i] = begin[i];
      par_result_kv_num[i] = G;
    }
  };
  {
    std::vector<std::thread> th;
    for (size_t i = 1; i < lanes.size(); ++i) th.emplace_back(gpu_worker, std::ref(lanes[i]));
    for (int i = 0; i < cpu_threads; ++i) th.emplace_back(cpu_worker);
    gpu_worker(lanes[0]);
    for (auto &x : th) x.join();
  }
  if (cl.err.load()) return fail_free(fail(ctx, (tga_status)cl.err.load(), "%s", cl.err_text.c_str()));
  record(ctx, 2, cl, std::chrono::duration<double>(clk::now() - t0).count());
  // the merged list replaces the caller's list and is drained (intra:294, pop2device)
  ps->parts.assign(np, tga_partset::Part());
  *out_keys = ok;
  *out_vals = ov;
  *n_merged = np;
  return TGA_OK;
}
