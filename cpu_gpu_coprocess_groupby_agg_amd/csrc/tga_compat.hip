// tga_compat.hip -- the reference's two phase entry points over host buffers
// (u32 keys, u32 values, u32 wrap-around SUM), rebuilt on the gfx950 kernels.
//
//   tga_compat_partition        <- groupby_agg_partition<K,V>
//                                  T/src/group_then_partition.cu:427-528
//   tga_compat_intra_partition  <- groupby_agg_intra_partition
//                                  T/src/groupby_agg_intra_partition.cu:283-411
//   tga_partset_*               <- class par_result, T/include/par_result.cuh:11-82
//
// Phase 1 per tile: host->device copy, the plain operator's pre-aggregation
// (k_agg into a global table, k_compact), then a hash partition of the tile's
// distinct rows into P partitions (k_compat_hist + host prefix + scatter) and
// a device->host copy over the caller's buffers at the tile offset, exactly
// where the reference writes its tile result (gtp:413-418).
// Phase 2 per merged partition: gather (intra:89-107), the reference's chunk
// loop with its overflow rule (intra:213-253) on the device table, and the
// compaction (intra:259-276).
#include <algorithm>
#include <queue>

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

namespace {

// The reference query: SUM over a u32 column (zero extended), u32 keys.
QuerySpec compat_spec() {
  tga_column key{nullptr, DT_U32, TGA_DEVICE};
  tga_column val{nullptr, DT_U32, TGA_DEVICE};
  tga_agg agg{OP_SUM, 0};
  QuerySpec q;
  build_spec(&key, &val, 1, &agg, 1, q);
  return q;
}

}  // namespace

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
  hipStream_t st = ctx->stream;
  QuerySpec q = compat_spec();
  const size_t tl = std::min(tile_len, std::max<size_t>(n, 1));
  DBuf dk, dv, pk, pv, cnt;
  if (!dk.alloc(ctx, tl * 4) || !dv.alloc(ctx, tl * 4) || !pk.alloc(ctx, tl * 4) || !pv.alloc(ctx, tl * 4) ||
      !cnt.alloc(ctx, (P + 1) * 4))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "compat partition buffers");
  auto *ps = new tga_partset();
  ps->ctx = ctx;
  ps->parts.resize(P);
  std::vector<uint32_t> hc(P), base(P + 1);
  const size_t tiles = (n + tile_len - 1) / tile_len;
  tga_status rc = TGA_OK;
  for (size_t t = 0; t < tiles && rc == TGA_OK; ++t) {
    const size_t b = t * tile_len, m = std::min(n, b + tile_len) - b;
    auto step = [&]() -> tga_status {
      TGA_HIP(ctx, hipMemcpyAsync(dk.p, host_keys + b, m * 4, hipMemcpyHostToDevice, st));
      TGA_HIP(ctx, hipMemcpyAsync(dv.p, host_vals + b, m * 4, hipMemcpyHostToDevice, st));
      QuerySpec qt = q;
      qt.schema.cols[0] = dv.p;
      uint64_t est = 0;
      tga_status s2 = estimate_groups(ctx, dk.p, 4, m, &est);
      if (s2) return s2;
      Plan plan = autotune(ctx, qt, m, est, nullptr);
      PlainOut po;
      uint64_t G = 0;
      s2 = run_plain(ctx, qt, dk.p, m, plan, &po, nullptr, nullptr, 0, &G);
      if (s2) return s2;
      // hash-partition the G distinct rows of the tile
      TGA_HIP(ctx, launch_fill_u32(cnt.as<uint32_t>(), P, 0, st));
      hipLaunchKernelGGL(k_compat_hist, dim3(grid_of(G)), dim3(256), P * 4, st, po.keys.as<uint32_t>(), G,
                         (uint32_t)P, cnt.as<uint32_t>());
      TGA_HIP(ctx, hipGetLastError());
      TGA_HIP(ctx, hipMemcpyAsync(hc.data(), cnt.p, P * 4, hipMemcpyDeviceToHost, st));
      TGA_HIP(ctx, hipStreamSynchronize(st));
      uint32_t run = 0;
      for (size_t p = 0; p < P; ++p) { base[p] = run; run += hc[p]; }
      base[P] = run;
      TGA_HIP(ctx, hipMemcpyAsync(cnt.p, base.data(), P * 4, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_compat_scatter, dim3(grid_of(G)), dim3(256), 0, st, po.keys.as<uint32_t>(),
                         po.cols[0].as<uint64_t>(), G, (uint32_t)P, cnt.as<uint32_t>(), pk.as<uint32_t>(),
                         pv.as<uint32_t>());
      TGA_HIP(ctx, hipGetLastError());
      // the tile result overwrites the caller's buffers at the tile offset
      TGA_HIP(ctx, hipMemcpyAsync(host_keys + b, pk.p, G * 4, hipMemcpyDeviceToHost, st));
      TGA_HIP(ctx, hipMemcpyAsync(host_vals + b, pv.p, G * 4, hipMemcpyDeviceToHost, st));
      TGA_HIP(ctx, hipStreamSynchronize(st));
      for (size_t p = 0; p < P; ++p) {  // update_par_result_task, par_result.cu:3-13
        const uint32_t len = base[p + 1] - base[p];
        ps->parts[p].seg.push_back({host_keys + b + base[p], host_vals + b + base[p], len});
        ps->parts[p].size += len;
      }
      return TGA_OK;
    };
    rc = step();
  }
  if (rc) { delete ps; return rc; }
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
  hipStream_t st = ctx->stream;
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
  QuerySpec q = compat_spec();
  // device table: at most max_load groups ever live in it (intra:215-236)
  uint64_t S = 1024;
  while (S < 2 * (uint64_t)max_load) S <<= 1;
  GTable g{};
  std::vector<DBuf> gbufs;
  DBuf flags, dk, dv, rk, rs, gcount;
  auto fail_free = [&](tga_status s) { hipHostFree(ok); hipHostFree(ov); return s; };
  if (gtable_alloc(ctx, q, S, g, gbufs, flags) != hipSuccess || !dk.alloc(ctx, (uint64_t)max_load * 4) ||
      !dv.alloc(ctx, (uint64_t)max_load * 4) || !rk.alloc(ctx, (S + 1) * 4) || !rs.alloc(ctx, (S + 1) * 8) ||
      !gcount.alloc(ctx, 8))
    return fail_free(fail(ctx, TGA_ERR_OUT_OF_MEMORY, "compat intra buffers"));
  std::vector<uint64_t> hs;
  for (size_t i = 0; i < np; ++i) {
    MergedPart &m = mp[i];
    // par_result_in_continous_mem: gather the segments into the output slice
    size_t w = begin[i];
    for (const auto &s : m.seg) {
      std::memcpy(ok + w, s.k, (size_t)s.len * 4);
      std::memcpy(ov + w, s.v, (size_t)s.len * 4);
      w += s.len;
    }
    auto run = [&]() -> tga_status {
      TGA_HIP(ctx, launch_gtable_init(4, g, q.schema, st));
      uint64_t rest = m.size, already = 0, tail = begin[i] + m.size;
      while (rest > 0) {
        const uint64_t next = max_load - already;
        uint64_t take;
        if (rest <= next) take = rest;
        else {
          if (already > min_load) return fail(ctx, TGA_ERR_TABLE_OVERFLOW, "hash table overflow");
          take = next;
        }
        // pop2device: the chunk comes from the back (par_result.cuh:48-74)
        TGA_HIP(ctx, hipMemcpyAsync(dk.p, ok + tail - take, take * 4, hipMemcpyHostToDevice, st));
        TGA_HIP(ctx, hipMemcpyAsync(dv.p, ov + tail - take, take * 4, hipMemcpyHostToDevice, st));
        QuerySpec qc = q;
        qc.schema.cols[0] = dv.p;
        tga_options o{};
        o.path = TGA_PATH_AUTO;
        Plan plan = autotune(ctx, qc, take, std::min<uint64_t>(take, max_load), &o);
        plan.gslots = S;
        std::vector<DBuf> keep;
        tga_status s2 = aggregate_into(ctx, qc, dk.p, take, plan, g, keep);
        if (s2) return s2;
        uint32_t hf[kFlagWords];
        TGA_HIP(ctx, launch_fill_u64(gcount.as<uint64_t>(), 1, 0, st));
        TGA_HIP(ctx, launch_count_groups(4, g, gcount.as<uint64_t>(), st));
        TGA_HIP(ctx, hipMemcpyAsync(hf, g.flags, sizeof hf, hipMemcpyDeviceToHost, st));
        TGA_HIP(ctx, hipMemcpyAsync(&already, gcount.p, 8, hipMemcpyDeviceToHost, st));
        TGA_HIP(ctx, hipStreamSynchronize(st));
        if (hf[FLAG_OVERFLOW]) return fail(ctx, TGA_ERR_TABLE_OVERFLOW, "hash table overflow");
        tail -= take;
        rest -= take;
      }
      OutDesc od{};
      od.keys = rk.p;
      od.cols[0] = rs.p;
      od.capacity = S + 1;
      uint32_t hf[kFlagWords];
      uint64_t G = 0;
      tga_status s2 = compact_into(ctx, q, g, od, hf, &G);
      if (s2) return s2;
      hs.resize(G);
      TGA_HIP(ctx, hipMemcpyAsync(ok + begin[i], rk.p, G * 4, hipMemcpyDeviceToHost, st));
      TGA_HIP(ctx, hipMemcpyAsync(hs.data(), rs.p, G * 8, hipMemcpyDeviceToHost, st));
      TGA_HIP(ctx, hipStreamSynchronize(st));
      for (uint64_t j = 0; j < G; ++j) ov[begin[i] + j] = (uint32_t)hs[j];
      par_kv_begin[i] = begin[i];
      par_result_kv_num[i] = G;
      return TGA_OK;
    };
    tga_status rc = run();
    if (rc) return fail_free(rc);
  }
  // the merged list replaces the caller's list and is drained (intra:294, pop2device)
  ps->parts.assign(np, tga_partset::Part());
  *out_keys = ok;
  *out_vals = ov;
  *n_merged = np;
  return TGA_OK;
}
