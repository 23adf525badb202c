// tga_noalloc.hip -- tga_groupby_into_noalloc: the plain operator into
// caller-owned device buffers with no device allocation at all, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// Every temporary of the call -- the global hash table, its status words, the
// work plan, the key sample and the compaction descriptors -- lives in one
// file-scope device array, `scratch`, of kNoAllocScratchBytes, which each GPU
// the library runs on holds until the process ends.  Its device address is
// looked up once per device (noalloc_scratch_init, called when a context is
// created on that device); host code only ever passes that address, never the
// array's name.  The call runs the direct path only (workgroup LDS tables
// flushed into the global table), so it needs no per-row storage; the number
// of groups is bounded by the table the array holds: half its slots, about
// 2^18 groups for four 8-byte aggregates of 64-bit keys.
//
// Every call on one GPU shares that GPU's array and overwrites it; the
// kernels clear what must start at zero (table, status words, plan ticket,
// compaction descriptors) on each call.  Calls on one context serialise on
// its lock; calls that overlap on one GPU from different contexts race.
#include <algorithm>
#include <mutex>

#include "tga_host.hpp"

namespace tga {

constexpr uint64_t kNoAllocScratchBytes = 32ULL << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices] = {};
hipError_t g_scratch_err[kNoAllocMaxDevices] = {};
bool g_scratch_done[kNoAllocMaxDevices] = {};

// Fixed part of the scratch layout (byte offsets); the table follows.
constexpr uint64_t kOffFlags = 0;                    // kFlagWords u32
constexpr uint64_t kOffPlan = 256;                   // pstart[2] u64 | cstart[2] u32 | ticket u32 | split u64
constexpr uint64_t kOffSample = 512;                 // 8 u64 of counts + 16384 sampled keys
constexpr uint64_t kSampleKeys = 16384;
constexpr uint64_t kOffDesc = kOffSample + (kSampleKeys + 8) * 8;  // compaction descriptors
constexpr uint64_t kDescBytes = 64 << 10;
constexpr uint64_t kOffTable = (kOffDesc + kDescBytes + 255) & ~255ULL;

// Largest power-of-two slot count S whose table (S + 1 slots, every array
// 256-byte aligned) fits behind the fixed part, and whose compaction
// descriptors fit theirs.
uint64_t max_slots(const QuerySpec &q) {
  uint64_t S = 1ULL << 30;
  while (S >= 1024) {
    uint64_t bytes = ((S + 1) * (uint64_t)q.key_bytes + 255) & ~255ULL;
    for (int a = 0; a < q.schema.nacc; ++a)
      bytes += ((S + 1) * (uint64_t)acc_global_bytes(q.schema.acc[a].kind) + 255) & ~255ULL;
    if (kOffTable + bytes <= kNoAllocScratchBytes && (compact_tiles(S + 1) + 2) * 8 + 64 <= kDescBytes) return S;
    S >>= 1;
  }
  return 0;
}
}  // namespace

// Looks up this device's address of `scratch` (the device must be current).
hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_done[device]) return g_scratch_err[device];
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_done[device] = true;
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
  if (e != hipSuccess) (void)hipGetLastError();
  return e;
}

static void *scratch_of(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return g_scratch_ptr[device];
}

}  // namespace tga

using namespace tga;

#define TGA_API extern "C" __attribute__((visibility("default")))

TGA_API tga_status tga_groupby_into_noalloc(tga_context *ctx, const tga_column *key, const tga_column *values,
                                            int32_t n_values, int64_t n_rows, const tga_agg *aggs, int32_t n_aggs,
                                            const tga_options *opt, void *out_keys, void *const *out_aggs,
                                            int64_t capacity, int64_t *n_groups) {
  if (!ctx || !out_keys || !out_aggs || capacity < 0 || !n_groups || n_rows < 0) return TGA_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  QuerySpec q;
  {
    const std::string e = build_spec(key, values, n_values, aggs, n_aggs, q);
    if (!e.empty()) return fail(ctx, TGA_ERR_INVALID_ARG, "%s", e.c_str());
  }
  if (key->location != TGA_DEVICE) return fail(ctx, TGA_ERR_INVALID_ARG, "no-alloc group-by: key column must be on the device");
  for (int c = 0; c < n_values; ++c)
    if (values[c].location != TGA_DEVICE)
      return fail(ctx, TGA_ERR_INVALID_ARG, "no-alloc group-by: value column %d must be on the device", c);
  if (n_rows > 0 && !key->data) return fail(ctx, TGA_ERR_INVALID_ARG, "key data is null");
  for (int c = 0; c < n_values; ++c)
    if (n_rows > 0 && !values[c].data) return fail(ctx, TGA_ERR_INVALID_ARG, "value column %d is null", c);
  for (int a = 0; a < n_aggs; ++a)
    if (!out_aggs[a]) return fail(ctx, TGA_ERR_INVALID_ARG, "output column %d is null", a);
  char *base = (char *)scratch_of(ctx->device);
  if (!base) return fail(ctx, TGA_ERR_HIP, "no-alloc scratch of device %d is not available", ctx->device);
  const uint64_t Smax = max_slots(q);
  const uint64_t max_groups = Smax / 2;
  if (opt && opt->est_groups > 0 && (uint64_t)opt->est_groups > max_groups)
    return fail(ctx, TGA_ERR_INVALID_ARG,
                "no-alloc group-by: est_groups %lld exceeds the %llu groups its scratch table holds",
                (long long)opt->est_groups, (unsigned long long)max_groups);
  if (opt && opt->path == TGA_PATH_PARTITIONED)
    return fail(ctx, TGA_ERR_UNSUPPORTED, "no-alloc group-by runs the direct path only");
  hipStream_t st = ctx->stream;
  *n_groups = 0;
  const uint64_t n = (uint64_t)n_rows;
  uint32_t *flags = (uint32_t *)(base + kOffFlags);
  uint64_t *sample = (uint64_t *)(base + kOffSample);
  uint64_t *desc = (uint64_t *)(base + kOffDesc);
  char *planb = base + kOffPlan;
  // cardinality: the caller's hint, else the key sample's statistics
  uint64_t est = opt && opt->est_groups > 0 ? (uint64_t)opt->est_groups : 0;
  if (!est && n > 0) {
    const uint64_t m = std::min<uint64_t>(n, kSampleKeys);
    TGA_HIP(ctx, launch_sample_stats(q.key_bytes, key->data, n, (uint32_t)m, sample + 8, sample, st));
    TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host, sample, 4 * 8, hipMemcpyDeviceToHost, st));
    TGA_HIP(ctx, hipStreamSynchronize(st));
    uint64_t h[4];
    std::memcpy(h, ctx->status_host, sizeof h);
    // counts incomplete (keys crowding the counting table): plan for the
    // largest table
    est = h[3] ? max_groups : estimate_from_stats(h[0], h[1], h[2], m, n);
  }
  tga_options o{};
  if (opt) o = *opt;
  o.path = TGA_PATH_DIRECT;
  Plan plan = autotune(ctx, q, n, std::max<uint64_t>(est, 1), &o);
  uint64_t S = std::min<uint64_t>(std::max<uint64_t>(plan.gslots, 1024), Smax);
  for (;;) {
    // the table: keys, then one array per aggregate, 256-byte aligned
    GTable g{};
    char *t = base + kOffTable;
    g.keys = t;
    t += ((S + 1) * (uint64_t)q.key_bytes + 255) & ~255ULL;
    for (int a = 0; a < q.schema.nacc; ++a) {
      g.acc[a] = t;
      t += ((S + 1) * (uint64_t)acc_global_bytes(q.schema.acc[a].kind) + 255) & ~255ULL;
    }
    g.S = S;
    g.flags = flags;
    TGA_HIP(ctx, launch_gtable_init(q.key_bytes, g, q.schema, st));  // also clears the status words
    if (n > 0) {
      AggPlan ap{(uint64_t *)planb, (uint32_t *)(planb + 16), 1, (uint32_t *)(planb + 32), (uint64_t *)(planb + 40)};
      const uint64_t nch = (n + plan.chunk_rows - 1) / plan.chunk_rows;
      TGA_HIP(ctx, launch_plan_direct(ap, n, (uint32_t)nch, st));
      const unsigned grid = (unsigned)std::min<uint64_t>(plan.workgroups, nch);
      TGA_HIP(ctx, launch_agg(q.key_bytes, key->data, q.schema, ap, plan.layout, g, plan.nbits, plan.max_probe, grid,
                              plan.block, st, 0, nullptr, nullptr, plan.fold));
    }
    const uint64_t ctiles = compact_tiles(S + 1);
    TGA_HIP(ctx, launch_fill_u64(desc, ctiles + 2, 0, st));
    OutDesc od{};
    od.keys = out_keys;
    for (int a = 0; a < q.schema.nacc; ++a) od.cols[a] = out_aggs[a];
    od.capacity = (uint64_t)capacity;
    uint64_t *count = desc + ctiles + 1;
    TGA_HIP(ctx, launch_compact(q.key_bytes, g, q.schema, od, desc, (uint32_t *)(desc + ctiles), count, st));
    TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host, flags, kFlagWords * 4, hipMemcpyDeviceToHost, st));
    TGA_HIP(ctx, hipMemcpyAsync(ctx->status_host + 4, count, 8, hipMemcpyDeviceToHost, st));
    TGA_HIP(ctx, hipStreamSynchronize(st));
    uint32_t hf[kFlagWords];
    std::memcpy(hf, ctx->status_host, sizeof hf);
    uint64_t G = 0;
    std::memcpy(&G, ctx->status_host + 4, 8);
    *n_groups = (int64_t)G;
    // (after a table overflow G is a lower bound of the group count)
    if (G > (uint64_t)capacity)
      return fail(ctx, TGA_ERR_CAPACITY, "%llu groups > capacity %lld", (unsigned long long)G, (long long)capacity);
    if (hf[FLAG_OVERFLOW]) {
      // more groups than half the table: a larger one, up to the scratch
      if (S >= Smax)
        return fail(ctx, TGA_ERR_OUT_OF_MEMORY,
                    "no-alloc group-by: more groups than its scratch table holds (about %llu)",
                    (unsigned long long)max_groups);
      S = std::min<uint64_t>(S * 4, Smax);
      continue;
    }
    if (hf[FLAG_RANGE]) return fail(ctx, TGA_ERR_RANGE, "fp64 SUM input outside [2^-96, 2^64) or not finite");
    ctx->last_plan = tga_options{};
    ctx->last_plan.path = TGA_PATH_DIRECT;
    ctx->last_plan.table_slots = (int32_t)plan.table_slots;
    ctx->last_plan.workgroups = (int32_t)plan.workgroups;
    ctx->last_plan.block_threads = (int32_t)plan.block;
    ctx->last_plan.chunk_rows = (int64_t)plan.chunk_rows;
    ctx->last_plan.est_groups = (int64_t)(S / 2);
    ctx->last_plan.attempts = 1;
    return TGA_OK;
  }
}
