// Runs the batched persistent path on the CPU execution model with the query
// mix of tests/test_gpu_batch_serve.py and checks every query against the
// oracle (tga_o_groupby), over several replays of the batch graph.
// Usage: batch_driver [mode] [replays]
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/tga_groupby.h"

extern "C" void emu_watch_batch(const tga_batch *b);
extern "C" int64_t tga_o_groupby(const void *keys, int key_bytes, int64_t n, const void *const *cols,
                                 const int *col_dtypes, int ncols, const int *agg_ops, const int *agg_cols, int naggs,
                                 uint64_t *out_keys, void *const *out_aggs, int64_t out_cap);

#define CHECK(x)                                                                      \
  do {                                                                                \
    tga_status st_ = (x);                                                             \
    if (st_) {                                                                        \
      std::fprintf(stderr, "%s:%d: %s -> %d (%s)\n", __FILE__, __LINE__, #x, (int)st_, \
                   tga_last_error(ctx));                                              \
      return 1;                                                                       \
    }                                                                                 \
  } while (0)

// est > 0: the hint passed instead of d (far below the truth: the query's
// table overflows and it is rerun alone, at every replay)
struct Spec { int64_t n; uint64_t d; int bits; bool hint; int64_t est = 0; };

int main(int argc, char **argv) {
  const int mode = argc > 1 ? std::atoi(argv[1]) : 0;
  const int replays = argc > 2 ? std::atoi(argv[2]) : 2;
  const std::vector<Spec> specs = {
      {1 << 10, 16, 32, true},       {5000, 5000, 32, true},     {1 << 16, 1 << 12, 32, true},
      {1 << 18, 1 << 16, 32, true},  {1 << 20, 1 << 18, 32, true}, {300000, 100000, 64, true},
      {1 << 17, 30, 64, false},      {0, 1, 32, true},           {1, 1, 32, true},
      {200000, 50, 32, false},       {100000, 90000, 32, false}, {70000, 64, 32, true},
      {1 << 18, 100000, 32, true, 10}, {20000, 6000, 64, true, 10},
      {1 << 20, 200, 32, true},  // direct over several row-range items: flushes, compaction
  };
  tga_context *ctx = nullptr;
  if (tga_context_create(0, 0, &ctx)) { std::fprintf(stderr, "context\n"); return 1; }
  const size_t nq = specs.size();
  std::vector<void *> dk(nq), dv(nq);
  std::vector<std::vector<uint64_t>> hk(nq);
  std::vector<std::vector<int32_t>> hv(nq);
  std::vector<tga_column> vcol(nq);
  std::vector<tga_query> qs(nq);
  const tga_agg aggs[4] = {{TGA_COUNT, 0}, {TGA_SUM, 0}, {TGA_MIN, 0}, {TGA_MAX, 0}};
  for (size_t i = 0; i < nq; ++i) {
    const Spec &s = specs[i];
    const int kb = s.bits / 8;
    CHECK(tga_device_alloc(ctx, std::max<int64_t>(1, s.n) * kb, &dk[i]));
    CHECK(tga_device_alloc(ctx, std::max<int64_t>(1, s.n) * 4, &dv[i]));
    if (s.n) {
      CHECK(tga_generate(ctx, s.bits == 32 ? TGA_GEN_KEYS_U32 : TGA_GEN_KEYS_U64, dk[i], s.n, s.d, 0, 0, 0, 100 + i, i));
      CHECK(tga_generate(ctx, TGA_GEN_VALS_I32, dv[i], s.n, (uint64_t)(-(int64_t)(1 << 30)), 1ULL << 31, 0, 0, 200 + i, 0));
    }
    hk[i].resize(s.n);
    hv[i].resize(s.n);
    std::vector<uint32_t> k32(s.bits == 32 ? s.n : 0);
    if (s.n) {
      if (s.bits == 32) {
        CHECK(tga_memcpy_d2h(ctx, k32.data(), dk[i], s.n * 4));
        for (int64_t r = 0; r < s.n; ++r) hk[i][r] = k32[r];
      } else {
        CHECK(tga_memcpy_d2h(ctx, hk[i].data(), dk[i], s.n * 8));
      }
      CHECK(tga_memcpy_d2h(ctx, hv[i].data(), dv[i], s.n * 4));
    }
    vcol[i] = tga_column{dv[i], TGA_I32, TGA_DEVICE};
    qs[i].key = tga_column{dk[i], s.bits == 32 ? TGA_U32 : TGA_U64, TGA_DEVICE};
    qs[i].values = &vcol[i];
    qs[i].n_values = 1;
    qs[i].aggs = aggs;
    qs[i].n_aggs = 4;
    qs[i].n_rows = s.n;
    qs[i].est_groups = s.est > 0 ? s.est : s.hint ? (int64_t)s.d : 0;
  }
  tga_batch *b = nullptr;
  CHECK(tga_batch_create(ctx, qs.data(), (int32_t)nq, mode, &b));
  emu_watch_batch(b);
  int bad = 0;
  // A second batch built between replays and held across the next one takes
  // its small output buffers from the arena's free lists, where an overflowed
  // query's released buffers would sit: the first batch's kernels must never
  // write into memory it gave back (a rerun must not release the buffers its
  // device record still targets).
  const size_t hq = 11;
  int64_t hsum = 0;
  for (int32_t v : hv[hq]) hsum += v;
  for (int rep = 0; rep < replays; ++rep) {
    std::printf("replay %d\n", rep);
    std::fflush(stdout);
    tga_batch *held = nullptr;
    if (rep > 0) {
      CHECK(tga_batch_create(ctx, &qs[hq], 1, 0, &held));
      CHECK(tga_batch_run(held));
    }
    CHECK(tga_batch_run(b));
    if (held) {
      int64_t g = 0;
      CHECK(tga_batch_num_groups(held, 0, &g));
      std::vector<int64_t> c(std::max<int64_t>(g, 1)), sm(std::max<int64_t>(g, 1));
      CHECK(tga_batch_copy_to_host(held, 0, 0, c.data(), g));
      CHECK(tga_batch_copy_to_host(held, 0, 1, sm.data(), g));
      int64_t tc = 0, ts = 0;
      for (int64_t r = 0; r < g; ++r) { tc += c[r]; ts += sm[r]; }
      if (g != (int64_t)specs[hq].d || tc != specs[hq].n || ts != hsum) {
        std::printf("held batch result changed by the replay: %lld groups, count %lld, sum %lld\n", (long long)g,
                    (long long)tc, (long long)ts);
        ++bad;
      }
      tga_batch_destroy(held);
    }
    for (size_t i = 0; i < nq; ++i) {
      const int64_t n = specs[i].n;
      // oracle on 64-bit keys (the 32-bit key values widened keep their order and identity)
      std::vector<uint64_t> ok(std::max<int64_t>(n, 1));
      std::vector<int64_t> oc(ok.size()), os(ok.size());
      std::vector<int32_t> omin(ok.size()), omax(ok.size());
      const void *cols[1] = {hv[i].data()};
      const int cdt[1] = {TGA_I32};
      const int ops[4] = {TGA_COUNT, TGA_SUM, TGA_MIN, TGA_MAX}, acs[4] = {0, 0, 0, 0};
      void *outs[4] = {oc.data(), os.data(), omin.data(), omax.data()};
      const int64_t G = tga_o_groupby(hk[i].data(), 8, n, cols, cdt, 1, ops, acs, 4, ok.data(), outs, (int64_t)ok.size());
      int64_t g = -1;
      CHECK(tga_batch_num_groups(b, (int32_t)i, &g));
      if (g != G) { std::printf("query %zu: %lld groups, oracle %lld\n", i, (long long)g, (long long)G); ++bad; continue; }
      const int kb = specs[i].bits / 8;
      std::vector<unsigned char> keys(std::max<int64_t>(1, g) * kb);
      std::vector<int64_t> c(std::max<int64_t>(1, g)), sm(std::max<int64_t>(1, g));
      std::vector<int32_t> mn(std::max<int64_t>(1, g)), mx(std::max<int64_t>(1, g));
      if (g) {
        CHECK(tga_batch_copy_to_host(b, (int32_t)i, -1, keys.data(), g));
        CHECK(tga_batch_copy_to_host(b, (int32_t)i, 0, c.data(), g));
        CHECK(tga_batch_copy_to_host(b, (int32_t)i, 1, sm.data(), g));
        CHECK(tga_batch_copy_to_host(b, (int32_t)i, 2, mn.data(), g));
        CHECK(tga_batch_copy_to_host(b, (int32_t)i, 3, mx.data(), g));
      }
      std::vector<int64_t> order(g);
      for (int64_t r = 0; r < g; ++r) order[r] = r;
      auto key_at = [&](int64_t r) -> uint64_t {
        if (kb == 4) { uint32_t v; std::memcpy(&v, &keys[r * 4], 4); return v; }
        uint64_t v; std::memcpy(&v, &keys[r * 8], 8); return v;
      };
      std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return key_at(x) < key_at(y); });
      int64_t mism = 0;
      for (int64_t r = 0; r < g; ++r) {
        const int64_t j = order[r];
        if (key_at(j) != ok[r] || c[j] != oc[r] || sm[j] != os[r] || mn[j] != omin[r] || mx[j] != omax[r]) ++mism;
      }
      if (mism) { std::printf("query %zu: %lld of %lld rows differ\n", i, (long long)mism, (long long)g); ++bad; }
    }
  }
  tga_batch_destroy(b);
  std::printf(bad ? "FAIL %d\n" : "OK\n", bad);
  return bad ? 1 : 0;
}
