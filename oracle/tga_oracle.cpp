// tga_oracle.cpp -- CPU oracle for the tiled hash group-by.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
//
// It holds three things, each citing the reference (paths relative to the
// reference repository root; `T/` = tiled_groupby_agg_partition/):
//
//  1. Counter-based synthetic-data generators whose integer streams are
//     bit-identical with the device generators of the product library.
//     (The reference generator is test/include/data_generator.hpp:8-105; its
//     std::rand/minstd streams cannot be reproduced on a GPU, so the workload
//     uses SplitMix64 keyed by (seed, row) instead.)
//
//  2. tga_o_groupby: an exact hash group-by with COUNT / SUM / MIN / MAX, the
//     order-free definition of what the reference computes
//     (T/src/kernel.cu:48,152,334,340 -- atomicAdd of the value into the
//     slot of its key), widened to int64 SUM, COUNT, MIN, MAX and a
//     compensated (Neumaier) fp64 SUM.  Output is sorted by key.
//
//  3. A restatement of the reference's two-phase algorithm on host threads:
//       phase 1  T/src/group_then_partition.cu:24-425  (tile pre-aggregation
//                with the one-third single-probe table K1 and the
//                linear-probing spill table K4, compaction K2/K3/K5, hash
//                partitioning K6/K7, write-back of the tile result into the
//                caller's buffers and the per-partition segment lists of
//                T/include/par_result.cuh:11-82 / T/src/par_result.cu:3-13)
//       merge    T/src/groupby_agg_intra_partition.cu:13-48
//       phase 2  T/src/groupby_agg_intra_partition.cu:89-279  (gather,
//                chunked linear-probing aggregation with the overflow rule of
//                :213-253, scan + compaction K10)
//     It is templated on the key type and carries a full partial state
//     (count, int64 sum, min, max) so that it serves both as the reference
//     path for the u32/u32 case and as the CPU baseline for int64 keys.
//     Deviation: empty slots are tracked with an occupancy byte, so the key
//     0xffffffff (the reference's sentinel, T/include/util.cuh:8) is
//     aggregated instead of being dropped.
//
// Build: oracle/Makefile -> oracle/liboracle.so (extern "C" API below).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <queue>
#include <thread>
#include <vector>

#define OAPI extern "C"

// ---------------------------------------------------------------------------
// 1. generators (must match csrc/tga_gen.hip bit for bit)
// ---------------------------------------------------------------------------

static inline uint64_t o_mix64(uint64_t z) {
  // SplitMix64 finalizer (Steele, Lea & Flood 2014); a bijection on 64 bits.
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static inline uint32_t o_fmix32(uint32_t h) {
  // MurmurHash3 32-bit finalizer; a bijection on 32 bits.
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}

OAPI uint64_t tga_o_splitmix(uint64_t seed, uint64_t i) {
  return o_mix64(seed + (i + 1) * 0x9e3779b97f4a7c15ULL);
}

// Lemire's multiply-shift reduction of a 32-bit draw onto [0, d).
static inline uint64_t o_fastrange(uint32_t u, uint64_t d) {
  return (uint64_t)(((unsigned __int128)u * d) >> 32);
}

// Key of distinct-index j: a bijection of j, so D indices give D distinct keys.
OAPI uint32_t tga_o_key_u32(uint64_t j, uint64_t salt) { return o_fmix32((uint32_t)j ^ (uint32_t)salt); }
OAPI uint64_t tga_o_key_u64(uint64_t j, uint64_t salt) { return o_mix64(j ^ salt); }

// row i draws index fastrange(low32(splitmix(seed, i)), D)
OAPI void tga_o_gen_keys_u32(uint32_t *out, int64_t n, uint64_t distinct, uint64_t seed, uint64_t salt) {
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u = (uint32_t)tga_o_splitmix(seed, (uint64_t)i);
    out[i] = tga_o_key_u32(o_fastrange(u, distinct), salt);
  }
}
OAPI void tga_o_gen_keys_u64(uint64_t *out, int64_t n, uint64_t distinct, uint64_t seed, uint64_t salt) {
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u = (uint32_t)tga_o_splitmix(seed, (uint64_t)i);
    out[i] = tga_o_key_u64(o_fastrange(u, distinct), salt);
  }
}
// value uniform in [lo, hi): lo + fastrange(high32(splitmix(seed, i)), hi - lo)
OAPI void tga_o_gen_vals_i32(int32_t *out, int64_t n, int64_t lo, int64_t hi, uint64_t seed) {
  uint64_t span = (uint64_t)(hi - lo);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t u = (uint32_t)(tga_o_splitmix(seed, (uint64_t)i) >> 32);
    out[i] = (int32_t)(lo + (int64_t)o_fastrange(u, span));
  }
}

// ---------------------------------------------------------------------------
// 2. exact group-by
// ---------------------------------------------------------------------------

enum { O_U32 = 1, O_I32 = 2, O_U64 = 3, O_I64 = 4, O_F64 = 5 };
enum { O_COUNT = 0, O_SUM = 1, O_MIN = 2, O_MAX = 3 };

static inline int64_t o_read_int(const void *col, int dt, int64_t i) {
  switch (dt) {
    case O_U32: return (int64_t)((const uint32_t *)col)[i];
    case O_I32: return (int64_t)((const int32_t *)col)[i];
    case O_U64: return (int64_t)((const uint64_t *)col)[i];
    default: return ((const int64_t *)col)[i];
  }
}

struct OAcc {
  int64_t cnt;
  uint64_t isum;   // wraps mod 2^64 like the device int64 atomics
  int64_t imin, imax;
  uint64_t umin, umax;
  double fs, fc;   // Neumaier running sum and compensation
  double fmin, fmax;
};

static inline void o_neumaier(double &s, double &c, double x) {
  double t = s + x;
  if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x; else c += (x - t) + s;
  s = t;
}

// Returns the group count (or -1 if out_cap is too small).  Outputs are sorted
// by the key bits read as unsigned integers.  out_aggs[a] has the dtype:
// COUNT -> int64; SUM -> int64 (integer column) or double (f64 column);
// MIN/MAX -> the column's own dtype.
OAPI int64_t tga_o_groupby(const void *keys, int key_bytes, int64_t n,
                      const void *const *cols, const int *col_dtypes, int ncols,
                      const int *agg_ops, const int *agg_cols, int naggs,
                      uint64_t *out_keys, void *const *out_aggs, int64_t out_cap) {
  (void)ncols;
  uint64_t cap = 16;
  while (cap < (uint64_t)n * 2 + 16) cap <<= 1;
  std::vector<uint64_t> tk(cap);
  std::vector<int64_t> slot_group(cap, -1);
  std::vector<uint64_t> gkeys;
  // accumulator per (group, agg)
  std::vector<OAcc> acc;
  auto new_acc = [&]() {
    for (int a = 0; a < naggs; ++a) {
      OAcc z;
      z.cnt = 0; z.isum = 0; z.imin = INT64_MAX; z.imax = INT64_MIN;
      z.umin = UINT64_MAX; z.umax = 0; z.fs = 0; z.fc = 0;
      z.fmin = INFINITY; z.fmax = -INFINITY;
      acc.push_back(z);
    }
  };
  for (int64_t i = 0; i < n; ++i) {
    uint64_t k = key_bytes == 4 ? (uint64_t)((const uint32_t *)keys)[i] : ((const uint64_t *)keys)[i];
    uint64_t s = o_mix64(k) & (cap - 1);
    int64_t g;
    for (;;) {
      g = slot_group[s];
      if (g < 0) {
        g = (int64_t)gkeys.size();
        slot_group[s] = g; tk[s] = k; gkeys.push_back(k); new_acc();
        break;
      }
      if (tk[s] == k) break;
      s = (s + 1) & (cap - 1);
    }
    for (int a = 0; a < naggs; ++a) {
      OAcc &z = acc[(size_t)g * naggs + a];
      int op = agg_ops[a];
      if (op == O_COUNT) { z.cnt++; continue; }
      int c = agg_cols[a];
      int dt = col_dtypes[c];
      if (dt == O_F64) {
        double x = ((const double *)cols[c])[i];
        if (op == O_SUM) o_neumaier(z.fs, z.fc, x);
        else if (op == O_MIN) z.fmin = x < z.fmin ? x : z.fmin;
        else z.fmax = x > z.fmax ? x : z.fmax;
      } else if (dt == O_U64) {
        uint64_t x = ((const uint64_t *)cols[c])[i];
        if (op == O_SUM) z.isum += x;
        else if (op == O_MIN) z.umin = x < z.umin ? x : z.umin;
        else z.umax = x > z.umax ? x : z.umax;
      } else {
        int64_t x = o_read_int(cols[c], dt, i);
        if (op == O_SUM) z.isum += (uint64_t)x;
        else if (op == O_MIN) z.imin = x < z.imin ? x : z.imin;
        else z.imax = x > z.imax ? x : z.imax;
      }
    }
  }
  int64_t G = (int64_t)gkeys.size();
  if (G > out_cap) return -1;
  std::vector<int64_t> order((size_t)G);
  for (int64_t g = 0; g < G; ++g) order[(size_t)g] = g;
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return gkeys[(size_t)a] < gkeys[(size_t)b]; });
  for (int64_t r = 0; r < G; ++r) {
    int64_t g = order[(size_t)r];
    out_keys[r] = gkeys[(size_t)g];
    for (int a = 0; a < naggs; ++a) {
      const OAcc &z = acc[(size_t)g * naggs + a];
      int op = agg_ops[a];
      if (op == O_COUNT) { ((int64_t *)out_aggs[a])[r] = z.cnt; continue; }
      int dt = col_dtypes[agg_cols[a]];
      if (op == O_SUM) {
        if (dt == O_F64) ((double *)out_aggs[a])[r] = z.fs + z.fc;
        else ((int64_t *)out_aggs[a])[r] = (int64_t)z.isum;
        continue;
      }
      bool mn = op == O_MIN;
      switch (dt) {
        case O_F64: ((double *)out_aggs[a])[r] = mn ? z.fmin : z.fmax; break;
        case O_U64: ((uint64_t *)out_aggs[a])[r] = mn ? z.umin : z.umax; break;
        case O_I64: ((int64_t *)out_aggs[a])[r] = mn ? z.imin : z.imax; break;
        case O_U32: ((uint32_t *)out_aggs[a])[r] = (uint32_t)(mn ? z.imin : z.imax); break;
        default: ((int32_t *)out_aggs[a])[r] = (int32_t)(mn ? z.imin : z.imax); break;
      }
    }
  }
  return G;
}

// Neumaier compensated sum of a plain array (the fp64 oracle of one group).
OAPI double tga_o_neumaier(const double *x, int64_t n) {
  double s = 0, c = 0;
  for (int64_t i = 0; i < n; ++i) o_neumaier(s, c, x[i]);
  return s + c;
}

// Inclusive / exclusive prefix sums (the textbook definition the reference's
// CUB calls at T/src/util.cu:59,82 compute).
OAPI void tga_o_scan_u64(const uint64_t *in, uint64_t *out, int64_t n, int exclusive) {
  uint64_t run = 0;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t v = in[i];
    if (exclusive) { out[i] = run; run += v; } else { run += v; out[i] = run; }
  }
}

// ---------------------------------------------------------------------------
// 3. restatement of the reference's two-phase algorithm
// ---------------------------------------------------------------------------

// MurmurHash3_x86_32 (Appleby, public domain) over the key's bytes: the hash
// family the reference uses at T/src/kernel.cu:22,127,196,261,309.
static inline uint32_t o_rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
static uint32_t o_murmur3_32(const void *data, int len, uint32_t seed) {
  const uint8_t *p = (const uint8_t *)data;
  uint32_t h = seed;
  const uint32_t c1 = 0xcc9e2d51u, c2 = 0x1b873593u;
  int nblocks = len / 4;
  for (int b = 0; b < nblocks; ++b) {
    uint32_t k;
    std::memcpy(&k, p + 4 * b, 4);
    k *= c1; k = o_rotl32(k, 15); k *= c2;
    h ^= k; h = o_rotl32(h, 13); h = h * 5 + 0xe6546b64u;
  }
  // keys are 4 or 8 bytes: no tail
  h ^= (uint32_t)len;
  return o_fmix32(h);
}
OAPI uint32_t tga_o_murmur3_u32(uint32_t key, uint32_t seed) { return o_murmur3_32(&key, 4, seed); }

struct OState {  // partial aggregate of one group
  uint64_t cnt;
  uint64_t sum;  // int64 sum, wraps mod 2^64; its low 32 bits are the reference's u32 sum
  int64_t mn, mx;
};

template <typename K>
struct ORun {  // rows of the partitioned tile results, kept SoA like the reference
  std::vector<K> key;
  std::vector<OState> st;
};

struct OSeg { size_t begin; uint32_t len; };  // a par_result segment: [begin, begin+len)

struct OPar {  // T/include/par_result.cuh:11-82
  std::vector<OSeg> seg;
  uint32_t size = 0;  // u32, as the reference (par_result.cuh:16)
  void push(size_t b, uint32_t n) { seg.push_back({b, n}); size += n; }
  void merge(const OPar &o) { seg.insert(seg.end(), o.seg.begin(), o.seg.end()); size += o.size; }
};

template <typename K>
struct OTable {  // open-addressing table with an occupancy byte instead of a sentinel key
  std::vector<K> key;
  std::vector<OState> st;
  std::vector<uint8_t> used;
  explicit OTable(size_t n) : key(n), st(n), used(n, 0) {}
  void reset(size_t n) {
    std::fill(used.begin(), used.begin() + (ptrdiff_t)n, 0);
  }
};

static inline void o_state_add(OState &a, const OState &b) {
  a.cnt += b.cnt; a.sum += b.sum;
  a.mn = b.mn < a.mn ? b.mn : a.mn; a.mx = b.mx > a.mx ? b.mx : a.mx;
}

template <typename K>
struct OPhase1 {
  const K *keys;
  const OState *rows;   // input rows as unit states (cnt=1, sum=min=max=value)
  size_t n, tile_len, P;
  ORun<K> out;          // partitioned tile results, tile t at [t*tile_len, ...)
  std::vector<uint32_t> par_pos;  // (P+1) per tile: exclusive offsets + result count
  std::vector<size_t> tile_result;  // r per tile
  std::atomic<size_t> counter{0};
};

// compute_groupby_agg_result (group_then_partition.cu:24-246) followed by
// compute_partition_result (:248-306) for one tile, sequentially.
template <typename K>
static void o_tile(OPhase1<K> &ph, size_t tile_id, OTable<K> &tab, std::vector<uint8_t> &flag,
                   std::vector<K> &rk, std::vector<OState> &rs) {
  size_t b = tile_id * ph.tile_len;
  size_t e = std::min(ph.n, b + ph.tile_len);
  size_t n = e - b;
  size_t third = n / 3;
  uint32_t seed1 = o_fmix32((uint32_t)(2 * tile_id + 1));   // reference: std::rand() (:54)
  uint32_t seed2 = o_fmix32((uint32_t)(2 * tile_id + 2));   // reference: std::rand() (:179)
  tab.reset(n);
  std::fill(flag.begin(), flag.begin() + (ptrdiff_t)n, 0);
  // K1 hash_into_one_third_ht (kernel.cu:9-54): one probe into the first third
  for (size_t i = 0; i < n; ++i) {
    K k = ph.keys[b + i];
    size_t loc = o_murmur3_32(&k, sizeof(K), seed1) % n;
    if (loc >= third) { flag[i] = 1; continue; }
    if (!tab.used[loc]) { tab.used[loc] = 1; tab.key[loc] = k; tab.st[loc] = ph.rows[b + i]; }
    else if (tab.key[loc] == k) o_state_add(tab.st[loc], ph.rows[b + i]);
    else flag[i] = 1;
  }
  // K2 collect_one_third_ht_result (kernel.cu:57-83): compact table 1 in slot order
  rk.clear(); rs.clear();
  std::vector<K> k1; std::vector<OState> s1;
  for (size_t s = 0; s < third; ++s)
    if (tab.used[s]) { k1.push_back(tab.key[s]); s1.push_back(tab.st[s]); }
  // K3 + K4 kv_hash_agg_into_ht (kernel.cu:85-158): flagged rows, linear probing over n slots
  tab.reset(n);
  for (size_t i = 0; i < n; ++i) {
    if (!flag[i]) continue;
    K k = ph.keys[b + i];
    size_t loc = o_murmur3_32(&k, sizeof(K), seed2) % n;
    for (;;) {
      if (!tab.used[loc]) { tab.used[loc] = 1; tab.key[loc] = k; tab.st[loc] = ph.rows[b + i]; break; }
      if (tab.key[loc] == k) { o_state_add(tab.st[loc], ph.rows[b + i]); break; }
      loc = (loc + 1) % n;
    }
  }
  // K5 combine_ht_result_with_previous_result (kernel.cu:160-186): table-2
  // results first, then table-1 results (result = [n-c2-c5, n), gtp:222-245)
  for (size_t s = 0; s < n; ++s)
    if (tab.used[s]) { rk.push_back(tab.key[s]); rs.push_back(tab.st[s]); }
  rk.insert(rk.end(), k1.begin(), k1.end());
  rs.insert(rs.end(), s1.begin(), s1.end());
  size_t r = rk.size();
  // K6 cal_kv_intra_group_ind + exclusive scan + K7 kv_inter_group_insert
  // (kernel.cu:189-291, gtp:248-306): p = h0(key) % P, stable within p
  uint32_t *pos = &ph.par_pos[tile_id * (ph.P + 1)];
  std::vector<uint32_t> cnt(ph.P, 0);
  std::vector<uint32_t> pid(r);
  for (size_t i = 0; i < r; ++i) {
    K k = rk[i];
    pid[i] = o_murmur3_32(&k, sizeof(K), 0u) % (uint32_t)ph.P;
    cnt[pid[i]]++;
  }
  uint32_t run = 0;
  for (size_t p = 0; p < ph.P; ++p) { pos[p] = run; run += cnt[p]; }
  pos[ph.P] = (uint32_t)r;  // gtp:408
  std::vector<uint32_t> cur(pos, pos + ph.P);
  for (size_t i = 0; i < r; ++i) {
    size_t d = b + cur[pid[i]]++;
    ph.out.key[d] = rk[i];
    ph.out.st[d] = rs[i];
  }
  ph.tile_result[tile_id] = r;
}

template <typename K>
struct OResult {
  std::vector<OPar> pars;
};

// groupby_agg_partition (gtp:427-528): tiles pulled from a shared counter by
// nthreads workers (gtp:332-339); segments appended per partition
// (update_par_result_task, par_result.cu:3-13) in tile order.
template <typename K>
static void o_phase1(OPhase1<K> &ph, int nthreads, std::vector<OPar> &pars) {
  size_t tiles = (ph.n + ph.tile_len - 1) / ph.tile_len;
  ph.out.key.resize(ph.n);
  ph.out.st.resize(ph.n);
  ph.par_pos.assign(tiles * (ph.P + 1), 0);
  ph.tile_result.assign(tiles, 0);
  ph.counter = 0;
  auto worker = [&]() {
    OTable<K> tab(ph.tile_len);
    std::vector<uint8_t> flag(ph.tile_len);
    std::vector<K> rk; std::vector<OState> rs;
    for (;;) {
      size_t t = ph.counter.fetch_add(1);
      if (t >= tiles) break;
      o_tile(ph, t, tab, flag, rk, rs);
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nthreads; ++i) th.emplace_back(worker);
  for (auto &x : th) x.join();
  pars.assign(ph.P, OPar());
  for (size_t t = 0; t < tiles; ++t) {
    const uint32_t *pos = &ph.par_pos[t * (ph.P + 1)];
    for (size_t p = 0; p < ph.P; ++p) {
      uint32_t hi = p + 1 < ph.P ? pos[p + 1] : pos[ph.P];
      pars[p].push(t * ph.tile_len + pos[p], hi - pos[p]);
    }
  }
}

// merge_partition (groupby_agg_intra_partition.cu:13-48), with the same
// std::priority_queue as the reference so ties break identically.
struct OParRef { OPar *p; };
static std::vector<OPar> o_merge_partition(std::vector<OPar> &in, size_t K) {
  auto cmp = [](const OPar &l, const OPar &r) { return l.size > r.size; };
  std::priority_queue<OPar, std::vector<OPar>, decltype(cmp)> H(in.begin(), in.end(), cmp);
  for (;;) {
    if (H.size() == 1) break;
    OPar tmp = H.top(); H.pop();
    OPar top = H.top();
    if ((size_t)tmp.size + top.size > K) { H.push(tmp); break; }
    do {
      tmp.merge(top);
      H.pop();
      if (H.size() == 0) break;
      top = H.top();
    } while ((size_t)top.size + tmp.size <= K);
    H.push(tmp);
  }
  std::vector<OPar> rv;
  size_t len = H.size();
  for (size_t i = 0; i < len; ++i) { rv.push_back(H.top()); H.pop(); }
  return rv;
}

template <typename K>
struct OPhase2 {
  const ORun<K> *src;
  std::vector<OPar> pars;
  size_t capacity, min_load, max_load;
  std::vector<size_t> begin, count;
  ORun<K> out;
  std::atomic<size_t> counter{0};
  std::atomic<int> overflow{0};
};

// groupby_agg_intra_partition_thread (intra:159-279) for one merged partition.
template <typename K>
static void o_intra(OPhase2<K> &ph, size_t pi, OTable<K> &tab) {
  const OPar &pr = ph.pars[pi];
  size_t base = ph.begin[pi];
  // par_result_in_continous_mem (intra:89-107): gather segments in order
  size_t w = base;
  for (const OSeg &s : pr.seg)
    for (uint32_t j = 0; j < s.len; ++j, ++w) {
      ph.out.key[w] = ph.src->key[s.begin + j];
      ph.out.st[w] = ph.src->st[s.begin + j];
    }
  size_t C = ph.capacity;
  tab.reset(C);
  size_t rest = pr.size, already = 0;
  size_t tail = base + pr.size;  // pop2device pops from the back (par_result.cuh:48-74)
  std::vector<K> ck; std::vector<OState> cs;
  while (rest > 0) {
    size_t next = ph.max_load - already;
    size_t take;
    if (rest <= next) take = rest;
    else {
      if (already > ph.min_load) { ph.overflow = 1; return; }  // intra:232-236
      take = next;
    }
    size_t lo = tail - take;
    size_t added = 0;
    // K8 groupby_agg_and_statistic_kernel (kernel.cu:294-351): h0 % C, linear probing
    for (size_t i = lo; i < tail; ++i) {
      K k = ph.out.key[i];
      size_t loc = o_murmur3_32(&k, sizeof(K), 0u) % C;
      for (;;) {
        if (!tab.used[loc]) { tab.used[loc] = 1; tab.key[loc] = k; tab.st[loc] = ph.out.st[i]; added++; break; }
        if (tab.key[loc] == k) { o_state_add(tab.st[loc], ph.out.st[i]); break; }
        loc = (loc + 1) % C;
      }
    }
    tail = lo;
    rest -= take;
    already += added;
  }
  // exclusive scan + K10 collect_kv_in_ht (intra:259-276): slot order
  size_t o = base;
  for (size_t s = 0; s < C; ++s)
    if (tab.used[s]) { ph.out.key[o] = tab.key[s]; ph.out.st[o] = tab.st[s]; ++o; }
  ph.count[pi] = already;
}

template <typename K>
static int o_phase2(OPhase2<K> &ph, int nthreads) {
  size_t np = ph.pars.size();
  ph.begin.assign(np, 0);
  ph.count.assign(np, 0);
  size_t total = 0;
  for (size_t i = 0; i < np; ++i) { ph.begin[i] = total; total += ph.pars[i].size; }
  ph.out.key.resize(total);
  ph.out.st.resize(total);
  ph.counter = 0;
  ph.overflow = 0;
  int nt = (int)std::min<size_t>((size_t)nthreads, np);
  auto worker = [&]() {
    OTable<K> tab(ph.capacity);
    for (;;) {
      size_t i = ph.counter.fetch_add(1);
      if (i >= np || ph.overflow.load()) break;
      o_intra(ph, i, tab);
    }
  };
  std::vector<std::thread> th;
  for (int i = 0; i < nt; ++i) th.emplace_back(worker);
  for (auto &x : th) x.join();
  return ph.overflow.load() ? 4 : 0;
}

// --- extern "C" driver over the restatement ---------------------------------

extern "C" struct tga_o_run_stats {
  int64_t tiles;
  int64_t phase1_rows;     // sum of tile result counts r
  int64_t merged_parts;
  int64_t groups;
  double t_phase1_s, t_merge_s, t_phase2_s;
};

template <typename K, typename V>
static int o_run_ref(const K *keys, const V *vals, size_t n, size_t tile_len, size_t P,
                     size_t capacity, size_t min_load, size_t max_load, int nthreads,
                     K *out_keys, uint64_t *out_cnt, int64_t *out_sum, int64_t *out_min,
                     int64_t *out_max, size_t out_cap, int64_t *n_groups,
                     int64_t *par_begin, int64_t *par_count, int64_t par_cap, int64_t *n_par,
                     tga_o_run_stats *stats) {
  using clk = std::chrono::steady_clock;
  std::vector<OState> rows(n);
  for (size_t i = 0; i < n; ++i) {
    int64_t v = (int64_t)vals[i];
    rows[i] = OState{1, (uint64_t)v, v, v};
  }
  OPhase1<K> p1;
  p1.keys = keys; p1.rows = rows.data(); p1.n = n; p1.tile_len = tile_len; p1.P = P;
  std::vector<OPar> pars;
  auto t0 = clk::now();
  o_phase1(p1, nthreads, pars);
  auto t1 = clk::now();
  OPhase2<K> p2;
  p2.src = &p1.out;
  p2.pars = o_merge_partition(pars, min_load);
  auto t2 = clk::now();
  p2.capacity = capacity; p2.min_load = min_load; p2.max_load = max_load;
  int rc = o_phase2(p2, nthreads);
  auto t3 = clk::now();
  if (stats) {
    stats->tiles = (int64_t)p1.tile_result.size();
    int64_t r = 0;
    for (size_t x : p1.tile_result) r += (int64_t)x;
    stats->phase1_rows = r;
    stats->merged_parts = (int64_t)p2.pars.size();
    stats->t_phase1_s = std::chrono::duration<double>(t1 - t0).count();
    stats->t_merge_s = std::chrono::duration<double>(t2 - t1).count();
    stats->t_phase2_s = std::chrono::duration<double>(t3 - t2).count();
  }
  if (rc) return rc;
  size_t G = 0;
  for (size_t c : p2.count) G += c;
  if (stats) stats->groups = (int64_t)G;
  *n_groups = (int64_t)G;
  *n_par = (int64_t)p2.pars.size();
  if (par_begin && (int64_t)p2.pars.size() <= par_cap)
    for (size_t i = 0; i < p2.pars.size(); ++i) { par_begin[i] = (int64_t)p2.begin[i]; par_count[i] = (int64_t)p2.count[i]; }
  if (!out_keys) return 0;
  if (G > out_cap) return 9;
  size_t o = 0;
  for (size_t i = 0; i < p2.pars.size(); ++i)
    for (size_t j = 0; j < p2.count[i]; ++j, ++o) {
      size_t s = p2.begin[i] + j;
      out_keys[o] = p2.out.key[s];
      const OState &st = p2.out.st[s];
      if (out_cnt) out_cnt[o] = st.cnt;
      if (out_sum) out_sum[o] = (int64_t)st.sum;
      if (out_min) out_min[o] = st.mn;
      if (out_max) out_max[o] = st.mx;
    }
  return 0;
}

// The reference's own types: u32 keys, u32 values; the u32 wrap-around SUM is
// the low 32 bits of out_sum.  Outputs are in the reference's gapped layout
// compacted per merged partition (par_begin/par_count describe it).
OAPI int tga_o_ref_u32(const uint32_t *keys, const uint32_t *vals, size_t n, size_t tile_len, size_t P,
                  size_t capacity, size_t min_load, size_t max_load, int nthreads,
                  uint32_t *out_keys, uint64_t *out_cnt, int64_t *out_sum, int64_t *out_min,
                  int64_t *out_max, size_t out_cap, int64_t *n_groups,
                  int64_t *par_begin, int64_t *par_count, int64_t par_cap, int64_t *n_par,
                  tga_o_run_stats *stats) {
  return o_run_ref<uint32_t, uint32_t>(keys, vals, n, tile_len, P, capacity, min_load, max_load,
                                       nthreads, out_keys, out_cnt, out_sum, out_min, out_max,
                                       out_cap, n_groups, par_begin, par_count, par_cap, n_par, stats);
}

// Widened types (int64 key bits, int32 values) for the CPU baseline.
OAPI int tga_o_ref_i64(const uint64_t *keys, const int32_t *vals, size_t n, size_t tile_len, size_t P,
                  size_t capacity, size_t min_load, size_t max_load, int nthreads,
                  uint64_t *out_keys, uint64_t *out_cnt, int64_t *out_sum, int64_t *out_min,
                  int64_t *out_max, size_t out_cap, int64_t *n_groups,
                  int64_t *par_begin, int64_t *par_count, int64_t par_cap, int64_t *n_par,
                  tga_o_run_stats *stats) {
  return o_run_ref<uint64_t, int32_t>(keys, vals, n, tile_len, P, capacity, min_load, max_load,
                                      nthreads, out_keys, out_cnt, out_sum, out_min, out_max,
                                      out_cap, n_groups, par_begin, par_count, par_cap, n_par, stats);
}

// --- stage probes for the per-stage invariant tests --------------------------

// Phase 1 only, u32/u32: writes the partitioned tile results back into
// keys/vals (as gtp:413-418 does) and returns, per partition, the segment
// lists flattened as (partition, begin, len) triples in push order.
OAPI int64_t tga_o_ref_phase1_u32(uint32_t *keys, uint32_t *vals, size_t n, size_t tile_len, size_t P,
                             int nthreads, int64_t *seg_out, int64_t seg_cap) {
  std::vector<OState> rows(n);
  for (size_t i = 0; i < n; ++i) rows[i] = OState{1, (uint64_t)vals[i], (int64_t)vals[i], (int64_t)vals[i]};
  OPhase1<uint32_t> p1;
  p1.keys = keys; p1.rows = rows.data(); p1.n = n; p1.tile_len = tile_len; p1.P = P;
  std::vector<OPar> pars;
  o_phase1(p1, nthreads, pars);
  size_t tiles = p1.tile_result.size();
  for (size_t t = 0; t < tiles; ++t)
    for (size_t i = 0; i < p1.tile_result[t]; ++i) {
      size_t d = t * tile_len + i;
      keys[d] = p1.out.key[d];
      vals[d] = (uint32_t)p1.out.st[d].sum;
    }
  int64_t k = 0;
  for (size_t p = 0; p < P; ++p)
    for (const OSeg &s : pars[p].seg) {
      if (k < seg_cap) { seg_out[3 * k] = (int64_t)p; seg_out[3 * k + 1] = (int64_t)s.begin; seg_out[3 * k + 2] = s.len; }
      ++k;
    }
  return k;
}

// merge_partition alone: sizes in, merged sizes out (ascending, as returned).
OAPI int64_t tga_o_merge_sizes(const uint32_t *sizes, int64_t n, uint64_t K, uint64_t *out, int64_t cap) {
  std::vector<OPar> v((size_t)n);
  for (int64_t i = 0; i < n; ++i) v[(size_t)i].push((size_t)i, sizes[i]);
  std::vector<OPar> m = o_merge_partition(v, (size_t)K);
  for (size_t i = 0; i < m.size() && (int64_t)i < cap; ++i) out[i] = m[i].size;
  return (int64_t)m.size();
}

OAPI uint32_t tga_o_partition_of_u32(uint32_t key, uint32_t P) { return o_murmur3_32(&key, 4, 0u) % P; }

