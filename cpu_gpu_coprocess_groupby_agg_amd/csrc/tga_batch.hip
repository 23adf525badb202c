// tga_batch.hip -- thousands of independent group-by queries in ONE persistent
// launch fed by a device work queue, captured in a hipGraph.
//
// The reference runs one query per call and blocks six times per tile
// (T/src/group_then_partition.cu:97,173,221,386,409,418); a batch of small
// queries would pay those round trips and per-kernel launch costs per query.
// Here the host builds, once per batch, an ordered list of work items:
//
//   INIT     clear a slice of a query's global table
//   COUNT    partition histogram of one tile           (large-cardinality queries)
//   DIRECT   LDS aggregation of a row range -> table   (small-cardinality queries)
//   SCAN     exclusive scan of a query's tile histogram
//   SCATTER  LDS counting sort + coalesced runs of one tile
//   PAGG     LDS aggregation of one partition -> table
//   COMPACT  compaction of a slice of a query's table into its output
//
// Workgroups pull items from an atomic ticket.  An item may depend only on
// items earlier in the list (INIT < COUNT < DIRECT < SCAN < SCATTER < PAGG <
// COMPACT), so every item it waits for is finished or held by a running
// workgroup: the spin waits always end.  An item's results are complete and
// visible to every XCD before its stage counter is bumped: INIT writes back
// its plain stores (an agent-scope release); the others hand on their results
// through agent-coherent stores or device atomics and wait for them to
// complete.  PAGG and COMPACT, which read with plain loads, acquire after
// observing the counter.  Two workgroups share each CU.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "tga_batch_internal.hpp"
#include "tga_device.hpp"

namespace tga {

constexpr int kBatchThreads = 512;
constexpr uint32_t kCompactSlice = 65536;
constexpr int kPhases = 8;
// partition tiles: kSubTiles sub-tiles of kScatterItems rows per thread, each
// staged through LDS on its own
constexpr int kScatterItems = 8;
constexpr uint32_t kSubRows = kBatchThreads * kScatterItems;
constexpr uint32_t kSubTiles = 16;
constexpr uint32_t kTileRows = kSubRows * kSubTiles;

// Device clock for the optional profile (100 MHz).
__device__ __forceinline__ uint64_t batch_clock() { return __builtin_amdgcn_s_memrealtime(); }

// Optional per-phase clock of an item (thread 0's view; every phase ends at
// a barrier, so it is the workgroup's): mark(k) adds the ticks since the
// previous mark to phase k.
struct PhaseClock {
  uint64_t *p;
  uint64_t t;
  __device__ explicit PhaseClock(uint64_t *p_) : p(p_), t(p_ && threadIdx.x == 0 ? batch_clock() : 0) {}
  __device__ void mark(int k) {
    if (!p) return;
    lds_barrier();
    if (threadIdx.x == 0) {
      const uint64_t n = batch_clock();
      g_add((unsigned long long *)p + k, (unsigned long long)(n - t));
      t = n;
    }
  }
};

// Data one workgroup writes and another reads inside the launch (tile
// counts, offsets, the partitioned rows) is stored and loaded coherently at
// agent scope, access by access.  With every such store complete before its
// item is counted done, a consumer needs no cache-wide acquire, and a
// producer no cache-wide release (which writes back the whole L2 of its XCD).
template <typename T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(gp(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_agent(const T *p) {
  return __hip_atomic_load(gp(const_cast<T *>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wait_done(const BQuery &Q, uint32_t type) {
  if (Q.need[type] == 0) return;
  uint32_t *p = &Q.done[type];
  while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < Q.need[type]) __builtin_amdgcn_s_sleep(2);
}

// Item functions read their query record through a constant-memory pointer
// (the record is not written while the kernel runs), made wave-uniform: its
// fields are loaded once, by scalar loads, never through generic
// instructions.
__device__ __forceinline__ const BQuery &query_of(const TGA_CAS BQuery *p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return *(const BQuery *)(const TGA_CAS BQuery *)(((uint64_t)hi << 32) | lo);
}

template <typename K>
__device__ __forceinline__ void item_init(const TGA_CAS BQuery *Qp, uint64_t a, uint64_t b) {
  const BQuery &Q = query_of(Qp);
  const GTable &g = Q.g;
  const uint64_t n = g.S + 1;
  K *keys = (K *)g.keys;
  for (uint64_t i = a + threadIdx.x; i < b; i += blockDim.x) gp(keys)[i] = KeyTraits<K>::kEmpty;
  #pragma unroll
  for (int c = 0; c < kMaxAccs; ++c) {
    if (c >= Q.s.nacc) break;
    const int kind = Q.s.acc[c].kind;
    if (kind == ACC_SUM_F) {
      for (int l = 0; l < kLimbs; ++l)
        for (uint64_t i = a + threadIdx.x; i < b; i += blockDim.x) gp((uint64_t *)g.acc[c])[(uint64_t)l * n + i] = 0;
    } else if (acc_global_bytes(kind) == 4) {
      const uint32_t id = (uint32_t)acc_identity(kind);
      for (uint64_t i = a + threadIdx.x; i < b; i += blockDim.x) gp((uint32_t *)g.acc[c])[i] = id;
    } else {
      const uint64_t id = acc_identity(kind);
      for (uint64_t i = a + threadIdx.x; i < b; i += blockDim.x) gp((uint64_t *)g.acc[c])[i] = id;
    }
  }
}

// Partitions of a query: the top bits of the key's multiplicative hash, as in
// the plain operator's partitioned path.
template <typename K>
__device__ __forceinline__ uint32_t batch_partition(K k, int bits) {
  return key_partition((uint64_t)k, bits);
}

// Partition histogram of one tile; each thread has 16 key loads in flight.
template <typename K>
__device__ __forceinline__ void item_count(const TGA_CAS BQuery *Qp, uint32_t tile, uint32_t *hist) {
  const BQuery &Q = query_of(Qp);
  constexpr int R = 16;
  const uint32_t P = 1u << Q.pbits;
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) hist[p] = 0;
  lds_barrier();
  const K *keys = (const K *)Q.keys;
  const uint64_t b = (uint64_t)tile * Q.tile_rows, e = min(Q.n, b + Q.tile_rows);
  for (uint64_t r0 = b; r0 < e; r0 += (uint64_t)R * blockDim.x) {
    K kk[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const uint64_t r = r0 + (uint64_t)i * blockDim.x + threadIdx.x;
      kk[i] = r < e ? gp(keys)[r] : KeyTraits<K>::kEmpty;
    }
#pragma unroll
    for (int i = 0; i < R; ++i)
      if (r0 + (uint64_t)i * blockDim.x + threadIdx.x < e) atomicAdd(&hist[batch_partition(kk[i], Q.pbits)], 1u);
  }
  lds_barrier();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) st_agent(Q.cnt + (uint64_t)p * Q.ntiles + tile, hist[p]);
}

// One workgroup scans the whole [P][ntiles] histogram of a query.  (Barriers
// inside items order LDS accesses only: global loads and the agent-coherent
// stores stay in flight across them, and the item's end waits for all.)
__device__ __forceinline__ void item_scan(const TGA_CAS BQuery *Qp, uint32_t *buf, uint32_t *scratch) {
  const BQuery &Q = query_of(Qp);
  const uint32_t P = 1u << Q.pbits;
  const uint64_t cells = (uint64_t)P * Q.ntiles;
  const uint32_t chunk = blockDim.x * 8;
  uint32_t run = 0;
  for (uint64_t base = 0; base < cells; base += chunk) {
    const uint32_t m = (uint32_t)min<uint64_t>(chunk, cells - base);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) buf[i] = ld_agent(Q.cnt + base + i);
    lds_barrier();
    const uint32_t tot = block_exclusive_scan_lds<LdsBarrier>(buf, m, scratch);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
      const uint64_t idx = base + i;
      st_agent(Q.offs + idx, run + buf[i]);
      if (idx % Q.ntiles == 0) st_agent(Q.pstart + idx / Q.ntiles, (uint64_t)(run + buf[i]));  // first tile of p
    }
    run += tot;
    lds_barrier();
  }
  if (threadIdx.x == 0) st_agent(Q.pstart + P, Q.n);
}

// Rows [b, b + kSubRows) of a tile into registers (b < e).  Rows past e
// re-read row e - 1: every lane issues the same loads, no branches, so the
// compiler knows how many memory operations follow them (see item_scatter).
template <typename K>
__device__ __forceinline__ void load_sub(const K *keys, const uint32_t *vals, uint64_t b, uint64_t e, K *kk,
                                         uint32_t *vv) {
#pragma unroll
  for (int i = 0; i < kScatterItems; ++i) {
    const uint64_t row = min(b + (uint64_t)i * kBatchThreads + threadIdx.x, e - 1);
    kk[i] = gp(keys)[row];
    vv[i] = gp(vals)[row];
  }
}

// A tile of rows into the query's partitioned copy, one sub-tile at a time:
// LDS counting sort by partition, then each partition's run written
// contiguously at its cursor.  The next sub-tile's loads are issued before
// the current one is sorted.  A destination outside [0, n) (offsets that do
// not add up) is dropped and reported as an overflow, never written.
template <typename K>
__device__ __forceinline__ void item_scatter(const TGA_CAS BQuery *Qp, uint32_t tile, char *smem, uint32_t *scratch,
                                          uint64_t *ph) {
  const BQuery &Q = query_of(Qp);
  PhaseClock pc(ph);
  constexpr int IT = kScatterItems;
  const uint32_t P = 1u << Q.pbits;
  uint32_t *hist = (uint32_t *)smem;
  uint32_t *goff = hist + P;  // output cursor per partition
  K *skey = (K *)(goff + P);
  uint32_t *sval = (uint32_t *)(skey + kSubRows);  // one 32-bit value column
  uint16_t *spid = (uint16_t *)(sval + kSubRows);
  const K *keys = (const K *)Q.keys;
  const uint32_t *vals = (const uint32_t *)Q.s.cols[0];
  K *pk = (K *)Q.pkeys;
  uint32_t *pv = (uint32_t *)Q.pcols[0];
  uint64_t *prec = (uint64_t *)Q.pkeys;  // 4-byte keys: (key, value) records
  const uint64_t tb = (uint64_t)tile * Q.tile_rows, te = min(Q.n, tb + Q.tile_rows);
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) goff[p] = ld_agent(Q.offs + (uint64_t)p * Q.ntiles + tile);
  K kk[IT];
  uint32_t vv[IT];
  load_sub<K>(keys, vals, tb, te, kk, vv);
  bool bad = false;
  for (uint64_t sb = tb; sb < te; sb += kSubRows) {
    for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) hist[p] = 0;
    lds_barrier();
    K ck[IT];
    uint32_t cv[IT], pp[IT], rk[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const bool v = sb + (uint64_t)i * blockDim.x + threadIdx.x < te;
      ck[i] = kk[i];
      cv[i] = vv[i];
      pp[i] = batch_partition(ck[i], Q.pbits);
      rk[i] = v ? atomicAdd(&hist[pp[i]], 1u) : 0u;
    }
    if (sb + kSubRows < te) load_sub<K>(keys, vals, sb + kSubRows, te, kk, vv);
    lds_barrier();
    pc.mark(0);
    block_exclusive_scan_lds<LdsBarrier>(hist, P, scratch);
#pragma unroll
    for (int i = 0; i < IT; ++i) {
      const uint64_t row = sb + (uint64_t)i * blockDim.x + threadIdx.x;
      if (row < te) {
        const uint32_t lp = hist[pp[i]] + rk[i];
        skey[lp] = ck[i];
        sval[lp] = cv[i];
        spid[lp] = (uint16_t)pp[i];
      }
    }
    lds_barrier();
    pc.mark(1);
    // A fixed number of stores per lane, none skipped: loads and stores
    // share one counter, so the wait for the next sub-tile's keys (issued
    // before these stores) can leave exactly these stores in flight; with a
    // data-dependent count the compiler would wait for every store to reach
    // memory first.  Lanes past the sub-tile's rows store its last row again
    // (same value, same place); a destination outside [0, n) (offsets that
    // do not add up) is reported and goes to row 0 of this query's copy,
    // whose result is then discarded and the query rerun alone.
    const uint32_t rows = (uint32_t)min<uint64_t>(kSubRows, te - sb);
#pragma unroll
    for (uint32_t r = 0; r < kScatterItems; ++r) {
      const uint32_t j = min(r * kBatchThreads + threadIdx.x, rows - 1);
      const uint32_t p = spid[j];
      uint64_t dst = (uint64_t)goff[p] + (j - hist[p]);
      bad |= dst >= Q.n;
      dst = dst < Q.n ? dst : 0;
      if constexpr (sizeof(K) == 4) {
        st_agent(prec + dst, (uint64_t)sval[j] << 32 | (uint64_t)(uint32_t)skey[j]);
      } else {
        st_agent(pk + dst, skey[j]);
        st_agent(pv + dst, sval[j]);
      }
    }
    lds_barrier();
    for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) goff[p] += (p + 1 < P ? hist[p + 1] : rows) - hist[p];
    lds_barrier();
    pc.mark(2);
  }
  if (bad) g_or(Q.g.flags + FLAG_OVERFLOW, 1u);
}

// DIRECT items (a row range) flush their table into the query's global
// table.  Items that see every row of their keys -- PAGG items (a whole
// partition: its keys occur in no other item) and the DIRECT item of a query
// small enough for one item -- emit their table straight into the query's
// output, unless a row had to spill to the global table; then the table is
// flushed there as well.
// part: the rows are the query's partitioned copy (a PAGG item), else its
// input columns.
template <typename K, bool FX, int NC>
__device__ __forceinline__ void item_agg(const TGA_CAS BQuery *Qp, bool part, uint64_t b, uint64_t e, char *smem,
                                      uint32_t *sentinel, bool whole, uint32_t *spilled, uint32_t *scratch,
                                      uint64_t *s_base, uint64_t *ph) {
  const BQuery &Q = query_of(Qp);
  const Schema &s = Q.s;
  const void *keys = part ? Q.pkeys : Q.keys;
  PhaseClock pc(ph);
  LdsTable<K> t{smem, Q.L, sentinel};
  lds_init(t, s);
  if (threadIdx.x == 0) *spilled = FX ? 1u : 0u;  // fp64 limbs: always through the global table
  lds_barrier();
  pc.mark(whole ? 6 : 3);
  if (sizeof(K) == 4 && part) {
    // the partitioned copy of a 4-byte-key query: 8-byte (key, value) records
    const void *cols[1] = {nullptr};
    agg_rows_int<K, 1, kVwRec4, false>((const K *)Q.pkeys, cols, b, e, t, Q.ap, Q.nbits, Q.max_probe, &Q.g, &Q.s,
                                       nullptr, spilled);
  } else if constexpr (FX) {
    const void *cols[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) cols[c] = part ? Q.pcols[c] : s.cols[c];
    agg_rows_int<K, NC, 0, false, 0, true>((const K *)keys, cols, b, e, t, Q.ap, Q.nbits, Q.max_probe, &Q.g, &Q.s,
                                           nullptr, spilled);
  } else {
    const void *cols[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) cols[c] = part ? Q.pcols[c] : s.cols[c];
    if constexpr (NC == 1) {
      if (dtype_bytes((int)(Q.ap.coldt & 15u)) == 8)
        agg_rows_int<K, 1, 8, false>((const K *)keys, cols, b, e, t, Q.ap, Q.nbits, Q.max_probe, &Q.g, &Q.s,
                                     nullptr, spilled);
      else
        agg_rows_int<K, 1, 4, false>((const K *)keys, cols, b, e, t, Q.ap, Q.nbits, Q.max_probe, &Q.g, &Q.s,
                                     nullptr, spilled);
    } else {
      agg_rows_int<K, NC, 0, false>((const K *)keys, cols, b, e, t, Q.ap, Q.nbits, Q.max_probe, &Q.g, &Q.s,
                                    nullptr, spilled);
    }
  }
  __syncthreads();
  pc.mark(whole ? 7 : 4);
  if (whole && wave_uniform(*spilled) == 0) lds_emit<K>(t, s, Q.od, Q.count_out, Q.g.flags, scratch, s_base);
  else lds_flush<K>(t, s, Q.g);
  pc.mark(5);
}

// Compacts table slots [a, b) of one query; output rows are reserved with an
// atomic on the query's counter (order inside a query is unspecified).
template <typename K>
__device__ __forceinline__ void item_compact(const TGA_CAS BQuery *Qp, uint64_t a, uint64_t b, uint32_t *scratch,
                                          uint64_t *s_base) {
  const BQuery &Q = query_of(Qp);
  const GTable &g = Q.g;
  const uint64_t n = g.S + 1;
  const K *keys = (const K *)g.keys;
  for (uint64_t base = a; base < b; base += blockDim.x) {
    const uint64_t slot = base + threadIdx.x;
    bool occ = false;
    if (slot < b) {
      if (slot < g.S) occ = gp(keys)[slot] != KeyTraits<K>::kEmpty;
      else occ = gp(g.flags)[FLAG_SENTINEL] != 0;
    }
    const uint64_t bal = __ballot(occ);
    const uint32_t pre = (uint32_t)__builtin_popcountll(bal & lanemask_lt());
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0) scratch[wave] = (uint32_t)__builtin_popcountll(bal);
    lds_barrier();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (uint32_t w = 0; w < blockDim.x / kWave; ++w) { const uint32_t c = scratch[w]; scratch[w] = run; run += c; }
      *s_base = run ? g_add((unsigned long long *)Q.count_out, (unsigned long long)run) : 0;
    }
    lds_barrier();
    if (occ) {
      const uint64_t pos = *s_base + scratch[wave] + pre;
      if (pos >= Q.od.capacity) g_or(Q.g.flags + FLAG_OVERFLOW, 1u);
      else {
        gp((K *)Q.od.keys)[pos] = slot < g.S ? gp(keys)[slot] : KeyTraits<K>::kEmpty;
        #pragma unroll
  for (int c = 0; c < kMaxAccs; ++c) {
    if (c >= Q.s.nacc) break;
          const int kind = Q.s.acc[c].kind;
          if (kind == ACC_COUNT || kind == ACC_SUM_I) gp((uint64_t *)Q.od.cols[c])[pos] = gp((const uint64_t *)g.acc[c])[slot];
          else if (kind == ACC_MIN_I32 || kind == ACC_MAX_I32) gp((uint32_t *)Q.od.cols[c])[pos] = gp((const uint32_t *)g.acc[c])[slot];
          else if (kind == ACC_SUM_F) {
            int64_t Lm[kLimbs];
            for (int l = 0; l < kLimbs; ++l) Lm[l] = gp((const int64_t *)g.acc[c])[(uint64_t)l * n + slot];
            gp((double *)Q.od.cols[c])[pos] = fx_to_double(Lm);
          } else {
            const int64_t v = gp((const int64_t *)g.acc[c])[slot];
            const int dt = Q.s.acc[c].col_dtype;
            if (dt == DT_U64) gp((uint64_t *)Q.od.cols[c])[pos] = (uint64_t)v ^ (1ULL << 63);
            else if (dt == DT_F64) gp((uint64_t *)Q.od.cols[c])[pos] = ordered_to_f64_bits(v);
            else if (dt == DT_U32 || dt == DT_I32) gp((uint32_t *)Q.od.cols[c])[pos] = (uint32_t)v;
            else gp((int64_t *)Q.od.cols[c])[pos] = v;
          }
        }
      }
    }
    lds_barrier();
  }
}

// The same items as functions of their own, each with its own register
// allocation: the kernel variants for fp64 SUM or several value columns call
// these, since all items inlined into one kernel would spill in their row
// loops; the one-column integer variant inlines every item (no call saves
// and restores of registers per item).
template <typename K>
__device__ __noinline__ void item_init_fn(const TGA_CAS BQuery *Qp, uint64_t a, uint64_t b) {
  item_init<K>(Qp, a, b);
}
template <typename K>
__device__ __noinline__ void item_count_fn(const TGA_CAS BQuery *Qp, uint32_t tile, uint32_t *hist) {
  item_count<K>(Qp, tile, hist);
}
__device__ __noinline__ void item_scan_fn(const TGA_CAS BQuery *Qp, uint32_t *buf, uint32_t *scratch) {
  item_scan(Qp, buf, scratch);
}
template <typename K>
__device__ __noinline__ void item_scatter_fn(const TGA_CAS BQuery *Qp, uint32_t tile, char *smem, uint32_t *scratch,
                                             uint64_t *ph) {
  item_scatter<K>(Qp, tile, smem, scratch, ph);
}
template <typename K, bool FX, int NC>
__device__ __noinline__ void item_agg_fn(const TGA_CAS BQuery *Qp, bool part, uint64_t b, uint64_t e, char *smem,
                                         uint32_t *sentinel, bool whole, uint32_t *spilled, uint32_t *scratch,
                                         uint64_t *s_base, uint64_t *ph) {
  item_agg<K, FX, NC>(Qp, part, b, e, smem, sentinel, whole, spilled, scratch, s_base, ph);
}
template <typename K>
__device__ __noinline__ void item_compact_fn(const TGA_CAS BQuery *Qp, uint64_t a, uint64_t b, uint32_t *scratch,
                                             uint64_t *s_base) {
  item_compact<K>(Qp, a, b, scratch, s_base);
}

template <typename K, bool FX, int NC>
__device__ __forceinline__ void run_item(const TGA_CAS BQuery *Qp, const BItem &it, char *smem, uint32_t *scratch,
                                         uint32_t *sentinel, uint64_t *s_base, uint32_t *spilled, uint64_t *ph) {
  constexpr bool kInl = !FX && NC == 1;
  const BQuery &Q = query_of(Qp);
  switch (it.type) {
    case IT_INIT:
      if constexpr (kInl) item_init<K>(Qp, it.a, it.b);
      else item_init_fn<K>(Qp, it.a, it.b);
      break;
    case IT_COUNT:
      if constexpr (kInl) item_count<K>(Qp, (uint32_t)it.a, (uint32_t *)smem);
      else item_count_fn<K>(Qp, (uint32_t)it.a, (uint32_t *)smem);
      break;
    case IT_DIRECT: {
      const bool whole = it.a == 0 && it.b == Q.n;
      if constexpr (kInl) item_agg<K, FX, NC>(Qp, false, it.a, it.b, smem, sentinel, whole, spilled, scratch, s_base, ph);
      else item_agg_fn<K, FX, NC>(Qp, false, it.a, it.b, smem, sentinel, whole, spilled, scratch, s_base, ph);
      break;
    }
    case IT_SCAN:
      if constexpr (kInl) item_scan(Qp, (uint32_t *)smem, scratch);
      else item_scan_fn(Qp, (uint32_t *)smem, scratch);
      break;
    case IT_SCATTER:
      if constexpr (kInl) item_scatter<K>(Qp, (uint32_t)it.a, smem, scratch, ph);
      else item_scatter_fn<K>(Qp, (uint32_t)it.a, smem, scratch, ph);
      break;
    case IT_PAGG: {
      // a range outside [0, n) (offsets that do not add up) is cut to it
      const uint64_t e = min(ld_agent(Q.pstart + it.a + 1), Q.n);
      const uint64_t b = min(ld_agent(Q.pstart + it.a), e);
      if constexpr (kInl) item_agg<K, FX, NC>(Qp, true, b, e, smem, sentinel, true, spilled, scratch, s_base, ph);
      else item_agg_fn<K, FX, NC>(Qp, true, b, e, smem, sentinel, true, spilled, scratch, s_base, ph);
      break;
    }
    case IT_COMPACT:
      if constexpr (kInl) item_compact<K>(Qp, it.a, it.b, scratch, s_base);
      else item_compact_fn<K>(Qp, it.a, it.b, scratch, s_base);
      break;
    default: break;
  }
}

// prof (optional, 4 * kItemTypes + kPhases words): per item type the clock
// ticks its items spent working, the ticks spent waiting for the items they
// depend on, the item count, and the ticks spent acquiring and publishing
// around each item; then the ticks of the phases inside items
// (scatter: rank, local sort, write-out; row-range aggregation: table
// set-up, rows; both aggregations: flush or emission; partition
// aggregation: table set-up, rows).
// Two workgroups per CU (each with at most half of the LDS, at most 128
// VGPRs): one's barriers and memory waits overlap the other's work.
template <bool FX, int NC>
__global__ void __launch_bounds__(kBatchThreads, 2 * kBatchThreads / kWave / 4) k_batch(const BQuery *__restrict__ qs, const BItem *__restrict__ items,
                                                         uint32_t nitems, uint32_t *ticket, uint64_t *prof) {
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t s_item, s_sentinel, s_spilled;
  __shared__ uint32_t scratch[kBatchThreads / kWave + 1];
  __shared__ uint64_t s_base;
  // the item finished last by this workgroup; its completion is published
  // together with the next ticket fetch (one lane-0 region per iteration)
  uint32_t prev_q = 0xffffffffu, prev_type = 0;
  for (;;) {
    if (threadIdx.x == 0) {
      if (prev_q != 0xffffffffu)
        __hip_atomic_fetch_add(&qs[prev_q].done[prev_type], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_item = atomicAdd(ticket, 1u);
    }
    __syncthreads();
    const uint32_t i = wave_uniform(s_item);
    if (i >= nitems) break;
    const BItem it = items[i];
    const BQuery &Q = qs[it.q];
    const uint64_t c0 = prof && threadIdx.x == 0 ? batch_clock() : 0;
    if (threadIdx.x == 0) {
      switch (it.type) {
        case IT_DIRECT: wait_done(Q, IT_INIT); break;
        case IT_SCAN: wait_done(Q, IT_COUNT); break;
        case IT_SCATTER: wait_done(Q, IT_SCAN); break;
        case IT_PAGG: wait_done(Q, IT_SCATTER); wait_done(Q, IT_INIT); break;
        case IT_COMPACT: wait_done(Q, IT_DIRECT); wait_done(Q, IT_PAGG); wait_done(Q, IT_INIT); break;
        default: break;
      }
    }
    __syncthreads();
    const uint64_t c1 = prof && threadIdx.x == 0 ? batch_clock() : 0;
    // PAGG reads the partitioned rows and COMPACT the global table with plain
    // loads: lines of them cached in this XCD's L2 before their writers
    // finished are dropped first
    if (it.type == IT_PAGG || it.type == IT_COMPACT) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint64_t c1f = prof && threadIdx.x == 0 ? batch_clock() : 0;
    uint64_t *ph = prof ? prof + 4 * kItemTypes : nullptr;
    const TGA_CAS BQuery *Qp = (const TGA_CAS BQuery *)(qs + it.q);
    if (Q.key_bytes == 4) run_item<uint32_t, FX, NC>(Qp, it, smem, scratch, &s_sentinel, &s_base, &s_spilled, ph);
    else run_item<uint64_t, FX, NC>(Qp, it, smem, scratch, &s_sentinel, &s_base, &s_spilled, ph);
    // profiling: every wave's work ends before the release is timed
    if (prof) __syncthreads();
    const uint64_t c2w = prof && threadIdx.x == 0 ? batch_clock() : 0;
    // Publication before the item is counted done (by a relaxed increment
    // after the barrier).  INIT's plain stores are written back from L2.
    // Every other item hands on its results through agent-coherent stores
    // or device-scope atomics, which are visible to every XCD once complete:
    // the wave waits for all of its memory operations.
    if (it.type == IT_INIT) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    } else {
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      __builtin_amdgcn_s_waitcnt(0);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
    }
    __syncthreads();
    if (prof && threadIdx.x == 0) {
      const uint64_t c2 = batch_clock();
      atomicAdd((unsigned long long *)&prof[it.type], (unsigned long long)(c2w - c1f));
      atomicAdd((unsigned long long *)&prof[kItemTypes + it.type], (unsigned long long)(c1 - c0));
      atomicAdd((unsigned long long *)&prof[2 * kItemTypes + it.type], 1ull);
      atomicAdd((unsigned long long *)&prof[3 * kItemTypes + it.type], (unsigned long long)((c1f - c1) + (c2 - c2w)));
    }
    prev_q = it.q;
    prev_type = it.type;
  }
}

__global__ void k_batch_reset(uint32_t *done, uint64_t ndone, uint64_t *counts, uint64_t nq, uint32_t *ticket,
                              uint32_t *flags, uint64_t *prof) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < ndone; i += (uint64_t)gridDim.x * blockDim.x) done[i] = 0;
  if (prof && blockIdx.x == 0 && threadIdx.x < 4 * kItemTypes + kPhases) prof[threadIdx.x] = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nq * kFlagWords; i += (uint64_t)gridDim.x * blockDim.x)
    flags[i] = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nq; i += (uint64_t)gridDim.x * blockDim.x) counts[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) *ticket = 0;
}

}  // namespace tga


namespace tga {

namespace {
uint64_t np2(uint64_t x) { uint64_t p = 1; while (p < x) p <<= 1; return p; }
int lg2(uint64_t x) { int b = 0; while ((1ULL << b) < x) ++b; return b; }

// Per-query shape decisions of the batched path (kept in one place so the
// admission estimate and the build agree).
struct QShape {
  bool part = false;
  uint32_t T = 256;
  uint64_t S = 256;
  uint64_t cap = 1;
  int pbits = 0;
  uint32_t tile_rows = 0, ntiles = 0;
};

// LDS per workgroup: two workgroups share a CU (less the kernel's static
// LDS, a few hundred bytes).
size_t batch_lds_budget(const tga_context *ctx) { return ctx->lds_max / 2 - 1024; }

// LDS of a scatter item: histogram and cursors per partition, one sub-tile
// of keys, values and partition ids.
size_t scatter_lds(int pbits, int key_bytes) {
  return ((size_t)8 << pbits) + (size_t)kSubRows * (key_bytes + 4 + 2);
}

QShape shape_of(const tga_context *ctx, const QuerySpec &q, uint64_t n, uint64_t est) {
  QShape z;
  const size_t lds_max = batch_lds_budget(ctx);
  const uint64_t G = std::max<uint64_t>(1, std::min<uint64_t>(est, std::max<uint64_t>(n, 1)));
  const uint32_t sb = lds_slot_bytes(q.schema, q.key_bytes);
  uint32_t tmax = 256;
  while ((uint64_t)(tmax * 2 + 2) * sb + 256 <= lds_max && tmax < 8192) tmax *= 2;
  z.part = q.schema.ncols == 1 && dtype_bytes(q.schema.col_dtype[0]) == 4 && 2 * G > tmax && n > 65536;
  z.T = z.part ? tmax : (uint32_t)std::min<uint64_t>(tmax, std::max<uint64_t>(256, np2(2 * G)));
  z.S = std::min<uint64_t>(np2(std::max<uint64_t>(256, 2 * G)), np2(std::max<uint64_t>(256, 2 * n)));
  z.cap = std::min<uint64_t>(z.S + 1, n + 1);
  // partitioned: whole partitions emit into the output; the global table
  // only takes the rows of partitions that spill (a miss is rerun alone)
  if (z.part) z.S = std::max<uint64_t>(4096, z.S / 32);
  if (z.part) {
    // partitions past the scatter's LDS are not made: larger partitions
    // spill part of their rows to the global table instead
    int pmax = 12;
    while (pmax > 1 && scatter_lds(pmax, q.key_bytes) > lds_max) --pmax;
    z.pbits = std::max(1, std::min(pmax, lg2(np2((2 * G + z.T - 1) / z.T))));
    z.tile_rows = kTileRows;
    z.ntiles = (uint32_t)((n + z.tile_rows - 1) / z.tile_rows);
  }
  return z;
}
}  // namespace

BatchExec::~BatchExec() {
  if (ctx) {
    ctx->pinned->put(h_counts, (uint64_t)std::max(1, nq) * 8);
    ctx->pinned->put(h_flags, (uint64_t)std::max(1, nq) * kFlagWords * 4);
    ctx->pinned->put(h_stage, stage_bytes);
  }
}

uint64_t batch_query_bytes(const tga_context *ctx, const QuerySpec &q, uint64_t n, uint64_t est) {
  const QShape z = shape_of(ctx, q, n, est);
  uint64_t b = (z.S + 1) * q.key_bytes;
  for (int a = 0; a < q.schema.nacc; ++a) b += (z.S + 1) * acc_global_bytes(q.schema.acc[a].kind);
  b += z.cap * q.key_bytes;
  for (int a = 0; a < q.schema.nacc; ++a) b += z.cap * dtype_bytes(q.out_dtype[a]);
  if (z.part) {
    const uint64_t cells = (uint64_t)(1u << z.pbits) * z.ntiles;
    b += cells * 8 + ((1ULL << z.pbits) + 1) * 8 + n * (q.key_bytes + 4);
  }
  return b + 64 * 1024;
}

tga_status batch_prepare(tga_context *ctx, const tga_query *queries, int32_t n, BatchExec &b) {
  b.ctx = ctx;
  b.nq = n;
  b.spec.resize(n);
  b.rows.resize(n);
  b.keys.resize(n);
  for (int32_t i = 0; i < n; ++i) {
    const tga_query &q = queries[i];
    std::string e = build_spec(&q.key, q.values, q.n_values, q.aggs, q.n_aggs, b.spec[i]);
    if (e.empty() && q.n_rows < 0) e = "negative row count";
    if (e.empty() && q.key.location != TGA_DEVICE) e = "batched queries need device-resident columns";
    for (int c = 0; e.empty() && c < q.n_values; ++c)
      if (q.values[c].location != TGA_DEVICE) e = "batched queries need device-resident columns";
    if (e.empty() && q.n_rows > 0 && !q.key.data) e = "null key column";
    if (!e.empty()) return fail(ctx, TGA_ERR_INVALID_ARG, "query %d: %s", i, e.c_str());
    b.rows[i] = (uint64_t)q.n_rows;
    b.keys[i] = q.key.data;
  }
  // estimates: the caller's hint, else up to 1024 keys per query at random rows
  // copied in one transfer and fed to the GEE estimator
  b.est.assign(n, 0);
  std::vector<int32_t> todo;
  for (int32_t i = 0; i < n; ++i) {
    if (queries[i].est_groups > 0) b.est[i] = (uint64_t)queries[i].est_groups;
    else if (queries[i].n_rows > 0) todo.push_back(i);
    else b.est[i] = 1;
  }
  if (todo.empty()) return TGA_OK;
  const uint64_t M = 1024;
  DBuf buf;
  if (!buf.alloc(ctx, todo.size() * M * 8)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch sample");
  std::vector<uint64_t> ms(todo.size());
  for (size_t j = 0; j < todo.size(); ++j) {
    const tga_query &q = queries[todo[j]];
    ms[j] = std::min<uint64_t>(M, (uint64_t)q.n_rows);
    TGA_HIP(ctx, launch_sample_keys(dtype_bytes(q.key.dtype), q.key.data, (uint64_t)q.n_rows, ms[j],
                                    buf.as<uint64_t>() + j * M, ctx->stream));
  }
  std::vector<uint64_t> h(todo.size() * M);
  TGA_HIP(ctx, hipMemcpyAsync(h.data(), buf.p, h.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (size_t j = 0; j < todo.size(); ++j)
    b.est[todo[j]] = estimate_from_sample(h.data() + j * M, ms[j], (uint64_t)queries[todo[j]].n_rows);
  return TGA_OK;
}

tga_status batch_build(BatchExec &b, const tga_query *qs, hipStream_t upload) {
  tga_context *ctx = b.ctx;
  const int32_t n = b.nq;
  b.hq.assign(n, BQuery{});
  b.cap.assign(n, 0);
  b.okeys.clear();
  b.okeys.resize(n);
  b.ocols.clear();
  b.ocols.resize(n);
  b.graph_out.assign(n, 1);
  std::vector<std::vector<BItem>> lists(kItemTypes);
  size_t lds_need = 0;
  const uint64_t total_done = (uint64_t)n * kItemTypes;
  if (!b.ddone.alloc(ctx, total_done * 4) || !b.dcounts.alloc(ctx, (uint64_t)n * 8) || !b.dticket.alloc(ctx, 64) ||
      !b.dflags.alloc(ctx, (uint64_t)n * kFlagWords * 4))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch counters");
  // the per-item-type profile, when kernel timing is on for the context
  if (ctx->prof.on && !b.dprof.alloc(ctx, (4 * kItemTypes + kPhases) * 8)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch profile");
  b.h_counts = (uint64_t *)ctx->pinned->get((uint64_t)std::max(1, n) * 8);
  b.h_flags = (uint32_t *)ctx->pinned->get((uint64_t)std::max(1, n) * kFlagWords * 4);
  if (!b.h_counts || !b.h_flags) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "pinned batch status");
  for (int32_t i = 0; i < n; ++i) {
    const QuerySpec &q = b.spec[i];
    BQuery &Q = b.hq[i];
    Q.s = q.schema;
    Q.keys = qs[i].key.data;
    Q.n = b.rows[i];
    Q.key_bytes = q.key_bytes;
    Q.done = b.ddone.as<uint32_t>() + (uint64_t)i * kItemTypes;
    Q.count_out = b.dcounts.as<uint64_t>() + i;
    const QShape z = shape_of(ctx, q, Q.n, b.est[i]);
    Q.path = z.part ? TGA_PATH_PARTITIONED : TGA_PATH_DIRECT;
    Q.L = make_layout(q.schema, q.key_bytes, z.T);
    Q.nbits = lg2((uint64_t)z.T + 1);
    Q.max_probe = (int)std::min<uint32_t>(z.T / 4, 64);  // in 4-slot buckets
    Q.ap = make_acc_prog(q.schema, Q.L);
    lds_need = std::max<size_t>(lds_need, Q.L.bytes);
    // global table; its status words live in the batch's flag block
    {
      const uint64_t S = z.S;
      DBuf k;
      if (!k.alloc(ctx, (S + 1) * q.key_bytes)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch tables");
      Q.g.keys = k.p;
      b.bufs.push_back(std::move(k));
      for (int a = 0; a < q.schema.nacc; ++a) {
        DBuf c;
        if (!c.alloc(ctx, (S + 1) * acc_global_bytes(q.schema.acc[a].kind)))
          return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch tables");
        Q.g.acc[a] = c.p;
        b.bufs.push_back(std::move(c));
      }
      Q.g.S = S;
      Q.g.flags = b.dflags.as<uint32_t>() + (uint64_t)i * kFlagWords;
    }
    b.cap[i] = z.cap;
    Q.od.capacity = z.cap;
    if (!b.okeys[i].alloc(ctx, z.cap * q.key_bytes)) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch outputs");
    Q.od.keys = b.okeys[i].p;
    for (int a = 0; a < q.schema.nacc; ++a) {
      if (!b.ocols[i][a].alloc(ctx, z.cap * dtype_bytes(q.out_dtype[a])))
        return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch outputs");
      Q.od.cols[a] = b.ocols[i][a].p;
    }
    for (uint64_t a = 0; a < z.S + 1; a += kCompactSlice)
      lists[IT_INIT].push_back({(uint32_t)i, IT_INIT, a, std::min<uint64_t>(z.S + 1, a + kCompactSlice)});
    if (Q.n > 0 && !z.part) {
      // enough rows per item to amortise the table flush (a flush of T slots
      // costs about as much as 8T rows)
      const uint64_t per = std::max<uint64_t>(262144, std::min<uint64_t>(1 << 20, 256 * (uint64_t)z.T));
      for (uint64_t a = 0; a < Q.n; a += per)
        lists[IT_DIRECT].push_back({(uint32_t)i, IT_DIRECT, a, std::min<uint64_t>(Q.n, a + per)});
    } else if (Q.n > 0) {
      Q.pbits = z.pbits;
      const uint32_t PP = 1u << z.pbits;
      Q.tile_rows = z.tile_rows;
      Q.ntiles = z.ntiles;
      const uint64_t cells = (uint64_t)PP * Q.ntiles;
      // the partitioned copy: 4-byte keys as 8-byte (key, value) records in
      // one array, 8-byte keys as a key array and a value array
      const bool rec = q.key_bytes == 4;
      DBuf c1, c2, c3, c4, c5;
      if (!c1.alloc(ctx, cells * 4) || !c2.alloc(ctx, cells * 4) || !c3.alloc(ctx, (PP + 1) * 8) ||
          !c4.alloc(ctx, Q.n * (rec ? 8 : q.key_bytes)) || (!rec && !c5.alloc(ctx, Q.n * 4)))
        return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch partition buffers");
      Q.cnt = c1.as<uint32_t>();
      Q.offs = c2.as<uint32_t>();
      Q.pstart = c3.as<uint64_t>();
      Q.pkeys = c4.p;
      Q.pcols[0] = rec ? nullptr : c5.p;
      b.bufs.push_back(std::move(c1));
      b.bufs.push_back(std::move(c2));
      b.bufs.push_back(std::move(c3));
      b.bufs.push_back(std::move(c4));
      if (!rec) b.bufs.push_back(std::move(c5));
      lds_need = std::max(lds_need, scatter_lds(z.pbits, q.key_bytes));
      lds_need = std::max<size_t>(lds_need, (size_t)4 * std::min<uint64_t>(cells, kBatchThreads * 8));
      for (uint32_t t = 0; t < Q.ntiles; ++t) lists[IT_COUNT].push_back({(uint32_t)i, IT_COUNT, t, 0});
      lists[IT_SCAN].push_back({(uint32_t)i, IT_SCAN, 0, 0});
      for (uint32_t t = 0; t < Q.ntiles; ++t) lists[IT_SCATTER].push_back({(uint32_t)i, IT_SCATTER, t, 0});
      for (uint32_t p = 0; p < PP; ++p) lists[IT_PAGG].push_back({(uint32_t)i, IT_PAGG, p, 0});
    }
    for (uint64_t a = 0; a < z.S + 1; a += kCompactSlice)
      lists[IT_COMPACT].push_back({(uint32_t)i, IT_COMPACT, a, std::min<uint64_t>(z.S + 1, a + kCompactSlice)});
  }
  for (auto &l : lists)
    for (auto &it : l) b.hq[it.q].need[it.type] += 1;
  if (lds_need > batch_lds_budget(ctx)) return fail(ctx, TGA_ERR_UNSUPPORTED, "batch needs %zu B of LDS", lds_need);
  // stage by stage; the largest queries first inside a stage
  std::vector<BItem> items;
  for (int t = 0; t < kItemTypes; ++t) {
    auto &l = lists[t];
    std::stable_sort(l.begin(), l.end(), [&](const BItem &x, const BItem &y) { return b.hq[x.q].n > b.hq[y.q].n; });
    items.insert(items.end(), l.begin(), l.end());
  }
  b.nitems = (uint32_t)items.size();
  if (!b.dq.alloc(ctx, std::max(1, n) * sizeof(BQuery)) || !b.ditems.alloc(ctx, std::max<size_t>(1, items.size()) * sizeof(BItem)))
    return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "batch queue");
  const uint64_t qbytes = (uint64_t)n * sizeof(BQuery), ibytes = items.size() * sizeof(BItem);
  if (upload) {
    // A server thread building the next flight must not wait here: while
    // other streams' kernels hold the CUs, an upload from pageable memory and
    // its wait cost it hundreds of microseconds.  The pinned staging stays
    // with the batch until it is destroyed, after its stream has finished.
    b.stage_bytes = std::max<uint64_t>(1, qbytes + ibytes);
    b.h_stage = ctx->pinned->get(b.stage_bytes);
    if (!b.h_stage) return fail(ctx, TGA_ERR_OUT_OF_MEMORY, "pinned batch queue");
    char *hs = (char *)b.h_stage;
    if (qbytes) std::memcpy(hs, b.hq.data(), qbytes);
    if (ibytes) std::memcpy(hs + qbytes, items.data(), ibytes);
    if (qbytes) TGA_HIP(ctx, hipMemcpyAsync(b.dq.p, hs, qbytes, hipMemcpyHostToDevice, upload));
    if (ibytes) TGA_HIP(ctx, hipMemcpyAsync(b.ditems.p, hs + qbytes, ibytes, hipMemcpyHostToDevice, upload));
  } else {
    if (qbytes) TGA_HIP(ctx, hipMemcpyAsync(b.dq.p, b.hq.data(), qbytes, hipMemcpyHostToDevice, ctx->stream));
    if (ibytes) TGA_HIP(ctx, hipMemcpyAsync(b.ditems.p, items.data(), ibytes, hipMemcpyHostToDevice, ctx->stream));
    TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  b.lds = std::max<size_t>(lds_need, 1024);
  // kernel variant: fp64 SUM anywhere in the batch / more than one value column
  bool fx = false, multi = false;
  for (int32_t i = 0; i < n; ++i) { fx |= b.spec[i].nfx > 0; multi |= b.spec[i].schema.ncols > 1; }
  b.kernel = fx ? (const void *)&k_batch<true, kMaxCols> : multi ? (const void *)&k_batch<false, kMaxCols>
                                                                 : (const void *)&k_batch<false, 1>;
  if (b.lds > 65536)
    TGA_HIP(ctx, raise_lds_limit(b.kernel, b.lds));
  const unsigned per_cu = (unsigned)std::max<size_t>(1, std::min<size_t>(4, ctx->lds_max / b.lds));
  b.grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ctx->cus * per_cu, b.nitems));
  return TGA_OK;
}

tga_status batch_enqueue(BatchExec &b, hipStream_t st) {
  tga_context *ctx = b.ctx;
  const uint64_t n = (uint64_t)b.nq;
  uint64_t *parg = b.dprof.p ? b.dprof.as<uint64_t>() : nullptr;
  hipLaunchKernelGGL(k_batch_reset, dim3(64), dim3(256), 0, st, b.ddone.as<uint32_t>(), n * kItemTypes,
                     b.dcounts.as<uint64_t>(), n, b.dticket.as<uint32_t>(), b.dflags.as<uint32_t>(), parg);
  TGA_HIP(ctx, hipGetLastError());
  if (b.nitems) {
    const BQuery *qarg = b.dq.as<BQuery>();
    const BItem *iarg = b.ditems.as<BItem>();
    uint32_t narg = b.nitems;
    uint32_t *targ = b.dticket.as<uint32_t>();
    void *args[] = {&qarg, &iarg, &narg, &targ, &parg};
    TGA_HIP(ctx, hipLaunchKernel(b.kernel, dim3(b.grid), dim3(kBatchThreads), args, b.lds, st));
  }
  if (n) {
    TGA_HIP(ctx, hipMemcpyAsync(b.h_counts, b.dcounts.p, n * 8, hipMemcpyDeviceToHost, st));
    TGA_HIP(ctx, hipMemcpyAsync(b.h_flags, b.dflags.p, n * kFlagWords * 4, hipMemcpyDeviceToHost, st));
  }
  return TGA_OK;
}

tga_status batch_finalize(BatchExec &b, std::vector<uint64_t> &groups) {
  tga_context *ctx = b.ctx;
  groups.assign(b.nq, 0);
  for (int32_t i = 0; i < b.nq; ++i) {
    const uint32_t *fl = b.h_flags + (uint64_t)i * kFlagWords;
    if (fl[FLAG_RANGE]) return fail(ctx, TGA_ERR_RANGE, "query %d: fp64 SUM input outside the exact window", i);
    groups[i] = b.h_counts[i];
    if (fl[FLAG_OVERFLOW] || b.h_counts[i] > b.cap[i]) {
      // the estimate was far too low: this query alone, with a larger table
      const uint64_t est = std::max<uint64_t>(b.h_counts[i], b.hq[i].g.S) * 4;
      Plan p = autotune(ctx, b.spec[i], b.rows[i], est, nullptr);
      PlainOut po;
      uint64_t G = 0;
      tga_status st = run_plain(ctx, b.spec[i], b.keys[i], b.rows[i], p, &po, nullptr, nullptr, 0, &G);
      if (st) return st;
      // The device query record (and a captured graph replaying it) still
      // targets the batch's own output buffers: they stay owned by the batch
      // until it is destroyed, and the rerun's result is kept beside them.
      if (i < (int32_t)b.graph_out.size() && b.graph_out[i]) {
        b.bufs.push_back(std::move(b.okeys[i]));
        for (int a = 0; a < b.spec[i].schema.nacc; ++a) b.bufs.push_back(std::move(b.ocols[i][a]));
        b.graph_out[i] = 0;
      }
      b.okeys[i] = std::move(po.keys);
      for (int a = 0; a < b.spec[i].schema.nacc; ++a) b.ocols[i][a] = std::move(po.cols[a]);
      b.cap[i] = po.capacity;
      groups[i] = G;
    }
  }
  return TGA_OK;
}

}  // namespace tga

using namespace tga;

struct tga_batch {
  tga_context *ctx = nullptr;
  int mode = 0;
  BatchExec ex;
  std::vector<uint64_t> groups;
  std::vector<Plan> plan;  // per-query mode
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

#define TGA_API extern "C" __attribute__((visibility("default")))

TGA_API tga_status tga_batch_create(tga_context *ctx, const tga_query *queries, int32_t n, int32_t mode,
                                    tga_batch **out) {
  if (!ctx || !out || n < 0 || (n > 0 && !queries) || (mode != TGA_BATCH_PERSISTENT && mode != TGA_BATCH_PER_QUERY))
    return TGA_ERR_INVALID_ARG;
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  auto *b = new tga_batch();
  b->ctx = ctx;
  b->mode = mode;
  b->groups.assign(n, 0);
  tga_status st = batch_prepare(ctx, queries, n, b->ex);
  if (!st && mode == TGA_BATCH_PERSISTENT) {
    st = batch_build(b->ex, queries);
    if (!st) {
      // one graph: reset kernel + persistent kernel + status copies
      hipError_t be = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
      if (be != hipSuccess) st = fail(ctx, TGA_ERR_HIP, "begin capture: %s", hipGetErrorString(be));
      else {
        tga_status est = batch_enqueue(b->ex, ctx->stream);
        hipGraph_t g = nullptr;
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (est || ce != hipSuccess) {
          if (g) hipGraphDestroy(g);
          st = est ? est : fail(ctx, TGA_ERR_HIP, "end capture: %s", hipGetErrorString(ce));
        } else {
          b->graph = g;
          hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
          if (ie != hipSuccess) st = fail(ctx, TGA_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
        }
      }
    }
  } else if (!st) {
    b->plan.resize(n);
    for (int32_t i = 0; i < n; ++i) b->plan[i] = autotune(ctx, b->ex.spec[i], b->ex.rows[i], b->ex.est[i], nullptr);
    b->ex.okeys.resize(n);
    b->ex.ocols.resize(n);
    b->ex.cap.assign(n, 0);
  }
  if (st) {
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    delete b;
    return st;
  }
  *out = b;
  return TGA_OK;
}

TGA_API tga_status tga_batch_run(tga_batch *b) {
  if (!b) return TGA_ERR_INVALID_ARG;
  tga_context *ctx = b->ctx;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  BatchExec &ex = b->ex;
  if (b->mode == TGA_BATCH_PER_QUERY) {
    for (int32_t i = 0; i < ex.nq; ++i) {
      PlainOut po;
      uint64_t G = 0;
      tga_status st = run_plain(ctx, ex.spec[i], ex.keys[i], ex.rows[i], b->plan[i], &po, nullptr, nullptr, 0, &G);
      if (st) return st;
      ex.okeys[i] = std::move(po.keys);
      for (int a = 0; a < ex.spec[i].schema.nacc; ++a) ex.ocols[i][a] = std::move(po.cols[a]);
      ex.cap[i] = po.capacity;
      b->groups[i] = G;
    }
    return TGA_OK;
  }
  TGA_HIP(ctx, timed(ctx, "batch_persistent", [&] { return hipGraphLaunch(b->exec, ctx->stream); }));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->prof.on) ctx->prof.resolve();
  return batch_finalize(ex, b->groups);
}

TGA_API tga_status tga_batch_num_groups(const tga_batch *b, int32_t q, int64_t *n) {
  if (!b || !n || q < 0 || q >= b->ex.nq) return TGA_ERR_INVALID_ARG;
  *n = (int64_t)b->groups[q];
  return TGA_OK;
}

TGA_API tga_status tga_batch_copy_to_host(const tga_batch *b, int32_t q, int32_t index, void *dst, int64_t cap) {
  if (!b || !dst || q < 0 || q >= b->ex.nq || cap < 0) return TGA_ERR_INVALID_ARG;
  const QuerySpec &s = b->ex.spec[q];
  if (index < -1 || index >= s.schema.nacc) return TGA_ERR_INVALID_ARG;
  tga_context *ctx = b->ctx;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  const uint64_t rows = std::min<uint64_t>((uint64_t)cap, b->groups[q]);
  if (!rows) return TGA_OK;
  const int dt = index < 0 ? s.key_dtype : s.out_dtype[index];
  const void *src = index < 0 ? b->ex.okeys[q].p : b->ex.ocols[q][index].p;
  TGA_HIP(ctx, hipMemcpyAsync(dst, src, rows * dtype_bytes(dt), hipMemcpyDeviceToHost, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return TGA_OK;
}

TGA_API tga_status tga_batch_profile(const tga_batch *b, double *work_ms, double *wait_ms, int64_t *items,
                                     double *fence_ms, double *phase_ms) {
  if (!b || !work_ms || !wait_ms || !items || !fence_ms) return TGA_ERR_INVALID_ARG;
  if (!b->ex.dprof.p) return TGA_ERR_UNSUPPORTED;
  tga_context *ctx = b->ctx;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  uint64_t h[4 * kItemTypes + kPhases];
  TGA_HIP(ctx, hipMemcpyAsync(h, b->ex.dprof.p, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
  TGA_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int t = 0; t < kItemTypes; ++t) {
    work_ms[t] = (double)h[t] * 1e-5;  // 100 MHz ticks
    wait_ms[t] = (double)h[kItemTypes + t] * 1e-5;
    items[t] = (int64_t)h[2 * kItemTypes + t];
    fence_ms[t] = (double)h[3 * kItemTypes + t] * 1e-5;
  }
  if (phase_ms)
    for (int k = 0; k < kPhases; ++k) phase_ms[k] = (double)h[4 * kItemTypes + k] * 1e-5;
  return TGA_OK;
}

TGA_API tga_status tga_batch_destroy(tga_batch *b) {
  if (!b) return TGA_ERR_INVALID_ARG;
  tga_context *ctx = b->ctx;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  delete b;
  return TGA_OK;
}
