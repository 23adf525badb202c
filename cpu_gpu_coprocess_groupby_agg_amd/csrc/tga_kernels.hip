// tga_kernels.hip -- gfx950 kernels of the plain group-by operator.
//
//   (a) k_agg           LDS open-addressing pre-aggregation (tga_agg.hip);
//                       here its work plans
//   (b) k_part_scatter, k_part_scatter_wc
//                       count-free partitioning: LDS counting sort of a tile,
//                       then each partition's run appended to the workgroup's
//                       open 256-row bucket of that partition (whole cache
//                       lines through LDS, one or two passes)
//   (c) k_bucket_count, k_bucket_plan, k_bucket_list
//                       bucket headers -> per-partition bucket lists and the
//                       work-item plan of k_agg
//   (d) k_compact       one-pass compaction of the global table (own look-back)
//       k_scan_u32      decoupled-look-back exclusive scan (the scan entry point)
//
// Reference counterparts (T/ = tiled_groupby_agg_partition/):
//   K1-K5 pre-aggregation      T/src/kernel.cu:9-186     -> k_agg
//   K6/K7 + CUB scan            T/src/kernel.cu:189-291, T/src/util.cu:68-88
//                                                        -> k_part_bucket + (c)
//   K8/K9 intra-partition agg   T/src/kernel.cu:294-378  -> k_agg over partitions
//   S5 + K10 compaction         T/src/kernel.cu:380-406  -> k_compact
#include <map>
#include <mutex>

#include "tga_device.hpp"
#include "tga_internal.hpp"

namespace tga {

// ---------------------------------------------------------------------------
// fills
// ---------------------------------------------------------------------------

__global__ void k_fill_u64(uint64_t *p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// 16-byte streaming copy, the bandwidth reference for the roofline and the
// counters: each lane has four 16-byte loads in flight before it stores, and
// loads and stores are non-temporal (the data is touched once).  A sweep of
// this loop over unroll (1, 4, 8), workgroup size (256..1024), workgroups per
// CU (2..32 wave-equivalents) and cache hints put one 1024-thread workgroup
// per CU with both hints at the top (about 5.7 TB/s for 2 x 2 GiB, against
// 4.5-4.9 TB/s for the plain one-load-per-iteration loop).
__global__ void __launch_bounds__(1024) k_copy16(const uint4 *__restrict__ a0, uint4 *__restrict__ b0, uint64_t n) {
  typedef unsigned int v4u __attribute__((ext_vector_type(4)));
  constexpr int U = 4;
  const v4u *a = (const v4u *)a0;
  v4u *b = (v4u *)b0;
  const uint64_t step = (uint64_t)gridDim.x * blockDim.x * U;
  for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x * U + threadIdx.x; base < n; base += step) {
    v4u v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = base + (uint64_t)u * blockDim.x;
      if (i < n) v[u] = __builtin_nontemporal_load(&a[i]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t i = base + (uint64_t)u * blockDim.x;
      if (i < n) __builtin_nontemporal_store(v[u], &b[i]);
    }
  }
}

template <typename K>
__global__ void k_gtable_init(GTable g, Schema s, WordInit wi) {
  if (g.S_dev) g.S = *g.S_dev;
  const uint64_t n = g.S + 1;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  K *keys = (K *)g.keys;
  for (uint64_t i = t0; i < n; i += stride) keys[i] = KeyTraits<K>::kEmpty;
  #pragma unroll
  for (int a = 0; a < kMaxAccs; ++a) {
    if (a >= s.nacc) break;
    const int kind = s.acc[a].kind;
    if (kind == ACC_SUM_F) {
      uint64_t *p = (uint64_t *)g.acc[a];
      for (uint64_t i = t0; i < n * kLimbs; i += stride) p[i] = 0;
    } else if (acc_global_bytes(kind) == 4) {
      uint32_t *p = (uint32_t *)g.acc[a];
      const uint32_t id = (uint32_t)acc_identity(kind);
      for (uint64_t i = t0; i < n; i += stride) p[i] = id;
    } else {
      uint64_t *p = (uint64_t *)g.acc[a];
      const uint64_t id = acc_identity(kind);
      for (uint64_t i = t0; i < n; i += stride) p[i] = id;
    }
  }
  if (t0 < kFlagWords) g.flags[t0] = 0;
  // the caller's counters, set in the same launch
  for (int k = 0; k < 2; ++k)
    if (wi.p[k] && t0 < wi.n[k]) wi.p[k][t0] = wi.v[k];
}

// ---------------------------------------------------------------------------
// (a)/(c) work plans of the LDS aggregation (the kernel is in tga_agg.hip)
// ---------------------------------------------------------------------------

// Single-segment plan over rows [0, n) in nchunks slices.
__global__ void k_plan_direct(AggPlan plan, uint64_t n, uint32_t nchunks) {
  if (threadIdx.x == 0) {
    plan.pstart[0] = 0;
    plan.pstart[1] = n;
    plan.cstart[0] = 0;
    plan.cstart[1] = nchunks;
    *plan.ticket = 0;
  }
}

// Rows and count of the partitions cut into several work items, and the
// global-table size they imply (AggPlan::gslots).  One thread.
__device__ __forceinline__ void plan_split_totals(const AggPlan &plan, uint64_t srows, uint32_t nsplit) {
  if (plan.split_rows) *plan.split_rows = srows;
  if (!plan.gslots) return;
  const uint64_t want = min(2 * srows, (uint64_t)plan.gs_per_split * nsplit);
  uint64_t S = 4096;
  while (S < want && S < plan.gs_cap) S <<= 1;
  *plan.gslots = min(S, plan.gs_cap);
}

// (c) Partition-contiguous layout: partition p has rows [pbase[p], pbase[p+1]); it becomes
// ceil(rows/chmax) work items.
__global__ void __launch_bounds__(1024) k_plan_chunks(const uint32_t *pbase, uint32_t P, uint64_t n, uint64_t chmax,
                                                      AggPlan plan) {
  __shared__ uint32_t scratch[17];
  __shared__ unsigned long long s_srows;
  __shared__ uint32_t s_nsplit;
  if (threadIdx.x == 0) { s_srows = 0; s_nsplit = 0; }
  __syncthreads();
  extern __shared__ uint32_t cnt[];  // P + 1
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    const uint64_t b = pbase[p];
    const uint64_t e = p + 1 < P ? pbase[p + 1] : n;
    plan.pstart[p] = b;
    const uint64_t len = e - b;
    cnt[p] = (uint32_t)((len + chmax - 1) / chmax);
    if (cnt[p] > 1) {
      atomicAdd(&s_srows, (unsigned long long)len);
      atomicAdd(&s_nsplit, 1u);
    }
  }
  if (threadIdx.x == 0) { plan.pstart[P] = n; *plan.ticket = 0; }
  __syncthreads();
  const uint32_t total = block_exclusive_scan_lds(cnt, P, scratch);
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) plan.cstart[p] = cnt[p];
  if (threadIdx.x == 0) {
    plan.cstart[P] = total;
    plan_split_totals(plan, s_srows, s_nsplit);
  }
}

// ---------------------------------------------------------------------------
// (b) partitioning
// ---------------------------------------------------------------------------

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup w runs on
// XCD w % 8, each XCD with its own L2).  This renumbering gives the
// workgroups of one XCD a contiguous range of items, so neighbouring items
// -- whose outputs share cache lines -- meet in one L2.
constexpr uint32_t kXcds = 8;
// grid for `count` items: a multiple of the XCD count
__host__ __device__ inline uint32_t xcd_grid(uint32_t count) { return (count + kXcds - 1) / kXcds * kXcds; }
// item of workgroup w: XCD x = w % 8 owns items [x*per, (x+1)*per), its k-th
// workgroup (w / 8) takes the k-th; padding workgroups get an item >= count
__device__ __forceinline__ uint32_t xcd_local_index(uint32_t w, uint32_t count) {
  const uint32_t per = (count + kXcds - 1) / kXcds;
  return (w % kXcds) * per + w / kXcds;
}

// Per-tile partition histogram, tile-major: cnt[tile][p] (one contiguous
// row of P counters per tile).
template <typename K>
__global__ void __launch_bounds__(1024) k_part_count(const K *__restrict__ keys, uint64_t n, int pbits,
                                                     uint32_t tile_rows, uint32_t ntiles, uint32_t *cnt) {
  extern __shared__ uint32_t hist[];
  const uint32_t P = 1u << pbits;
  const uint32_t tile = xcd_local_index(blockIdx.x, ntiles);
  if (tile >= ntiles) return;  // uniform per workgroup
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) hist[p] = 0;
  __syncthreads();
  const uint64_t b = (uint64_t)tile * tile_rows;
  const uint64_t e = min(n, b + tile_rows);
#pragma unroll 8
  for (uint64_t r = b + threadIdx.x; r < e; r += blockDim.x)
    atomicAdd(&hist[key_partition((uint64_t)keys[r], pbits)], 1u);
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) cnt[(uint64_t)tile * P + p] = hist[p];
}

// Column-wise exclusive scan of the tile-major histogram: for every
// partition p, offs[t][p] = sum of cnt[t'][p] over tiles t' < t, and
// total[p] = the partition's row count.  Workgroups take chunks of kColTiles
// tiles in ticket order; every column does its own decoupled look-back
// (Merrill & Garland 2016) over the chunks before it, one lane per column,
// so the whole scan reads and writes rows of P contiguous counters (a
// partition-major layout would make every tile touch P scattered lines).
constexpr uint32_t kColTiles = 32;
__global__ void __launch_bounds__(256) k_colscan(const uint32_t *__restrict__ cnt, uint32_t ntiles, uint32_t P,
                                                 uint64_t *desc, uint32_t *ticket, uint32_t *offs, uint32_t *total) {
  __shared__ uint32_t s_chunk;
  if (threadIdx.x == 0) s_chunk = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t c = wave_uniform(s_chunk);
  const uint32_t nchunks = (ntiles + kColTiles - 1) / kColTiles;
  if (c >= nchunks) return;
  const uint32_t t0 = c * kColTiles, t1 = min(ntiles, t0 + kColTiles);
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    uint64_t sum = 0;
    for (uint32_t t = t0; t < t1; ++t) sum += cnt[(uint64_t)t * P + p];
    uint64_t excl = 0;
    if (c == 0) {
      lb_store(&desc[p], kLbInc | sum);
    } else {
      lb_store(&desc[(uint64_t)c * P + p], kLbAgg | sum);
      for (int64_t k = (int64_t)c - 1; k >= 0; --k) {  // every chunk k < c holds an earlier ticket
        uint64_t d = lb_load(&desc[(uint64_t)k * P + p]);
        while ((d >> 62) == 0) {
          __builtin_amdgcn_s_sleep(1);
          d = lb_load(&desc[(uint64_t)k * P + p]);
        }
        excl += d & kLbVal;
        if ((d >> 62) == 2) break;  // an inclusive prefix ends the walk
      }
      lb_store(&desc[(uint64_t)c * P + p], kLbInc | (excl + sum));
    }
    uint64_t run = excl;
    for (uint32_t t = t0; t < t1; ++t) {
      offs[(uint64_t)t * P + p] = (uint32_t)run;
      run += cnt[(uint64_t)t * P + p];
    }
    if (c == nchunks - 1) total[p] = (uint32_t)run;
  }
}

// pbase[p] = rows of the partitions before p (pbase[P] = all rows): the
// start of partition p in the partitioned copy.  One workgroup.
__global__ void __launch_bounds__(1024) k_colbase(const uint32_t *total, uint32_t P, uint32_t *pbase) {
  extern __shared__ uint32_t cnt[];  // P
  __shared__ uint32_t scratch[17];
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) cnt[p] = total[p];
  __syncthreads();
  const uint32_t all = block_exclusive_scan_lds(cnt, P, scratch);
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) pbase[p] = cnt[p];
  if (threadIdx.x == 0) pbase[P] = all;
}

// Exclusive scan of n u32 (total < 2^32), 4096 values per tile, tiles in
// ticket order, offsets from the decoupled look-back.
constexpr int kScanThreads = 256, kScanItems = 16;
__global__ void __launch_bounds__(kScanThreads) k_scan_u32(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                           uint64_t n, uint64_t *desc, uint32_t *ticket) {
  __shared__ uint32_t s_tile, s_excl;
  __shared__ uint32_t scratch[kScanThreads / kWave + 1];
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t base = tile * (kScanThreads * kScanItems) + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t local = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0u;
    local += v[i];
  }
  // block exclusive scan of the thread sums
  const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
  uint32_t x = local;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) scratch[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int w = 0; w < kScanThreads / kWave; ++w) { const uint32_t t = scratch[w]; scratch[w] = run; run += t; }
    scratch[kScanThreads / kWave] = run;
  }
  __syncthreads();
  if (wave == 0) {
    const uint64_t ex = lookback_exclusive(desc, tile, scratch[kScanThreads / kWave]);
    if (lane == 0) s_excl = (uint32_t)ex;
  }
  __syncthreads();
  uint32_t run = s_excl + scratch[wave] + x - local;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
}

// Count-free partitioning into buckets (replaces the reference's histogram
// K6, CUB scan and scatter K7 with one pass over the rows).
//
// The rows are cut into tiles of blockDim*ITEMS rows and a persistent
// workgroup takes one contiguous range of tiles.  For each tile it ranks
// every row inside its partition with an LDS histogram atomic and
// counting-sorts the tile into LDS, which serves as the write-combining
// buffer.  Each partition's run is then appended to that partition's open
// bucket of this workgroup: a block of kBucketRows rows of the partitioned
// copy.  A workgroup keeps one open bucket per partition, so the run of a
// partition from its next tile continues the same bucket and completes the
// cache lines this tile left partial, in the same XCD's L2; buckets are
// line-aligned, so no line is shared between workgroups.  No histogram pass
// has to run first: a partition's rows need not be contiguous, only grouped
// in buckets, which k_bucket_count / k_bucket_plan / k_bucket_list turn into
// one bucket list per partition for k_agg.
//
// Bucket ids come from an atomic pool counter.  A run that needs one fresh
// bucket (the common case: runs are much shorter than a bucket) takes it
// from its wave's reserve, refilled kReserve ids at a time with one atomic,
// so most tiles need no global atomic at all; a run needing several takes
// them, contiguous, with its own atomic.  Ids left in the reserves at the end
// are marked unused (bpart = ~0).
//
// Bucket headers: bpart[id] = partition (~0: unused); bcnt[id] = rows
// (kBucketRows for every bucket except the last one of each (workgroup,
// partition)).  At most ceil(n / kBucketRows) + gridDim.x * (P + blockDim.x)
// ids are taken (full buckets, one open bucket per workgroup and partition,
// and a reserve of < kReserve per wave); the host sizes `cap` to that bound,
// and no id >= cap is ever written.
//
// Per tile the workgroup goes through two barrier intervals:
//   [scan the tile's counts, extend the bucket chains, stage the rows]
//   [write the staged rows out, then rank the next tile's rows]
// so the stores of one tile overlap the hashing and LDS atomics of the next
// (the counts go to the other of two histograms).  The barriers order LDS
// only (lds_barrier), so loads stay in flight across them.  With one value
// column the rows of the tile after the next are loaded as soon as the next
// one is ranked, into the register set the current tile was staged from:
// two register sets alternate, in a loop unrolled twice, so no register
// move (and no wait for a load in flight) is ever needed between them.
// With more columns (whose second register set would spill) the next tile is
// loaded after the staging, during the writes.
// VW: width of every value cell when NC == 1 (4 or 8), 0 = per column.
constexpr uint32_t kReserve = 64;

__host__ __device__ inline uint32_t bucket_of(uint32_t k, uint32_t first, uint32_t delta) {
  return k == 0 ? first : delta + k;
}

// LDS and register state of one k_part_scatter workgroup.
template <typename K, int NC>
struct ScatterState {
  const K *keys;
  const Schema *s;
  uint64_t n;
  uint32_t P, TR, cap;
  int pbits, ncol;
  uint4 *wdesc;     // {q0, first, delta, f | c << 16}: row j of p's run -> position j + q0 of
                    // its chain; f = rows the open bucket held before the run, c = the
                    // run's rows (bucket layout)
  uint32_t *hist;   // the tile's counts, then exclusive offsets
  uint32_t *hnext;  // the next tile's counts
  uint32_t *fill;   // rows in the open bucket (0: none open)
  uint32_t *cur;    // the open bucket
  K *skey;
  char *scol;
  uint16_t *spid;
  uint32_t colb[NC];
  size_t coff[NC];
  uint32_t *scratch;
  uint32_t *bpart, *bcnt, *pool;  // bucket layout
  const uint32_t *offs, *pbase;     // partition-contiguous layout
  K *okeys;
  PartOut o;
  uint32_t rlo, rhi;  // this wave's reserve of bucket ids (wave-uniform)
  uint32_t pid_base;  // bucket headers record partition pid_base + p
};

// Rows of tile t into registers (rows past the end re-read the last row).
template <typename K, int ITEMS, int NC, typename CellT>
__device__ __forceinline__ void scatter_load(const ScatterState<K, NC> &S, uint32_t t, K (&kk)[ITEMS],
                                             CellT (&cell)[ITEMS][NC]) {
  const uint64_t b0 = (uint64_t)t * S.TR;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint64_t row = b0 + (uint64_t)i * blockDim.x + threadIdx.x;
    const uint64_t r = row < S.n ? row : S.n - 1;
    kk[i] = S.keys[r];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c >= S.ncol) { cell[i][c] = 0; continue; }
      cell[i][c] = S.colb[c] == 4 ? (CellT)((const uint32_t *)S.s->cols[c])[r]
                                  : (CellT)((const uint64_t *)S.s->cols[c])[r];
    }
  }
}

// The registers of one tile, pinned (see pin_reg) once their loads are done.
template <typename K, int ITEMS, int NC, typename CellT>
__device__ __forceinline__ void scatter_pin(K (&kk)[ITEMS], CellT (&cell)[ITEMS][NC]) {
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    pin_reg(kk[i]);
#pragma unroll
    for (int c = 0; c < NC; ++c) pin_reg(cell[i][c]);
  }
}

// Rank of each row of tile t inside its partition, counted into h.
template <typename K, int ITEMS, int NC>
__device__ __forceinline__ void scatter_rank(const ScatterState<K, NC> &S, uint32_t t, const K (&kk)[ITEMS],
                                             uint32_t *h, uint32_t (&pp)[ITEMS], uint32_t (&rk)[ITEMS]) {
  const uint64_t b0 = (uint64_t)t * S.TR;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint64_t row = b0 + (uint64_t)i * blockDim.x + threadIdx.x;
    pp[i] = key_partition((uint64_t)kk[i], S.pbits);
    rk[i] = row < S.n ? atomicAdd(&h[pp[i]], 1u) : 0u;
  }
}

// Write descriptors of tile t's runs (rows in all), and the next histogram
// zeroed.  Partition-contiguous layout (!BK): partition p's run of tile t
// starts at pbase[p] + offs[t][p].  Bucket layout (BK): every partition's
// bucket chain is extended by its run; runs needing one fresh bucket take it
// from the wave's reserve, runs needing several take them, contiguous, with
// their own atomic.
template <bool BK, typename K, int NC>
__device__ __forceinline__ void scatter_extend(ScatterState<K, NC> &S, uint32_t rows, uint32_t t) {
  if constexpr (!BK) {
    for (uint32_t p = threadIdx.x; p < S.P; p += blockDim.x) {
      S.wdesc[p] = make_uint4(S.pbase[p] + S.offs[(uint64_t)t * S.P + p] - S.hist[p], 0u, 0u, 0u);
      S.hnext[p] = 0;
    }
    return;
  }
  const int lane = (int)__lane_id();
  for (uint32_t p0 = 0; p0 < S.P; p0 += blockDim.x) {
    const uint32_t p = p0 + threadIdx.x;
    const bool on = p < S.P;
    uint32_t off = 0, c = 0, f = 0, first = 0, fresh = 0;
    if (on) {
      off = S.hist[p];
      c = (p + 1 < S.P ? S.hist[p + 1] : rows) - off;
      f = S.fill[p];
      first = S.cur[p];
      if (c) fresh = ((f + c + kBucketRows - 1) >> kBucketShift) - (f ? 1u : 0u);
    }
    const bool one = fresh == 1;
    const uint64_t m1 = __ballot(one);
    const uint32_t n1 = (uint32_t)__builtin_popcountll(m1);
    const uint32_t j = (uint32_t)__builtin_popcountll(m1 & lanemask_lt());
    uint32_t base = 0;
    if (n1) {
      const uint32_t have = S.rhi - S.rlo;
      uint32_t got = 0;
      if (have < n1) {  // refill (wave-uniform branch)
        if (lane == 0) got = atomicAdd(S.pool, kReserve);
        got = wave_uniform((uint32_t)__shfl((int)got, 0, kWave));
      }
      if (one) base = j < have ? S.rlo + j : got + (j - have);
      if (have < n1) {
        S.rlo = got + (n1 - have);
        S.rhi = got + kReserve;
      } else {
        S.rlo += n1;
      }
    }
    if (fresh > 1) base = atomicAdd(S.pool, fresh);
    uint32_t delta = 0;
    if (on && c) {
      for (uint32_t k = 0; k < fresh; ++k)
        if (base + k < S.cap) S.bpart[base + k] = S.pid_base + p;
      delta = f ? base - 1 : base;
      if (!f) first = base;
      const uint32_t end = f + c;
      const uint32_t full = end >> kBucketShift;  // buckets this run completes
      for (uint32_t k = 0; k < full; ++k) {
        const uint32_t id = bucket_of(k, first, delta);
        if (id < S.cap) S.bcnt[id] = kBucketRows;
      }
      S.fill[p] = end & (kBucketRows - 1);
      S.cur[p] = bucket_of((end - 1) >> kBucketShift, first, delta);
    }
    if (on) {
      S.wdesc[p] = make_uint4(f - off, first, delta, f | c << 16);
      S.hnext[p] = 0;
    }
  }
}

// Counting sort of tile t's rows into the LDS stage.
template <typename K, int ITEMS, int NC, typename CellT>
__device__ __forceinline__ void scatter_stage(const ScatterState<K, NC> &S, uint32_t t, const K (&kk)[ITEMS],
                                              const CellT (&cell)[ITEMS][NC], const uint32_t (&pp)[ITEMS],
                                              const uint32_t (&rk)[ITEMS]) {
  const uint64_t b0 = (uint64_t)t * S.TR;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint64_t row = b0 + (uint64_t)i * blockDim.x + threadIdx.x;
    if (row < S.n) {
      const uint32_t lp = S.hist[pp[i]] + rk[i];
      S.skey[lp] = kk[i];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (c >= S.ncol) break;
        if (S.colb[c] == 4) ((uint32_t *)(S.scol + S.coff[c]))[lp] = (uint32_t)cell[i][c];
        else ((uint64_t *)(S.scol + S.coff[c]))[lp] = (uint64_t)cell[i][c];
      }
      S.spid[lp] = (uint16_t)pp[i];
    }
  }
}

// The staged rows to their destinations: consecutive rows of one run land in
// consecutive rows of the partition (of a bucket, with BK).  REC: the copy
// is (key, 4-byte value) records, one store per row -- one write stream
// instead of two, so half the partially written cache lines at run ends.
template <bool BK, bool REC, typename K, int NC>
__device__ __forceinline__ void scatter_write(const ScatterState<K, NC> &S, uint32_t rows) {
  for (uint32_t jr = threadIdx.x; jr < rows; jr += blockDim.x) {
    const uint4 d = S.wdesc[S.spid[jr]];
    const uint32_t q = jr + d.x;
    uint64_t dst = q;
    if constexpr (BK) {
      const uint32_t id = bucket_of(q >> kBucketShift, d.y, d.z);
      if (id >= S.cap) continue;
      dst = ((uint64_t)id << kBucketShift) | (q & (kBucketRows - 1));
    }
    if constexpr (REC) {
      Rec4<K>::store(S.okeys, dst, S.skey[jr], ((const uint32_t *)S.scol)[jr]);
      continue;
    }
    S.okeys[dst] = S.skey[jr];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c >= S.ncol) break;
      if (S.colb[c] == 4) ((uint32_t *)S.o.cols[c])[dst] = ((const uint32_t *)(S.scol + S.coff[c]))[jr];
      else ((uint64_t *)S.o.cols[c])[dst] = ((const uint64_t *)(S.scol + S.coff[c]))[jr];
    }
  }
}

template <typename K, int ITEMS, int NC, int VW, bool BK, bool REC>
__global__ void __launch_bounds__(1024) k_part_scatter(const K *__restrict__ keys, Schema s, uint64_t n, int pbits,
                                                       uint32_t ntiles, K *okeys, PartOut o, uint32_t *bpart,
                                                       uint32_t *bcnt, uint32_t *pool, uint32_t cap,
                                                       const uint32_t *offs, const uint32_t *pbase) {
  using CellT = typename std::conditional<VW == 4, uint32_t, uint64_t>::type;
  constexpr bool kEarly = NC == 1;
  static_assert(!REC || (NC == 1 && VW == 4), "records carry one 4-byte value");
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t scratch[17];
  ScatterState<K, NC> S;
  S.keys = keys;
  S.s = &s;
  S.n = n;
  S.P = 1u << pbits;
  S.TR = blockDim.x * ITEMS;
  S.cap = cap;
  S.pbits = pbits;
  S.ncol = NC == 1 ? 1 : s.ncols;
  S.wdesc = (uint4 *)smem;
  S.hist = (uint32_t *)(S.wdesc + S.P);
  S.hnext = S.hist + S.P;
  S.fill = S.hnext + S.P;
  S.cur = S.fill + S.P;
  char *stage = (char *)(S.cur + S.P);
  // stage layout: K keys[TR] | per column [TR] cells | u16 pid[TR]
  S.skey = (K *)stage;
  S.scol = stage + (size_t)S.TR * sizeof(K);
  size_t acc_off = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    S.colb[c] = VW ? (uint32_t)VW : c < S.ncol ? (uint32_t)dtype_bytes(s.col_dtype[c]) : 0u;
    S.coff[c] = acc_off;
    acc_off += (size_t)S.TR * S.colb[c];
  }
  S.spid = (uint16_t *)(S.scol + acc_off);
  S.scratch = scratch;
  S.bpart = bpart;
  S.bcnt = bcnt;
  S.pool = pool;
  S.offs = offs;
  S.pbase = pbase;
  S.okeys = okeys;
  S.o = o;
  S.rlo = S.rhi = 0;
  S.pid_base = 0;

  // Tiles of this workgroup: tile, tile + step, ... below tend.  Buckets: a
  // chunk of consecutive tiles (its runs of one partition append to one
  // bucket chain), chunks renumbered so the chunks of one XCD are neighbours.
  // Partition-contiguous: XCD x (workgroups are dispatched round-robin over
  // the XCDs; the grid is a multiple of their count) owns the tiles
  // [x * ntiles / 8, (x + 1) * ntiles / 8), and its m workgroups take them in
  // turn, so the tiles in flight on one XCD are adjacent and so are their
  // runs of each partition: the partial lines where one run ends and the
  // next begins are completed in that XCD's L2, and HBM sees each
  // partition's rows of m tiles as one region.
  uint32_t tile, tend, step;
  if constexpr (BK) {
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    tile = xcd_local_index(blockIdx.x, gridDim.x) * per;
    tend = min(ntiles, tile + per);
    step = 1;
  } else {
    const uint32_t x = blockIdx.x % kXcds, m = gridDim.x / kXcds;
    tile = (uint32_t)((uint64_t)ntiles * x / kXcds) + blockIdx.x / kXcds;
    tend = (uint32_t)((uint64_t)ntiles * (x + 1) / kXcds);
    step = m;
  }
  if (tile >= tend) return;  // uniform per workgroup
  for (uint32_t p = threadIdx.x; p < S.P; p += blockDim.x) {
    S.fill[p] = 0;
    S.cur[p] = 0;
    S.hist[p] = 0;
  }
  K ka[ITEMS], kb[ITEMS];
  CellT ca[ITEMS][NC], cb[ITEMS][NC];
  uint32_t pp[ITEMS], rk[ITEMS];
  scatter_load<K, ITEMS, NC, CellT>(S, tile, ka, ca);
  lds_barrier();
  scatter_rank<K, ITEMS, NC>(S, tile, ka, S.hist, pp, rk);
  // all of the first tile's loads done before the next tile's go out, so no
  // load is outstanding on set A where the loop starts (as after each tile)
  __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
  scatter_pin<K, ITEMS, NC, CellT>(ka, ca);
  if constexpr (kEarly) scatter_load<K, ITEMS, NC, CellT>(S, tile + step, kb, cb);  // clamped past the end

  // One tile: X holds its ranked rows; with kEarly, Y is receiving the next
  // tile's rows, and X is refilled with the tile after that.
#define TGA_SCATTER_TILE(X, XC, Y, YC)                                                           \
  {                                                                                              \
    lds_barrier(); /* the tile's counts are complete, the previous writes done */               \
    const uint32_t rows = block_exclusive_scan_lds<LdsBarrier>(S.hist, S.P, scratch);           \
    scatter_extend<BK, K, NC>(S, rows, tile);                                                              \
    scatter_stage<K, ITEMS, NC, CellT>(S, tile, X, XC, pp, rk);                                  \
    /* vmcnt counts loads and stores in order: Y's loads are waited for     */                   \
    /* here, before this tile's stores, not behind them                      */                  \
    if constexpr (kEarly) {                                                                      \
      __builtin_amdgcn_s_waitcnt(0x0070); /* vmcnt(0) lgkmcnt(0) */                              \
      scatter_pin<K, ITEMS, NC, CellT>(Y, YC);                                                   \
    }                                                                                            \
    lds_barrier(); /* staged; the next histogram is zeroed */                                    \
    const uint32_t next = tile + step;                                                           \
    if constexpr (!kEarly) {                                                                     \
      if (next < tend) scatter_load<K, ITEMS, NC, CellT>(S, next, Y, YC);                        \
    }                                                                                            \
    scatter_write<BK, REC, K, NC>(S, rows);                                                               \
    if (next >= tend) break;                                                                     \
    /* the next tile, ranked into the other histogram while the stores drain */                  \
    scatter_rank<K, ITEMS, NC>(S, next, Y, S.hnext, pp, rk);                                     \
    if constexpr (kEarly) scatter_load<K, ITEMS, NC, CellT>(S, next + step, X, XC);              \
    uint32_t *t_ = S.hist;                                                                       \
    S.hist = S.hnext;                                                                            \
    S.hnext = t_;                                                                                \
    tile = next;                                                                                 \
  }
  if constexpr (kEarly) {
    for (;;) {
      TGA_SCATTER_TILE(ka, ca, kb, cb)
      TGA_SCATTER_TILE(kb, cb, ka, ca)
    }
  } else {
    for (;;) TGA_SCATTER_TILE(ka, ca, ka, ca)
  }
#undef TGA_SCATTER_TILE
  if constexpr (BK) {
    // the open buckets are the last of their chains
    for (uint32_t p = threadIdx.x; p < S.P; p += blockDim.x)
      if (S.fill[p] && S.cur[p] < cap) bcnt[S.cur[p]] = S.fill[p];
    // unused reserve ids
    for (uint32_t id = S.rlo + (uint32_t)__lane_id(); id < S.rhi; id += kWave)
      if (id < cap) bpart[id] = ~0u;
  }
}

// ---------------------------------------------------------------------------
// Write-combining scatter: bucket layout, (key, 4-byte value) records.
//
// A partition's run within one tile is a few hundred bytes, so most of its
// cache lines are only partly written by the tile, and the L2 evicts them
// (as masked partial writes) before the next tile completes them.  Here no
// partial line leaves the workgroup until the end: each partition's
// unfinished line, < 32 dwords, is kept in an LDS carry buffer, and a tile
// writes only complete, aligned 128-byte lines.  A bucket (kBucketRows
// records of RD dwords) is a whole number of lines, so a line never straddles
// two buckets.
//
// The partition's record stream is viewed as dwords: the carry (the first cd
// dwords of the line that holds the open bucket's fill point, cd =
// fill * RD mod 32) followed by the tile's run of c records (c * RD dwords);
// the tile writes its first (cd + c * RD) / 32 lines and carries the rest.
// Per tile the workgroup goes through two barrier intervals:
//   [scan counts and task counts, extend the chains, stage the records]
//   [tasks: write the complete lines and move the new partial lines to the
//    carry buffer; rank the next tile]
// then the next tile; the carries left at the end are written as partial
// lines.  Loads, stage and ranking are those of k_part_scatter.
//
// LDS: wdesc[P] uint4 | hist, hnext, fill, cur, loff [P] u32 |
//      carry [P][32] u32 | stage [TR * RD] u32 | lpid [tasks] u16
// ---------------------------------------------------------------------------
constexpr uint32_t kLineDw = 32;  // dwords per 128-byte cache line
constexpr uint32_t kWcMaxP = 512;  // the largest fanout whose carry buffer fits

// tasks of one tile: complete lines plus at most one carry-only task per
// partition
__host__ __device__ inline uint32_t wc_tasks_max(uint32_t P, uint32_t TR, uint32_t RD) {
  return (P * (kLineDw - 1) + TR * RD) / kLineDw + 1 + P;
}

size_t scatter_wc_lds_bytes(int key_bytes, int vd, int pbits, uint32_t tile_rows) {
  const uint32_t P = 1u << pbits, RD = (uint32_t)key_bytes / 4 + (uint32_t)vd;
  return (size_t)P * (16 + 5 * 4 + kLineDw * 4) + (size_t)tile_rows * RD * 4 +
         (((size_t)wc_tasks_max(P, tile_rows, RD) * 2 + 15) & ~(size_t)15);
}

// TWO: the second pass of a two-pass partitioning.  The input is then the
// first pass's records, read through its bucket lists in work items of at
// most items_per list entries of one first-pass partition p1 (taken from a
// ticket, so a hot partition is spread over many workgroups).  A row goes to
// sub-partition key_partition(k, hash_bits) mod 2^pbits of p1, recorded in the
// bucket headers as p1 * 2^pbits + that: the top hash_bits bits of the
// partition hash, as one pass would give.  Each work item ends with its
// carries written out and its buckets closed.
struct Pass2In {
  const void *rec;         // first-pass records
  const uint64_t *list;    // first-pass bucket lists, rows << 32 | id, grouped by partition
  const uint64_t *pstart;  // [P1 + 1] list offsets of the first-pass partitions
  const uint32_t *istart;  // [P1 + 1] first work item of each first-pass partition
  uint32_t *ticket;
  uint32_t P1, hash_bits, items_per;
};

// Value words of a record (first pass): word d of row r is vp[d][r << vsh[d]]
// (a 4-byte column is one word at stride 1, an 8-byte one two at stride 2).
struct RecSrc {
  const uint32_t *vp[kMaxCols];
  uint32_t vsh[kMaxCols];
};

template <typename K, int VD, int ITEMS, bool TWO>
__global__ void __launch_bounds__(1024) k_part_scatter_wc(const K *__restrict__ keys, Schema s, uint64_t n,
                                                          int pbits, uint32_t ntiles, void *orec, uint32_t *bpart,
                                                          uint32_t *bcnt, uint32_t *pool, uint32_t cap, RecSrc rs,
                                                          Pass2In p2) {
  using R = RecW<K, VD>;
  constexpr uint32_t KD = (uint32_t)R::KD;                  // key words per record
  constexpr uint32_t RD = (uint32_t)R::RD;                  // words per record
  constexpr uint32_t LB = kBucketRows * RD / kLineDw;       // lines per bucket
  static_assert(kBucketRows * RD % kLineDw == 0, "buckets are whole lines");
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t scratch[17];
  ScatterState<K, 1> S;
  S.keys = keys;
  S.s = &s;
  S.n = n;
  S.P = 1u << pbits;
  S.TR = blockDim.x * ITEMS;
  S.cap = cap;
  S.pbits = pbits;
  S.ncol = 1;
  S.colb[0] = 4;
  S.coff[0] = 0;
  S.wdesc = (uint4 *)smem;
  S.hist = (uint32_t *)(S.wdesc + S.P);
  S.hnext = S.hist + S.P;
  S.fill = S.hnext + S.P;
  S.cur = S.fill + S.P;
  uint32_t *loff = S.cur + S.P;       // first task of each partition in the tile's task list
  uint32_t *carry = loff + S.P;        // [P][kLineDw]
  uint32_t *stage = carry + S.P * kLineDw;  // the tile's records, grouped by partition
  uint16_t *lpid = (uint16_t *)(stage + S.TR * RD);  // partition of each task of the tile
  S.skey = nullptr;
  S.scol = nullptr;
  S.spid = nullptr;
  S.scratch = scratch;
  S.bpart = bpart;
  S.bcnt = bcnt;
  S.pool = pool;
  S.offs = nullptr;
  S.pbase = nullptr;
  S.okeys = nullptr;
  S.rlo = S.rhi = 0;
  S.pid_base = 0;
  uint32_t *out = (uint32_t *)orec;
  const uint32_t hbits = TWO ? p2.hash_bits : (uint32_t)pbits;
  constexpr uint32_t kBpt = ITEMS * 1024 / kBucketRows;  // first-pass buckets per tile (TWO)
  __shared__ uint32_t s_item;
  // TWO: the list entries of tiles t and t + 1, fetched two tiles ahead (at
  // the top of tile t - 2, by the last wave) so no row load waits on one
  __shared__ uint64_t s_ent[2][kBpt];

  // Rows of tile t into registers (rows past the end re-read an existing
  // one), with the mask of the valid ones.
  uint64_t ie0 = 0, ie1 = 0;  // TWO: the work item's list entries
  auto wc_load = [&](uint32_t t, K(&kk)[ITEMS], uint32_t(&cc)[ITEMS][VD], uint32_t &vm) {
    vm = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      if constexpr (!TWO) {
        const uint64_t row = (uint64_t)t * S.TR + (uint64_t)i * blockDim.x + threadIdx.x;
        const uint64_t r = row < n ? row : n - 1;
        kk[i] = keys[r];
#pragma unroll
        for (int d = 0; d < VD; ++d) cc[i][d] = rs.vp[d][r << rs.vsh[d]];
        vm |= (uint32_t)(row < n) << i;
      } else {
        const uint32_t j = (uint32_t)i * blockDim.x + threadIdx.x;
        const uint64_t e = ie0 + (uint64_t)t * kBpt + (j >> kBucketShift);
        const uint32_t rr = j & (kBucketRows - 1);
        const uint64_t ent = s_ent[t & 1][j >> kBucketShift];
        const bool ok = e < ie1 && rr < (uint32_t)(ent >> 32);
        R::load(p2.rec, ((ent & 0xffffffffULL) << kBucketShift) + (ok ? rr : 0u), kk[i], cc[i]);
        vm |= (uint32_t)ok << i;
      }
    }
  };
  auto ent_fetch = [&](uint32_t t) {
    if constexpr (TWO) {
      const uint32_t b = threadIdx.x - (blockDim.x - kBpt);
      if (threadIdx.x >= blockDim.x - kBpt) {
        const uint64_t e = ie0 + (uint64_t)t * kBpt + b;
        s_ent[t & 1][b] = p2.list[e < ie1 ? e : ie1 - 1];
      }
    }
  };
  // Rank of each row inside its partition's run of the tile.  TWO: a work
  // item of a hot first-pass partition sends most rows of a wave to one
  // sub-partition, whose counter would take one atomic per lane (a 64-way
  // same-address chain); the lanes sharing the first valid lane's
  // sub-partition take their ranks from one atomic instead.
  auto wc_rank = [&](const K(&kk)[ITEMS], uint32_t vm, uint32_t *h, uint32_t(&pp)[ITEMS], uint32_t(&rk)[ITEMS]) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      pp[i] = key_partition((uint64_t)kk[i], (int)hbits) & (S.P - 1);
      const bool v = (vm >> i) & 1u;
      if constexpr (TWO) {
        const uint64_t act = __ballot(v);
        const uint32_t lp = (uint32_t)__shfl((int)pp[i], act ? (int)__builtin_ctzll(act) : 0, kWave);
        const bool same = v && pp[i] == lp;
        const uint64_t sm = __ballot(same);
        const uint64_t below = sm & lanemask_lt();
        uint32_t base = 0;
        if (same && below == 0) base = atomicAdd(&h[lp], (uint32_t)__builtin_popcountll(sm));
        base = (uint32_t)__shfl((int)base, sm ? (int)__builtin_ctzll(sm) : 0, kWave);
        rk[i] = same ? base + (uint32_t)__builtin_popcountll(below) : v ? atomicAdd(&h[pp[i]], 1u) : 0u;
      } else {
        rk[i] = v ? atomicAdd(&h[pp[i]], 1u) : 0u;
      }
    }
  };

  uint32_t tile = 0, tend = 0;
  if constexpr (!TWO) {
    const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
    tile = xcd_local_index(blockIdx.x, gridDim.x) * per;
    tend = min(ntiles, tile + per);
    if (tile >= tend) return;  // uniform per workgroup
  }
  const uint32_t items_total = TWO ? p2.istart[p2.P1] : 1u;
  const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (;;) {  // work items (one: this workgroup's tiles, without TWO)
  if constexpr (TWO) {
    if (threadIdx.x == 0) s_item = atomicAdd(p2.ticket, 1u);
    __syncthreads();
    const uint32_t item = wave_uniform(s_item);
    if (item >= items_total) break;
    uint32_t lo = 0, hi = p2.P1;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (p2.istart[mid] <= item) lo = mid; else hi = mid;
    }
    ie0 = p2.pstart[lo] + (uint64_t)(item - p2.istart[lo]) * p2.items_per;
    ie1 = min(p2.pstart[lo + 1], ie0 + p2.items_per);
    S.pid_base = lo * S.P;
    tile = 0;
    tend = (uint32_t)((ie1 - ie0 + kBpt - 1) / kBpt);
  }
  for (uint32_t p = threadIdx.x; p < S.P; p += blockDim.x) {
    S.fill[p] = 0;
    S.cur[p] = 0;
    S.hist[p] = 0;
  }
  K ka[ITEMS], kb[ITEMS];
  uint32_t ca[ITEMS][VD], cb[ITEMS][VD];
  uint32_t va = 0, vb = 0;
  uint32_t pp[ITEMS], rk[ITEMS];
  if constexpr (TWO) {
    ent_fetch(tile);
    ent_fetch(tile + 1);
    __syncthreads();
  }
  wc_load(tile, ka, ca, va);
  lds_barrier();
  wc_rank(ka, va, S.hist, pp, rk);
  __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
  scatter_pin<K, ITEMS, VD, uint32_t>(ka, ca);
  wc_load(tile + 1, kb, cb, vb);  // clamped past the end

#define TGA_WC_TILE(X, XC, XV, Y, YC, YV)                                                        \
  {                                                                                              \
    lds_barrier(); /* counts complete; the previous tile's tasks done */                        \
    ent_fetch(tile + 2); /* its slot's entries were read by the load of this tile */             \
    /* rows | tasks << 16 per partition, scanned together: a partition with */                  \
    /* rows has max(1, complete lines) tasks; its task 0 also moves the new */                  \
    /* partial line to the carry buffer */                                                      \
    /* (by wave 0 alone: no barrier inside) */                                                  \
    if (wave == 0) {                                                                             \
      const uint32_t t_ = wave_exclusive_scan_fn<kWcMaxP / kWave>(                                \
          S.P,                                                                                   \
          [&](uint32_t p) {                                                                      \
            const uint32_t c = S.hist[p];                                                        \
            const uint32_t nl = (((S.fill[p] * RD) & (kLineDw - 1)) + c * RD) / kLineDw;         \
            return c | (c ? max(nl, 1u) : 0u) << 16;                                             \
          },                                                                                     \
          [&](uint32_t p, uint32_t v) {                                                          \
            S.hist[p] = v & 0xffffu;                                                             \
            loff[p] = v >> 16;                                                                   \
          });                                                                                    \
      if (lane == 0) scratch[0] = t_;                                                            \
    }                                                                                            \
    lds_barrier();                                                                               \
    const uint32_t tot = scratch[0];                                                             \
    const uint32_t rows = tot & 0xffffu, tasks = tot >> 16;                                      \
    scatter_extend<true, K, 1>(S, rows, tile);                                                   \
    for (uint32_t p = threadIdx.x; p < S.P; p += blockDim.x) {                                   \
      const uint32_t l0 = loff[p], l1 = p + 1 < S.P ? loff[p + 1] : tasks;                       \
      for (uint32_t l = l0; l < l1; ++l) lpid[l] = (uint16_t)p;                                  \
    }                                                                                            \
    {                                                                                            \
      _Pragma("unroll") for (int i = 0; i < ITEMS; ++i) {                                        \
        if ((XV >> i) & 1u) {                                                                    \
          uint32_t *r = stage + (S.hist[pp[i]] + rk[i]) * RD;                                    \
          r[0] = (uint32_t)(uint64_t)X[i];                                                       \
          if constexpr (KD == 2) r[1] = (uint32_t)((uint64_t)X[i] >> 32);                        \
          _Pragma("unroll") for (uint32_t d_ = 0; d_ < (uint32_t)VD; ++d_) r[KD + d_] = XC[i][d_];   \
        }                                                                                        \
      }                                                                                          \
    }                                                                                            \
    __builtin_amdgcn_s_waitcnt(0x0070); /* vmcnt(0) lgkmcnt(0): Y's loads */                     \
    scatter_pin<K, ITEMS, VD, uint32_t>(Y, YC);                                                  \
    lds_barrier(); /* staged; task list built */                                                 \
    /* tasks: 8 lanes x 16 bytes per line, 8 tasks per wave step.  Each */                     \
    /* lane of task 0 reads the old carry and writes the new one at the */                      \
    /* same 4 positions, reads first; no other task reads the carry */                          \
    for (uint32_t g0 = wave * 8; g0 < tasks; g0 += nw * 8) {                                     \
      const uint32_t g = g0 + (lane >> 3), part = lane & 7;                                      \
      if (g < tasks) {                                                                           \
        const uint32_t p = lpid[g];                                                              \
        const uint32_t l = g - loff[p];                                                          \
        const uint4 d = S.wdesc[p];                                                              \
        const uint32_t f = d.w & 0xffffu, c = d.w >> 16, off = f - d.x;                          \
        const uint32_t cd = (f * RD) & (kLineDw - 1), all = cd + c * RD;                         \
        const uint32_t nl = all / kLineDw;                                                       \
        if (l < nl) {                                                                            \
          uint32_t v[4];                                                                         \
          _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                        \
            const uint32_t q = l * kLineDw + part * 4 + i;                                       \
            v[i] = q < cd ? carry[p * kLineDw + q] : stage[off * RD + q - cd];                   \
          }                                                                                      \
          const uint32_t lg = (f * RD) / kLineDw + l;                                            \
          const uint32_t id = bucket_of(lg / LB, d.y, d.z);                                      \
          if (id < S.cap)                                                                        \
            *(uint4 *)(out + (uint64_t)id * (kBucketRows * RD) + (lg % LB) * kLineDw + part * 4) = \
                make_uint4(v[0], v[1], v[2], v[3]);                                              \
        }                                                                                        \
        if (l == 0) {                                                                            \
          _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                        \
            const uint32_t j = part * 4 + i, q = nl * kLineDw + j;                               \
            if (q >= cd && q < all) carry[p * kLineDw + j] = stage[off * RD + q - cd];           \
          }                                                                                      \
        }                                                                                        \
      }                                                                                          \
    }                                                                                            \
    const uint32_t next = tile + 1;                                                              \
    if (next >= tend) break;                                                                     \
    wc_rank(Y, YV, S.hnext, pp, rk);                                                             \
    wc_load(next + 1, X, XC, XV);                                                                \
    uint32_t *t_ = S.hist;                                                                       \
    S.hist = S.hnext;                                                                            \
    S.hnext = t_;                                                                                \
    tile = next;                                                                                 \
  }
  for (;;) {
    TGA_WC_TILE(ka, ca, va, kb, cb, vb)
    TGA_WC_TILE(kb, cb, vb, ka, ca, va)
  }
#undef TGA_WC_TILE
  lds_barrier();  // the last carries are in place
  // carries left: partial lines at the fill point of the open buckets
  for (uint32_t e = threadIdx.x; e < S.P * kLineDw; e += blockDim.x) {
    const uint32_t p = e / kLineDw, i = e % kLineDw;
    const uint32_t f = S.fill[p], id = S.cur[p];
    if (i < ((f * RD) & (kLineDw - 1)) && id < cap)
      out[(uint64_t)id * (kBucketRows * RD) + ((f * RD) & ~(kLineDw - 1)) + i] = carry[e];
  }
  for (uint32_t p = threadIdx.x; p < S.P; p += blockDim.x)
    if (S.fill[p] && S.cur[p] < cap) bcnt[S.cur[p]] = S.fill[p];
  if constexpr (!TWO) break;
  __syncthreads();  // the item's state is read; the next item resets it
  }  // work items
  for (uint32_t id = S.rlo + (uint32_t)__lane_id(); id < S.rhi; id += kWave)
    if (id < cap) bpart[id] = ~0u;
}

// Per-partition bucket and row totals from the bucket headers (pbk, prow
// zeroed by the caller).
__global__ void __launch_bounds__(1024) k_bucket_count(const uint32_t *__restrict__ bpart,
                                                       const uint32_t *__restrict__ bcnt, const uint32_t *pool,
                                                       uint32_t cap, uint32_t P, uint32_t *pbk, uint64_t *prow) {
  extern __shared__ __align__(16) char smem[];
  // one workgroup's rows of a partition stay below 2^31 (a partitioning
  // pass covers fewer than 2^31 rows)
  uint32_t *lrow = (uint32_t *)smem;
  uint32_t *lbk = lrow + P;
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    lrow[p] = 0;
    lbk[p] = 0;
  }
  __syncthreads();
  const uint32_t nb = min(*pool, cap);
  const uint32_t per = (nb + gridDim.x - 1) / gridDim.x;
  const uint32_t lo = blockIdx.x * per, hi = min(nb, lo + per);
  for (uint32_t id = lo + threadIdx.x; id < hi; id += blockDim.x) {
    const uint32_t p = bpart[id];
    if (p >= P) continue;
    atomicAdd(&lbk[p], 1u);
    atomicAdd(&lrow[p], bcnt[id]);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    if (!lbk[p]) continue;
    atomicAdd(&pbk[p], lbk[p]);
    atomicAdd((unsigned long long *)&prow[p], (unsigned long long)lrow[p]);
  }
}

// (c) Work plan over the bucket lists: partition p owns list entries
// [pstart[p], pstart[p+1]) and is cut into ceil(rows / chmax) work items;
// bcur[p] = pstart[p] (the fill cursor of k_bucket_list).  One workgroup.
__global__ void __launch_bounds__(1024) k_bucket_plan(const uint32_t *pbk, const uint64_t *prow, uint32_t P,
                                                      uint64_t chmax, uint32_t *bcur, uint32_t *bend, AggPlan plan) {
  extern __shared__ uint32_t cnt[];  // [2P]
  __shared__ uint32_t scratch[17];
  __shared__ unsigned long long s_srows;
  __shared__ uint32_t s_nsplit;
  uint32_t *nbk = cnt, *nit = cnt + P;
  if (threadIdx.x == 0) { s_srows = 0; s_nsplit = 0; }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    const uint64_t rows = prow[p];
    nbk[p] = pbk[p];
    nit[p] = (uint32_t)((rows + chmax - 1) / chmax);
    if (nit[p] > 1) {
      atomicAdd(&s_srows, (unsigned long long)rows);
      atomicAdd(&s_nsplit, 1u);
    }
  }
  __syncthreads();
  const uint32_t buckets = block_exclusive_scan_lds(nbk, P, scratch);
  const uint32_t items = block_exclusive_scan_lds(nit, P, scratch);
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    plan.pstart[p] = nbk[p];
    bcur[p] = nbk[p];
    bend[p] = p + 1 < P ? nbk[p + 1] : buckets;
    plan.cstart[p] = nit[p];
  }
  if (threadIdx.x == 0) {
    plan.pstart[P] = buckets;
    plan.cstart[P] = items;
    *plan.ticket = 0;
    plan_split_totals(plan, s_srows, s_nsplit);
  }
}

// Buckets grouped by partition: blist[pstart[p] ..] = the entries
// rows << 32 | id of the buckets of p, the full ones first (any order), then
// the partly filled ones (one per scatter workgroup that saw the partition).
// Each workgroup counts its range of buckets per partition in LDS (full ones
// in the low, partly filled ones in the high half of a word), reserves its
// share of every list with one atomic from the front (bcur[p]) and one from
// the back (bend[p], which starts at pstart[p+1]), then places its entries,
// the full ones first.
__global__ void __launch_bounds__(1024) k_bucket_list(const uint32_t *__restrict__ bpart,
                                                      const uint32_t *__restrict__ bcnt, const uint32_t *pool,
                                                      uint32_t cap, uint32_t P, uint32_t *bcur, uint32_t *bend,
                                                      uint64_t *blist) {
  extern __shared__ uint32_t cnt[];  // [2P]
  uint32_t *lc = cnt, *lb = cnt + P;
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) lc[p] = 0;
  __syncthreads();
  const uint32_t nb = min(*pool, cap);
  const uint32_t per = (nb + gridDim.x - 1) / gridDim.x;
  const uint32_t lo = blockIdx.x * per, hi = min(nb, lo + per);
  // per < 2^16: a workgroup's counts of one partition fit half a word
  const bool packed = per < (1u << 16);
  for (uint32_t id = lo + threadIdx.x; id < hi; id += blockDim.x) {
    const uint32_t p = bpart[id];
    if (p < P && (bcnt[id] == kBucketRows || packed)) atomicAdd(&lc[p], bcnt[id] == kBucketRows ? 1u : 1u << 16);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    const uint32_t full = packed ? lc[p] & 0xffffu : lc[p];
    lb[p] = full ? atomicAdd(&bcur[p], full) : 0u;
    // the back share: its lowest slot, kept in lc's high half's place
    lc[p] = packed ? lc[p] >> 16 : 0u;
  }
  __syncthreads();
  if (packed) {
    for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
      const uint32_t part = lc[p];
      lc[p] = part ? atomicAdd(&bend[p], 0u - part) - part : 0u;
    }
  }
  __syncthreads();
  // full buckets: lb[p] + rank; partly filled: lc[p] + rank (packed), or a
  // slot of their own from the back; ranks count up from the bases
  for (uint32_t id = lo + threadIdx.x; id < hi; id += blockDim.x) {
    const uint32_t p = bpart[id];
    if (p >= P) continue;
    const uint32_t rows = bcnt[id];
    const uint32_t at = rows == kBucketRows ? atomicAdd(&lb[p], 1u)
                        : packed           ? atomicAdd(&lc[p], 1u)
                                           : atomicAdd(&bend[p], ~0u) - 1u;
    blist[at] = ((uint64_t)rows << 32) | id;
  }
}

// Pairs the partly filled buckets of partition blockIdx.x (list entries
// [bend[p], pstart[p+1])) into blocks of at most kBucketRows rows, so k_agg,
// whose wave block is one bucket, does not spend a whole block on each.
// Sorted by rows a[0] <= ... <= a[m-1] (a counting sort), the largest k for which every pair
// (a[i], a[2k-1-i]), i < k, fits is the most pairs any pairing reaches
// (if the 2k smallest can be paired at all, they can be paired this way, and
// the largest feasible k is found by bisection: feasibility only shrinks as
// k grows).  Written: the k pairs (larger bucket in blist, smaller in
// blist2), then the m - 2k unpaired buckets, then k entries of 0 rows, which
// end the segment.  blist2 of the full buckets is 0, and of all buckets
// when pair is 0.
constexpr uint32_t kPairMax = 1024;
__global__ void __launch_bounds__(256) k_bucket_pair(const uint64_t *__restrict__ pstart,
                                                     const uint32_t *__restrict__ bend, uint64_t *blist,
                                                     uint64_t *blist2, uint32_t pair) {
  __shared__ uint64_t a[kPairMax];
  __shared__ uint32_t hist[kBucketRows];
  __shared__ uint32_t scratch[17];
  __shared__ uint32_t s_bad;
  const uint32_t p = blockIdx.x;
  const uint64_t b = pstart[p], end = pstart[p + 1], lo = bend[p];
  for (uint64_t k = b + threadIdx.x; k < lo; k += blockDim.x) blist2[k] = 0;
  const uint32_t m = (uint32_t)(end - lo);
  if (!pair || m < 2 || m > kPairMax) {
    for (uint64_t k = lo + threadIdx.x; k < end; k += blockDim.x) blist2[k] = 0;
    return;
  }
  // counting sort by rows (1..kBucketRows-1): rank inside the row count's
  // bin from the LDS histogram, bin offsets from its scan
  for (uint32_t r = threadIdx.x; r < kBucketRows; r += blockDim.x) hist[r] = 0;
  __syncthreads();
  uint64_t ent[kPairMax / 256];
  uint32_t rank[kPairMax / 256];
#pragma unroll
  for (uint32_t j = 0; j < kPairMax / 256; ++j) {
    const uint32_t i = threadIdx.x + j * 256;
    ent[j] = i < m ? blist[lo + i] : 0;
    rank[j] = i < m ? atomicAdd(&hist[(uint32_t)(ent[j] >> 32) & (kBucketRows - 1)], 1u) : 0;
  }
  __syncthreads();
  block_exclusive_scan_lds(hist, kBucketRows, scratch);
  __syncthreads();
#pragma unroll
  for (uint32_t j = 0; j < kPairMax / 256; ++j) {
    const uint32_t i = threadIdx.x + j * 256;
    if (i < m) a[hist[(uint32_t)(ent[j] >> 32) & (kBucketRows - 1)] + rank[j]] = ent[j];
  }
  __syncthreads();
  uint32_t klo = 0, khi = m / 2;
  while (klo < khi) {
    const uint32_t k = (klo + khi + 1) / 2;
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < k; i += blockDim.x)
      if ((a[i] >> 32) + (a[2 * k - 1 - i] >> 32) > kBucketRows) s_bad = 1;
    __syncthreads();
    const bool bad = s_bad != 0;
    __syncthreads();
    if (bad) khi = k - 1;
    else klo = k;
  }
  const uint32_t k = klo;
  for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) {
    uint64_t x = 0, y = 0;
    if (i < k) {
      x = a[2 * k - 1 - i];
      y = a[i];
    } else if (i < m - k) {
      x = a[k + i];
    }
    blist[lo + i] = x;
    blist2[lo + i] = y;
  }
}

// ---------------------------------------------------------------------------
// (d) compaction: occupied slots of the global table -> dense output, in slot
// order, one pass with the look-back for the output offsets.
// ---------------------------------------------------------------------------

constexpr int kCompactThreads = 256, kCompactItems = 8;

template <typename K>
__global__ void __launch_bounds__(kCompactThreads) k_compact(GTable g, Schema s, OutDesc o, uint64_t *desc,
                                                             uint32_t *ticket, uint64_t *count_out,
                                                             const uint64_t *base) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_excl;
  __shared__ uint32_t wcnt[kCompactItems][kCompactThreads / kWave];
  __shared__ uint32_t s_total;
  if (g.S_dev) g.S = *g.S_dev;
  const uint64_t n = g.S + 1;
  const uint64_t TS = (uint64_t)kCompactThreads * kCompactItems;
  // workgroups beyond the used tiles leave at once (the others take them all)
  if ((uint64_t)blockIdx.x * TS >= n) return;
  const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
  const K *keys = (const K *)g.keys;
  // Tiles in ticket order until the table's last one: the look-back waits
  // only on lower tickets, all taken by workgroups already running.
  for (;;) {
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint64_t tile = s_tile;
    const uint64_t tile_base = tile * TS;
    if (tile_base >= n) break;
    bool occ[kCompactItems];
    uint32_t wpre[kCompactItems];
#pragma unroll
    for (int i = 0; i < kCompactItems; ++i) {
      const uint64_t slot = tile_base + (uint64_t)i * kCompactThreads + threadIdx.x;
      bool oc = false;
      if (slot < g.S) oc = keys[slot] != KeyTraits<K>::kEmpty;
      else if (slot == g.S) oc = g.flags[FLAG_SENTINEL] != 0;
      occ[i] = oc;
      const uint64_t bal = __ballot(oc);
      wpre[i] = (uint32_t)__builtin_popcountll(bal & lanemask_lt());
      if (lane == 0) wcnt[i][wave] = (uint32_t)__builtin_popcountll(bal);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t run = 0;
      for (int i = 0; i < kCompactItems; ++i)
        for (int w = 0; w < kCompactThreads / kWave; ++w) { const uint32_t c = wcnt[i][w]; wcnt[i][w] = run; run += c; }
      s_total = run;
    }
    __syncthreads();
    if (wave == 0) {
      const uint64_t ex = lookback_exclusive(desc, tile, s_total);
      if (lane == 0) {
        const uint64_t b0 = base ? *base : 0;
        s_excl = b0 + ex;
        if (tile_base + TS >= n) *count_out = b0 + ex + s_total;  // last tile
      }
    }
    __syncthreads();
    const uint64_t excl = s_excl;
#pragma unroll
    for (int i = 0; i < kCompactItems; ++i) {
      if (!occ[i]) continue;
      const uint64_t slot = tile_base + (uint64_t)i * kCompactThreads + threadIdx.x;
      const uint64_t pos = excl + wcnt[i][wave] + wpre[i];
      if (pos >= o.capacity) {  // output too small: the caller grows it or reports the count
        atomicOr(&g.flags[FLAG_OVERFLOW], 1u);
        continue;
      }
      ((K *)o.keys)[pos] = slot < g.S ? keys[slot] : KeyTraits<K>::kEmpty;
#pragma unroll
      for (int a = 0; a < kMaxAccs; ++a) {
        if (a >= s.nacc) break;
        const int kind = s.acc[a].kind;
        switch (kind) {
          case ACC_COUNT: case ACC_SUM_I:
            ((uint64_t *)o.cols[a])[pos] = ((const uint64_t *)g.acc[a])[slot];
            break;
          case ACC_MIN_I32: case ACC_MAX_I32:
            ((uint32_t *)o.cols[a])[pos] = ((const uint32_t *)g.acc[a])[slot];
            break;
          case ACC_SUM_F: {
            int64_t L[kLimbs];
#pragma unroll
            for (int l = 0; l < kLimbs; ++l) L[l] = ((const int64_t *)g.acc[a])[(uint64_t)l * n + slot];
            ((double *)o.cols[a])[pos] = fx_to_double(L);
            break;
          }
          default: {  // ACC_MIN_I64 / ACC_MAX_I64 back to the column dtype
            const int64_t v = ((const int64_t *)g.acc[a])[slot];
            const int dt = s.acc[a].col_dtype;
            if (dt == DT_U64) ((uint64_t *)o.cols[a])[pos] = (uint64_t)v ^ (1ULL << 63);
            else if (dt == DT_F64) ((uint64_t *)o.cols[a])[pos] = ordered_to_f64_bits(v);
            else if (dt == DT_U32 || dt == DT_I32) ((uint32_t *)o.cols[a])[pos] = (uint32_t)v;
            else ((int64_t *)o.cols[a])[pos] = v;
            break;
          }
        }
      }
    }
    __syncthreads();  // s_tile, wcnt and s_excl are rewritten by the next tile
  }
}

// Groups in a global table (occupied slots plus the empty-key slot), added
// to *out: one atomic per wave.
template <typename K>
__global__ void k_count_groups(GTable g, uint64_t *out) {
  if (g.S_dev) g.S = *g.S_dev;
  const K *keys = (const K *)g.keys;
  uint32_t c = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < g.S; i += (uint64_t)gridDim.x * blockDim.x)
    c += keys[i] != KeyTraits<K>::kEmpty;
  if (blockIdx.x == 0 && threadIdx.x == 0) c += g.flags[FLAG_SENTINEL] != 0;
  const int lane = (int)__lane_id();
  for (int o = kWave / 2; o > 0; o >>= 1) c += (uint32_t)__shfl((int)c, lane ^ o, kWave);
  if (lane == 0 && c) atomicAdd((unsigned long long *)out, (unsigned long long)c);
}

// ---------------------------------------------------------------------------
// synthetic data (counter-based; integer streams match the host generators)
// ---------------------------------------------------------------------------

TGA_HD uint64_t splitmix_at(uint64_t seed, uint64_t i) { return mix64(seed + (i + 1) * 0x9e3779b97f4a7c15ULL); }
TGA_HD uint64_t fastrange32(uint32_t u, uint64_t d) { return (uint64_t)(((unsigned __int128)u * d) >> 32); }

__global__ void k_gen_keys_u32(uint32_t *out, uint64_t n, uint64_t distinct, uint64_t seed, uint64_t salt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = (uint32_t)splitmix_at(seed, i);
    out[i] = fmix32((uint32_t)fastrange32(u, distinct) ^ (uint32_t)salt);
  }
}
__global__ void k_gen_keys_u64(uint64_t *out, uint64_t n, uint64_t distinct, uint64_t seed, uint64_t salt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = (uint32_t)splitmix_at(seed, i);
    out[i] = mix64(fastrange32(u, distinct) ^ salt);
  }
}
__global__ void k_gen_vals_i32(int32_t *out, uint64_t n, int64_t lo, uint64_t span, uint64_t seed) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = (uint32_t)(splitmix_at(seed, i) >> 32);
    out[i] = (int32_t)(lo + (int64_t)fastrange32(u, span));
  }
}

// Closed-form workload: key[i] = a bijective mix of (i mod D) (the fmix32 /
// SplitMix64 finalizer of (i mod D) ^ salt) and value[i] = lo + i (two's
// complement, low 32 bits).  Group k then holds c = floor((n-1-k)/D) + 1 rows
// with SUM = c*(lo+k) + D*c*(c-1)/2, MIN = lo + k and MAX = lo + k + D*(c-1)
// (while lo + n - 1 fits int32), so every group's aggregates are known
// without aggregating.
__global__ void k_gen_mod_u32(uint32_t *out, uint64_t n, uint64_t distinct, uint64_t salt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = fmix32((uint32_t)(i % distinct) ^ (uint32_t)salt);
}
__global__ void k_gen_mod_u64(uint64_t *out, uint64_t n, uint64_t distinct, uint64_t salt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = mix64((i % distinct) ^ salt);
}
__global__ void k_gen_seq_i32(int32_t *out, uint64_t n, uint64_t lo) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (int32_t)(uint32_t)(lo + i);
}

// Zipf(s) ranks on [1, D] by rejection-inversion (Hoermann & Derflinger 1996),
// mapped to distinct 64-bit keys mix64((rank-1) ^ salt).
struct ZipfConst { double s, hx1, hn, sh; uint64_t D; };
__device__ __forceinline__ double zh1(double x) { return fabs(x) > 1e-8 ? log1p(x) / x : 1.0 - x * (0.5 - x * (1.0 / 3.0)); }
__device__ __forceinline__ double zh2(double x) { return fabs(x) > 1e-8 ? expm1(x) / x : 1.0 + x * 0.5 * (1.0 + x * (1.0 / 3.0)); }
__device__ __forceinline__ double zH(double x, double s) { const double lx = log(x); return zh2((1.0 - s) * lx) * lx; }
__device__ __forceinline__ double zh(double x, double s) { return exp(-s * log(x)); }
__device__ __forceinline__ double zHinv(double x, double s) {
  double t = x * (1.0 - s);
  if (t < -1.0) t = -1.0;
  return exp(zh1(t) * x);
}
__global__ void k_gen_zipf_u64(uint64_t *out, uint64_t n, ZipfConst z, uint64_t seed, uint64_t salt) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = 1;
    for (int it = 0; it < 64; ++it) {
      const double u01 = (double)(splitmix_at(seed, i * 64 + it) >> 11) * 0x1.0p-53;
      const double u = z.hn + u01 * (z.hx1 - z.hn);
      const double x = zHinv(u, z.s);
      double kx = floor(x + 0.5);
      if (kx < 1.0) kx = 1.0;
      if (kx > (double)z.D) kx = (double)z.D;
      k = (uint64_t)kx;
      if (kx - x <= z.sh || u >= zH(kx + 0.5, z.s) - zh(kx, z.s)) break;
    }
    out[i] = mix64((k - 1) ^ salt);
  }
}
__global__ void k_gen_normal_f64(double *out, uint64_t n, double mean, double sigma, uint64_t seed) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const double u1 = ((double)(splitmix_at(seed, 2 * i) >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
    const double u2 = (double)(splitmix_at(seed, 2 * i + 1) >> 11) * 0x1.0p-53;
    out[i] = mean + sigma * sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
  }
}

// The reference's workload (test/test.cu:96-129, data_generator.hpp:7-85),
// restated as a counter-based generator: `card` distinct u32 keys (a 32-bit
// bijection of their index, never the empty marker 0xffffffff), each written
// once, the other n - card rows copying the key of an index drawn uniformly
// (dist 0), from the Zipf distribution of Gray et al. 1994 with theta < 1
// (dist 1: the zipfian_int_distribution the reference includes), or
// self-similar with skew h (dist 2: index = card * u^(log h / log(1 - h))),
// and the rows shuffled by a bijection of [0, n) (a 4-round Feistel network
// on the next even bit width, cycle-walked into range).  So the group count
// is exactly `card`.
struct RefGenConst {
  uint64_t n, card, seed;
  uint32_t dist, half;  // Feistel half width in bits
  double zetan, alpha, eta, zcut2, ss_exp;
};

__device__ __forceinline__ uint32_t mix32(uint32_t x) {  // a bijection of the 32-bit integers
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ uint64_t feistel(uint64_t x, uint32_t half, uint64_t seed) {
  const uint64_t mask = (1ULL << half) - 1;
  uint64_t l = x >> half, r = x & mask;
#pragma unroll
  for (uint32_t round = 0; round < 4; ++round) {
    const uint64_t f = mix32((uint32_t)r ^ (uint32_t)(seed >> (16 * round)) ^ (round * 0x9e3779b9U)) & mask;
    const uint64_t t = l ^ f;
    l = r;
    r = t;
  }
  return (l << half) | r;
}

__global__ void k_gen_ref_u32(uint32_t *out, RefGenConst c) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < c.n; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t s = feistel(i, c.half, c.seed);
    while (s >= c.n) s = feistel(s, c.half, c.seed);  // cycle-walk: ends, the domain is < 4n
    uint64_t idx = s;
    if (s >= c.card) {
      const double u = (double)(splitmix_at(c.seed ^ 0xa5a5a5a5ULL, s) >> 11) * 0x1.0p-53;  // [0, 1)
      double x;
      if (c.dist == 0) {
        x = u * (double)c.card;
      } else if (c.dist == 1) {
        const double uz = u * c.zetan;
        x = uz < 1.0 ? 0.0 : uz < c.zcut2 ? 1.0 : (double)c.card * pow(c.eta * u - c.eta + 1.0, c.alpha);
      } else {
        x = (double)c.card * pow(u, c.ss_exp);
      }
      idx = (uint64_t)x;
      if (idx >= c.card) idx = c.card - 1;
    }
    uint32_t key = mix32((uint32_t)idx);
    if (key == 0xffffffffU) key = mix32((uint32_t)c.card);  // the one index that maps to the empty marker
    out[i] = key;
  }
}

// key sample for the cardinality estimate
// stats (optional): the four counters k_sample_stats accumulates into,
// zeroed here.
template <typename K>
__global__ void k_sample_keys(const K *keys, uint64_t n, uint64_t m, uint64_t *out, uint64_t *stats) {
  if (stats && blockIdx.x == 0 && threadIdx.x < 4) stats[threadIdx.x] = 0;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t row = sample_row(i, m, n);
    out[i] = (uint64_t)keys[row];
  }
}

// Statistics of a key sample for the cardinality estimate: of the m <=
// kSampleMax keys in `smp` (gathered by k_sample_keys), out = {distinct keys,
// keys seen once, largest count of one key, overflow}, accumulated with
// device atomics into counters k_sample_keys zeroed.  kSampleGroups
// workgroups each count the keys whose hash has its top 4 bits equal to the
// workgroup's index, in an LDS hash table (the sample is read whole by each:
// 128 KB from the L2); the key equal to the empty marker is counted by
// workgroup 0 on its own.  overflow = 1: some workgroup's keys did not fit
// its table (only keys chosen against the hash can do that), the counts are
// incomplete.
constexpr uint32_t kSampleMax = 16384, kSampleGroups = 16, kSampleSlots = 4096;

// Cardinality from a hash slice: the distinct keys whose partition hash
// (key_partition's top kSliceBits bits) is 0, counted exactly over all n
// rows, so D ~ count * 2^kSliceBits with relative error ~ (D/1024)^-1/2
// whatever the key frequencies.  Each workgroup takes a contiguous range of
// rows and dedups its slice keys in an LDS set (a hot key only reads it);
// then it inserts them into the global set, counting new keys with one
// atomic per wave.  Past half full, a workgroup's keys go to the global set
// directly.  out[0] += the count; out[1] != 0: the global set overflowed and
// the count is a lower bound.  The empty-marker key is not counted.
constexpr uint32_t kSliceLds = 4096;

template <typename K>
__device__ __forceinline__ uint32_t slice_set_insert(K *set, uint32_t slots, K k, unsigned long long *out) {
  uint32_t s = (uint32_t)(((uint64_t)key_slot_hash((uint64_t)k) * slots) >> 32);
  for (uint32_t i = 0; i < 256; ++i) {
    const K cur = set[s];
    if (cur == k) return 0;
    if (cur == KeyTraits<K>::kEmpty) {
      const K prev = atomicCAS(&set[s], KeyTraits<K>::kEmpty, k);
      if (prev == KeyTraits<K>::kEmpty) return 1;
      if (prev == k) return 0;
    }
    s = (s + 1 == slots) ? 0 : s + 1;
  }
  atomicOr(&out[1], 1ULL);
  return 0;
}

// A slice key seen by a workgroup: deduplicated in its LDS set while that is
// under half full, else inserted into the global set (new keys counted into
// *added).  Keys outside the slice and the empty marker are skipped.
template <typename K>
__device__ __forceinline__ void slice_visit(K k, K *lset, uint32_t *s_fill, K *set, uint32_t slots,
                                            unsigned long long *out, uint32_t &added) {
  if (key_partition((uint64_t)k, kSliceBits) != 0 || k == KeyTraits<K>::kEmpty) return;
  if (*(volatile uint32_t *)s_fill < kSliceLds / 2) {
    uint32_t s = key_slot_hash((uint64_t)k) & (kSliceLds - 1);
    for (int i = 0; i < 64; ++i) {
      const K cur = lset[s];
      if (cur == k) return;
      if (cur == KeyTraits<K>::kEmpty) {
        const K prev = atomicCAS(&lset[s], KeyTraits<K>::kEmpty, k);
        if (prev == KeyTraits<K>::kEmpty) atomicAdd(s_fill, 1u);
        if (prev == KeyTraits<K>::kEmpty || prev == k) return;
      }
      s = (s + 1) & (kSliceLds - 1);
    }
  }
  added += slice_set_insert<K>(set, slots, k, out);
}

// The workgroup's LDS set into the global set, and its count of new keys
// into out[0] (one atomic per wave).  After a barrier; every thread calls it.
template <typename K>
__device__ __forceinline__ void slice_finish(const K *lset, K *set, uint32_t slots, unsigned long long *out,
                                             uint32_t added) {
  for (uint32_t i = threadIdx.x; i < kSliceLds; i += blockDim.x) {
    const K k = lset[i];
    if (k != KeyTraits<K>::kEmpty) added += slice_set_insert<K>(set, slots, k, out);
  }
  const int lane = (int)__lane_id();
  for (int o = kWave / 2; o > 0; o >>= 1) added += (uint32_t)__shfl((int)added, lane ^ o, kWave);
  if (lane == 0 && added) atomicAdd(&out[0], (unsigned long long)added);
}

// The same count over a partitioned copy of records (rd words each, the key
// first) in bucket layout, whose fanout is at most 2^kSliceBits: the slice
// lies inside partition 0, so only partition 0's buckets (list entries
// [pstart[0], pstart[1]), rows << 32 | bucket id) are read.  Each workgroup
// takes four buckets at a time, one row a thread.
template <typename K>
__global__ void __launch_bounds__(1024) k_slice_records(const uint32_t *__restrict__ rec, uint32_t rd,
                                                        const uint64_t *__restrict__ list,
                                                        const uint64_t *__restrict__ pstart, K *set, uint32_t slots,
                                                        unsigned long long *out) {
  __shared__ K lset[kSliceLds];
  __shared__ uint32_t s_fill;
  for (uint32_t i = threadIdx.x; i < kSliceLds; i += blockDim.x) lset[i] = KeyTraits<K>::kEmpty;
  if (threadIdx.x == 0) s_fill = 0;
  __syncthreads();
  uint32_t added = 0;
  const uint64_t e0 = pstart[0], ne = pstart[1] - e0;
  const uint32_t sub = threadIdx.x >> kBucketShift, r = threadIdx.x & (kBucketRows - 1);
  const uint32_t per = blockDim.x >> kBucketShift;
  for (uint64_t b = (uint64_t)blockIdx.x * per + sub; b < ne; b += (uint64_t)gridDim.x * per) {
    const uint64_t ent = list[e0 + b];
    if (r >= (uint32_t)(ent >> 32)) continue;
    const uint32_t *w = rec + (((ent & 0xffffffffULL) << kBucketShift) + r) * rd;
    K k = (K)w[0];
    if constexpr (sizeof(K) == 8) k = (K)((uint64_t)w[1] << 32 | w[0]);
    slice_visit<K>(k, lset, &s_fill, set, slots, out, added);
  }
  __syncthreads();
  slice_finish<K>(lset, set, slots, out, added);
}

// VEC: keys 16-byte aligned, read 16 bytes a lane, 4 loads in flight.
template <typename K, bool VEC>
__global__ void __launch_bounds__(1024) k_slice_distinct(const K *__restrict__ keys, uint64_t n, K *set,
                                                         uint32_t slots, unsigned long long *out) {
  __shared__ K lset[kSliceLds];
  __shared__ uint32_t s_fill;
  for (uint32_t i = threadIdx.x; i < kSliceLds; i += blockDim.x) lset[i] = KeyTraits<K>::kEmpty;
  if (threadIdx.x == 0) s_fill = 0;
  __syncthreads();
  uint32_t added = 0;
  auto visit = [&](K k) { slice_visit<K>(k, lset, &s_fill, set, slots, out, added); };
  constexpr uint32_t E = VEC ? 16 / sizeof(K) : 1;  // keys per load
  constexpr int U = 4;
  const uint64_t nv = n / E;
  const uint64_t per = (nv + gridDim.x - 1) / gridDim.x;
  const uint64_t b = blockIdx.x * per, e = min(nv, b + per);
  for (uint64_t r0 = b + threadIdx.x; r0 < e; r0 += (uint64_t)U * blockDim.x) {
    K ks[U][E];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t r = r0 + (uint64_t)u * blockDim.x;
      if (r >= e) continue;
      if constexpr (VEC) {
        const uint4 q = ((const uint4 *)keys)[r];
        __builtin_memcpy(ks[u], &q, 16);
      } else {
        ks[u][0] = keys[r];
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (r0 + (uint64_t)u * blockDim.x >= e) continue;
#pragma unroll
      for (uint32_t j = 0; j < E; ++j) visit(ks[u][j]);
    }
  }
  if (VEC && blockIdx.x == 0)  // rows past the last whole vector
    for (uint64_t r = nv * E + threadIdx.x; r < n; r += blockDim.x) visit(keys[r]);
  __syncthreads();
  slice_finish<K>(lset, set, slots, out, added);
}


__global__ void __launch_bounds__(1024) k_sample_stats(const uint64_t *__restrict__ smp, uint32_t m, uint64_t *out) {
  extern __shared__ __align__(16) char smem[];
  uint64_t *tk = (uint64_t *)smem;               // [kSampleSlots]
  uint32_t *tc = (uint32_t *)(tk + kSampleSlots);  // [kSampleSlots]
  __shared__ uint32_t s_d, s_f1, s_top, s_empty, s_over;
  const uint32_t tid = threadIdx.x, nt = blockDim.x, grp = blockIdx.x;
  for (uint32_t sl = tid; sl < kSampleSlots; sl += nt) {
    tk[sl] = ~0ULL;
    tc[sl] = 0;
  }
  if (tid == 0) s_d = s_f1 = s_top = s_empty = s_over = 0;
  __syncthreads();
  uint32_t empties = 0;
  for (uint32_t i = tid; i < m; i += nt) {
    const uint64_t key = smp[i];
    if (key == ~0ULL) {
      empties += grp == 0;
      continue;
    }
    const uint64_t h = key_hash(key);
    if ((uint32_t)(h >> 60) != grp) continue;
    uint32_t sl = (uint32_t)h & (kSampleSlots - 1);
    bool done = false;
    for (uint32_t p = 0; p < kSampleSlots && !done; ++p) {
      const uint64_t prev = atomicCAS((unsigned long long *)&tk[sl], ~0ULL, (unsigned long long)key);
      if (prev == ~0ULL || prev == key) {
        atomicAdd(&tc[sl], 1u);
        done = true;
      } else {
        sl = (sl + 1) & (kSampleSlots - 1);
      }
    }
    if (!done) s_over = 1;
  }
  if (empties) atomicAdd(&s_empty, empties);
  __syncthreads();
  uint32_t d = 0, f1 = 0, top = 0;
  for (uint32_t sl = tid; sl < kSampleSlots; sl += nt) {
    const uint32_t c = tc[sl];
    d += c != 0;
    f1 += c == 1;
    top = max(top, c);
  }
  if (d) atomicAdd(&s_d, d);
  if (f1) atomicAdd(&s_f1, f1);
  if (top) atomicMax(&s_top, top);
  __syncthreads();
  if (tid == 0) {
    const uint32_t c = s_empty;  // nonzero in workgroup 0 only
    const uint64_t dd = s_d + (c != 0), ff = s_f1 + (c == 1), tt = max(s_top, c);
    if (dd) atomicAdd((unsigned long long *)&out[0], (unsigned long long)dd);
    if (ff) atomicAdd((unsigned long long *)&out[1], (unsigned long long)ff);
    if (tt) atomicMax((unsigned long long *)&out[2], (unsigned long long)tt);
    if (s_over) atomicOr((unsigned long long *)&out[3], 1ULL);
  }
}

__global__ void k_gather_words(WordGather wg, uint32_t *dst) {
  uint32_t o = 0;
  for (uint32_t i = 0; i < wg.count; ++i) {
    for (uint32_t j = threadIdx.x; j < wg.n[i]; j += blockDim.x) dst[o + j] = wg.src[i][j];
    o += wg.n[i];
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------

static inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 16384) {
  uint64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_copy16(const void *src, void *dst, uint64_t n16, hipStream_t st) {
  if (n16 == 0) return hipSuccess;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  hipLaunchKernelGGL(k_copy16, dim3(grid_for(n16, 4096, (uint64_t)cus)), dim3(1024), 0, st, (const uint4 *)src,
                     (uint4 *)dst, n16);
  return hipGetLastError();
}

hipError_t launch_gtable_init(int key_bytes, const GTable &g, const Schema &s, hipStream_t st, const WordInit &wi) {
  // sized for the allocation; with S_dev set most of it may go unused
  const unsigned grid = grid_for((g.S + 1) * kLimbs, 256, g.S_dev ? 2048 : 8192);
  if (key_bytes == 4) hipLaunchKernelGGL(k_gtable_init<uint32_t>, dim3(grid), dim3(256), 0, st, g, s, wi);
  else hipLaunchKernelGGL(k_gtable_init<uint64_t>, dim3(grid), dim3(256), 0, st, g, s, wi);
  return hipGetLastError();
}

// A kernel's dynamic-LDS limit, raised past 64 KB once per (device, kernel)
// and size: the attribute call costs host time on every launch otherwise.
hipError_t raise_lds_limit(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, size_t> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> g(mu);
  size_t &have = done[{dev, fn}];
  if (have >= bytes) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
  if (e == hipSuccess) have = bytes;
  return e;
}

template <typename K>
static hipError_t set_lds_limit(const void *fn, size_t bytes) {
  (void)sizeof(K);
  return raise_lds_limit(fn, bytes);
}

bool copy_as_records(const Schema &s) {
  return s.nfx == 0 && s.ncols == 1 && dtype_bytes(s.col_dtype[0]) == 4;
}

int record_value_dwords(const Schema &s) {
  if (s.ncols < 1) return 0;
  int vd = 0;
  for (int c = 0; c < s.ncols; ++c) vd += dtype_bytes(s.col_dtype[c]) / 4;
  return vd <= kMaxRecValueDwords ? vd : 0;
}

hipError_t launch_plan_direct(const AggPlan &plan, uint64_t n, uint32_t nchunks, hipStream_t st) {
  hipLaunchKernelGGL(k_plan_direct, dim3(1), dim3(64), 0, st, plan, n, nchunks);
  return hipGetLastError();
}

uint64_t scan_tiles(uint64_t n) { return (n + kScanThreads * kScanItems - 1) / (kScanThreads * kScanItems); }

hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *desc, uint32_t *ticket,
                           hipStream_t st) {
  const uint64_t tiles = scan_tiles(n);
  if (tiles == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_u32, dim3((unsigned)tiles), dim3(kScanThreads), 0, st, in, out, n, desc, ticket);
  return hipGetLastError();
}

size_t scatter_lds_bytes(const Schema &s, int key_bytes, int pbits, uint32_t tile_rows) {
  size_t b = (size_t)32 << pbits;  // wdesc + 2 histograms + fill + cur
  size_t row = (size_t)key_bytes + 2;
  for (int c = 0; c < s.ncols; ++c) row += (size_t)dtype_bytes(s.col_dtype[c]);
  return b + row * tile_rows;
}

// Rows per thread of the plain scatter: 8 down to 1, so a tile of 1024 to
// 8192 rows always has 1024 threads.  At high fanouts the per-partition LDS
// arrays leave room for 1024-row tiles only; with 8 rows per thread those ran
// on 128 threads, and the per-tile work over all partitions (scan, bucket
// chains) fell on 2 waves (2^29 rows of a three-column schema at 2^12
// partitions: 80 ms).
static int scatter_items(uint32_t tile_rows) {
  return tile_rows >= 8192 ? 8 : tile_rows >= 4096 ? 4 : tile_rows >= 2048 ? 2 : 1;
}

template <typename K, int NC, int VW, bool BK, bool REC, int ITEMS>
static hipError_t launch_scatter_items(const void *keys, const Schema &s, uint64_t n, int pbits, uint32_t tile_rows,
                                       uint32_t ntiles, void *okeys, const PartOut &o, const BucketBufs *bb,
                                       const uint32_t *offs, const uint32_t *pbase, unsigned grid, hipStream_t st) {
  const unsigned block = tile_rows / ITEMS;
  const size_t lds = scatter_lds_bytes(s, sizeof(K), pbits, tile_rows);
  hipError_t e = set_lds_limit<K>((const void *)&k_part_scatter<K, ITEMS, NC, VW, BK, REC>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_part_scatter<K, ITEMS, NC, VW, BK, REC>), dim3(grid), dim3(block), lds, st, (const K *)keys, s, n,
                     pbits, ntiles, (K *)okeys, o, bb ? bb->bpart : nullptr, bb ? bb->bcnt : nullptr,
                     bb ? bb->pool : nullptr, bb ? bb->cap : 0u, offs, pbase);
  return hipGetLastError();
}

template <typename K, int NC, int VW, bool BK, bool REC = false>
static hipError_t launch_scatter_t(const void *keys, const Schema &s, uint64_t n, int pbits, uint32_t tile_rows,
                                   uint32_t ntiles, void *okeys, const PartOut &o, const BucketBufs *bb,
                                   const uint32_t *offs, const uint32_t *pbase, unsigned grid, hipStream_t st) {
  switch (scatter_items(tile_rows)) {
    case 8: return launch_scatter_items<K, NC, VW, BK, REC, 8>(keys, s, n, pbits, tile_rows, ntiles, okeys, o, bb, offs, pbase, grid, st);
    case 4: return launch_scatter_items<K, NC, VW, BK, REC, 4>(keys, s, n, pbits, tile_rows, ntiles, okeys, o, bb, offs, pbase, grid, st);
    case 2: return launch_scatter_items<K, NC, VW, BK, REC, 2>(keys, s, n, pbits, tile_rows, ntiles, okeys, o, bb, offs, pbase, grid, st);
    default: return launch_scatter_items<K, NC, VW, BK, REC, 1>(keys, s, n, pbits, tile_rows, ntiles, okeys, o, bb, offs, pbase, grid, st);
  }
}

unsigned bucket_grid(uint32_t ntiles, unsigned max_grid) {
  // at least 8 tiles per workgroup, so its buckets fill up
  unsigned g = std::max<unsigned>(1, std::min<unsigned>(max_grid, ntiles / 8));
  return xcd_grid(g);
}

unsigned scatter_block(uint32_t tile_rows, int wc) { return wc ? 1024u : tile_rows / (unsigned)scatter_items(tile_rows); }

uint64_t bucket_cap(uint64_t n, unsigned grid, int pbits, unsigned block) {
  return (n + kBucketRows - 1) / kBucketRows + (uint64_t)grid * ((1ULL << pbits) + block) + 1;
}

// Records of one value word take 8192-, 6144- or 4096-row tiles; wider ones
// 4096 rows only (their two register sets of 8 rows a lane would not fit the
// 128 registers of a 1024-thread workgroup).
uint32_t scatter_wc_tile(int key_bytes, int vd, int pbits, size_t lds_max) {
  if ((1u << pbits) > kWcMaxP || vd < 1 || vd > kMaxRecValueDwords) return 0;
  for (uint32_t items : {8u, 6u, 4u}) {
    if (vd > 1 && items != 4) continue;
    const uint32_t tr = items * 1024;
    if (scatter_wc_lds_bytes(key_bytes, vd, pbits, tr) + 256 <= lds_max) return tr;
  }
  return 0;
}

static RecSrc rec_src(const Schema &s) {
  RecSrc r{};
  int d = 0;
  for (int c = 0; c < s.ncols && d < kMaxCols; ++c) {
    const uint32_t *p = (const uint32_t *)s.cols[c];
    if (dtype_bytes(s.col_dtype[c]) == 4) {
      r.vp[d] = p;
      r.vsh[d++] = 0;
    } else {
      r.vp[d] = p;
      r.vsh[d++] = 1;
      if (d < kMaxCols) {
        r.vp[d] = p + 1;
        r.vsh[d++] = 1;
      }
    }
  }
  return r;
}

template <typename K, int VD, int ITEMS, bool TWO = false>
static hipError_t launch_scatter_wc_t(const void *keys, const Schema &s, uint64_t n, int pbits, uint32_t tile_rows,
                                      uint32_t ntiles, void *orec, const BucketBufs &bb, unsigned grid,
                                      hipStream_t st, const Pass2In &p2 = Pass2In{}) {
  const size_t lds = scatter_wc_lds_bytes(sizeof(K), VD, pbits, tile_rows);
  hipError_t e = set_lds_limit<K>((const void *)&k_part_scatter_wc<K, VD, ITEMS, TWO>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_part_scatter_wc<K, VD, ITEMS, TWO>), dim3(grid), dim3(1024), lds, st, (const K *)keys, s, n,
                     pbits, ntiles, orec, bb.bpart, bb.bcnt, bb.pool, bb.cap, rec_src(s), p2);
  return hipGetLastError();
}

// Work items of the second pass: first-pass partition p owns items
// [istart[p], istart[p+1]), one per items_per of its list entries; the
// ticket is zeroed.  One workgroup.
__global__ void __launch_bounds__(1024) k_pass2_plan(const uint64_t *pstart, uint32_t P1, uint32_t items_per,
                                                     uint32_t *istart, uint32_t *ticket) {
  __shared__ uint32_t scratch[1024 / kWave + 1];
  const uint32_t tot = block_exclusive_scan_fn<SyncBarrier>(
      P1, scratch,
      [&](uint32_t p) { return (uint32_t)((pstart[p + 1] - pstart[p] + items_per - 1) / items_per); },
      [&](uint32_t p, uint32_t v) { istart[p] = v; });
  if (threadIdx.x == 0) {
    istart[P1] = tot;
    *ticket = 0;
  }
}

hipError_t launch_pass2_plan(const uint64_t *pstart, uint32_t P1, uint32_t items_per, uint32_t *istart,
                             uint32_t *ticket, hipStream_t st) {
  if (P1 == 0 || P1 > 65536 || items_per == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pass2_plan, dim3(1), dim3(1024), 0, st, pstart, P1, items_per, istart, ticket);
  return hipGetLastError();
}

hipError_t launch_part_scatter2(int key_bytes, int vd, const void *rec, const uint64_t *list, const uint64_t *pstart,
                                const uint32_t *istart, uint32_t *ticket, uint32_t P1, int hash_bits,
                                uint32_t items_per, int pbits, void *orec, const BucketBufs &bb, unsigned grid,
                                hipStream_t st) {
  if (pbits < 1 || hash_bits <= pbits || (1u << pbits) > kWcMaxP || items_per == 0) return hipErrorInvalidValue;
  Pass2In p2{rec, list, pstart, istart, ticket, P1, (uint32_t)hash_bits, items_per};
  const Schema s{};
#define TGA_P2_CASE(K)                                                                                  \
  switch (vd) {                                                                                         \
    case 1: return launch_scatter_wc_t<K, 1, 8, true>(nullptr, s, 0, pbits, 8192, 0, orec, bb, grid, st, p2); \
    case 2: return launch_scatter_wc_t<K, 2, 4, true>(nullptr, s, 0, pbits, 4096, 0, orec, bb, grid, st, p2); \
    case 3: return launch_scatter_wc_t<K, 3, 4, true>(nullptr, s, 0, pbits, 4096, 0, orec, bb, grid, st, p2); \
    case 4: return launch_scatter_wc_t<K, 4, 4, true>(nullptr, s, 0, pbits, 4096, 0, orec, bb, grid, st, p2); \
    default: return hipErrorInvalidValue;                                                               \
  }
  if (key_bytes == 4) { TGA_P2_CASE(uint32_t) }
  TGA_P2_CASE(uint64_t)
#undef TGA_P2_CASE
}

uint32_t scatter2_tile(int vd) { return vd > 1 ? 4096u : 8192u; }

hipError_t launch_part_scatter(int key_bytes, const void *keys, const Schema &s, uint64_t n, int pbits,
                               uint32_t tile_rows, void *okeys, const PartOut &o, const BucketBufs *bb,
                               const uint32_t *offs, const uint32_t *pbase, unsigned grid, hipStream_t st, int wc) {
  const uint32_t ntiles = (uint32_t)((n + tile_rows - 1) / tile_rows);
  if (ntiles == 0) return hipSuccess;
  if (wc) {  // records into buckets, complete lines only
    const int vd = record_value_dwords(s);
    if (!bb || !vd || tile_rows % 1024 || (1u << pbits) > kWcMaxP || (vd > 1 && tile_rows != 4096))
      return hipErrorInvalidValue;
#define TGA_WC_CASE(K)                                                                                  \
  if (vd == 2) return launch_scatter_wc_t<K, 2, 4>(keys, s, n, pbits, tile_rows, ntiles, okeys, *bb, grid, st); \
  if (vd == 3) return launch_scatter_wc_t<K, 3, 4>(keys, s, n, pbits, tile_rows, ntiles, okeys, *bb, grid, st); \
  if (vd == 4) return launch_scatter_wc_t<K, 4, 4>(keys, s, n, pbits, tile_rows, ntiles, okeys, *bb, grid, st); \
  switch (tile_rows / 1024) {                                                                           \
    case 4: return launch_scatter_wc_t<K, 1, 4>(keys, s, n, pbits, tile_rows, ntiles, okeys, *bb, grid, st); \
    case 6: return launch_scatter_wc_t<K, 1, 6>(keys, s, n, pbits, tile_rows, ntiles, okeys, *bb, grid, st); \
    case 8: return launch_scatter_wc_t<K, 1, 8>(keys, s, n, pbits, tile_rows, ntiles, okeys, *bb, grid, st); \
    default: return hipErrorInvalidValue;                                                               \
  }
    if (key_bytes == 4) { TGA_WC_CASE(uint32_t) }
    TGA_WC_CASE(uint64_t)
#undef TGA_WC_CASE
  }
  const bool one = s.ncols <= 1;
  const bool w8 = s.ncols >= 1 && dtype_bytes(s.col_dtype[0]) == 8;
  const bool rec = copy_as_records(s);
#define TGA_SC_ARGS keys, s, n, pbits, tile_rows, ntiles, okeys, o, bb, offs, pbase, grid, st
#define TGA_SC_CASE(K, BK)                                                                          \
  if (rec) return launch_scatter_t<K, 1, 4, BK, true>(TGA_SC_ARGS);                                   \
  if (one) return w8 ? launch_scatter_t<K, 1, 8, BK>(TGA_SC_ARGS) : launch_scatter_t<K, 1, 4, BK>(TGA_SC_ARGS); \
  return launch_scatter_t<K, kMaxCols, 0, BK>(TGA_SC_ARGS);
  if (bb) {
    if (key_bytes == 4) { TGA_SC_CASE(uint32_t, true) }
    TGA_SC_CASE(uint64_t, true)
  }
  if (key_bytes == 4) { TGA_SC_CASE(uint32_t, false) }
  TGA_SC_CASE(uint64_t, false)
#undef TGA_SC_CASE
#undef TGA_SC_ARGS
}

hipError_t launch_plan_chunks(const uint32_t *pbase, uint32_t P, uint64_t n, uint64_t chmax, const AggPlan &plan,
                              hipStream_t st) {
  const size_t lds = (size_t)(P + 1) * 4;
  if (hipError_t e = raise_lds_limit((const void *)&k_plan_chunks, lds)) return e;
  hipLaunchKernelGGL(k_plan_chunks, dim3(1), dim3(1024), lds, st, pbase, P, n, chmax, plan);
  return hipGetLastError();
}

hipError_t launch_part_count(int key_bytes, const void *keys, uint64_t n, int pbits, uint32_t tile_rows,
                             uint32_t ntiles, uint32_t *cnt, hipStream_t st) {
  const size_t lds = (size_t)4 << pbits;
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_part_count<uint32_t>, dim3(xcd_grid(ntiles)), dim3(512), lds, st, (const uint32_t *)keys, n,
                       pbits, tile_rows, ntiles, cnt);
  else
    hipLaunchKernelGGL(k_part_count<uint64_t>, dim3(xcd_grid(ntiles)), dim3(512), lds, st, (const uint64_t *)keys, n,
                       pbits, tile_rows, ntiles, cnt);
  return hipGetLastError();
}

uint64_t colscan_desc_words(uint32_t ntiles, uint32_t P) {
  return (uint64_t)((ntiles + kColTiles - 1) / kColTiles) * P;
}

hipError_t launch_colscan(const uint32_t *cnt, uint32_t ntiles, uint32_t P, uint64_t *desc, uint32_t *ticket,
                          uint32_t *offs, uint32_t *total, uint32_t *pbase, hipStream_t st) {
  const uint32_t nchunks = (ntiles + kColTiles - 1) / kColTiles;
  hipLaunchKernelGGL(k_colscan, dim3(nchunks), dim3(256), 0, st, cnt, ntiles, P, desc, ticket, offs, total);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = (size_t)P * 4;
  e = raise_lds_limit((const void *)&k_colbase, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_colbase, dim3(1), dim3(1024), lds, st, total, P, pbase);
  return hipGetLastError();
}

static hipError_t lds_attr(const void *fn, size_t bytes) { return raise_lds_limit(fn, bytes); }

hipError_t launch_bucket_lists(const BucketBufs &bb, uint32_t P, uint64_t chmax, const AggPlan &plan,
                               hipStream_t st, bool for_agg, bool pair) {
  const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(256, bb.cap / 4096));
  size_t lds = (size_t)P * 8;
  hipError_t e = lds_attr((const void *)&k_bucket_count, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_count, dim3(grid), dim3(1024), lds, st, bb.bpart, bb.bcnt, bb.pool, bb.cap, P, bb.pbk,
                     bb.prow);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  lds = (size_t)P * 8;
  if ((e = lds_attr((const void *)&k_bucket_plan, lds)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_plan, dim3(1), dim3(1024), lds, st, bb.pbk, bb.prow, P, chmax, bb.bcur, bb.bend, plan);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = lds_attr((const void *)&k_bucket_list, lds)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_list, dim3(grid), dim3(1024), lds, st, bb.bpart, bb.bcnt, bb.pool, bb.cap, P, bb.bcur,
                     bb.bend, bb.blist);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return for_agg ? launch_bucket_pairs(bb, P, plan, st, pair) : hipSuccess;
}

hipError_t launch_bucket_pairs(const BucketBufs &bb, uint32_t P, const AggPlan &plan, hipStream_t st, bool pair) {
  hipLaunchKernelGGL(k_bucket_pair, dim3(P), dim3(256), 0, st, plan.pstart, bb.bend, bb.blist, bb.blist2,
                     (uint32_t)pair);
  return hipGetLastError();
}

hipError_t launch_count_groups(int key_bytes, const GTable &g, uint64_t *out, hipStream_t st) {
  const unsigned grid = grid_for(g.S + 1, 256, 2048);
  if (key_bytes == 4) hipLaunchKernelGGL(k_count_groups<uint32_t>, dim3(grid), dim3(256), 0, st, g, out);
  else hipLaunchKernelGGL(k_count_groups<uint64_t>, dim3(grid), dim3(256), 0, st, g, out);
  return hipGetLastError();
}

uint64_t compact_tiles(uint64_t slots_plus_one) {
  const uint64_t TS = (uint64_t)kCompactThreads * kCompactItems;
  return (slots_plus_one + TS - 1) / TS;
}

hipError_t launch_compact(int key_bytes, const GTable &g, const Schema &s, const OutDesc &o, uint64_t *desc,
                          uint32_t *ticket, uint64_t *count_out, hipStream_t st, const uint64_t *base) {
  // workgroups loop over the tiles (with S_dev set, over the used ones only)
  const unsigned grid = (unsigned)std::min<uint64_t>(compact_tiles(g.S + 1), 2048);
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_compact<uint32_t>, dim3(grid), dim3(kCompactThreads), 0, st, g, s, o, desc, ticket, count_out,
                       base);
  else
    hipLaunchKernelGGL(k_compact<uint64_t>, dim3(grid), dim3(kCompactThreads), 0, st, g, s, o, desc, ticket, count_out,
                       base);
  return hipGetLastError();
}

hipError_t launch_sample_keys(int key_bytes, const void *keys, uint64_t n, uint64_t m, uint64_t *out,
                              hipStream_t st, uint64_t *stats) {
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_sample_keys<uint32_t>, dim3(grid_for(m, 256)), dim3(256), 0, st, (const uint32_t *)keys, n, m,
                       out, stats);
  else
    hipLaunchKernelGGL(k_sample_keys<uint64_t>, dim3(grid_for(m, 256)), dim3(256), 0, st, (const uint64_t *)keys, n, m,
                       out, stats);
  return hipGetLastError();
}

hipError_t launch_gather_words(const WordGather &wg, uint32_t *dst, hipStream_t st) {
  uint32_t total = 0;
  for (uint32_t i = 0; i < wg.count; ++i) total += wg.n[i];
  if (wg.count > 8 || total > (uint32_t)kStatusWords) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_gather_words, dim3(1), dim3(64), 0, st, wg, dst);
  return hipGetLastError();
}

hipError_t launch_sample_stats(int key_bytes, const void *keys, uint64_t n, uint32_t m, uint64_t *smp, uint64_t *out,
                               hipStream_t st) {
  if (m == 0 || m > kSampleMax) return hipErrorInvalidValue;
  hipError_t e = launch_sample_keys(key_bytes, keys, n, m, smp, st, out);
  if (e != hipSuccess) return e;
  const size_t lds = (size_t)kSampleSlots * 12;
  e = set_lds_limit<uint64_t>((const void *)&k_sample_stats, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sample_stats, dim3(kSampleGroups), dim3(1024), lds, st, (const uint64_t *)smp, m, out);
  return hipGetLastError();
}

uint32_t slice_set_slots(uint64_t n) {
  uint64_t S = 4096;
  while (S < 4 * (n >> kSliceBits) && S < (1ULL << 31)) S <<= 1;
  return (uint32_t)S;
}

hipError_t launch_slice_distinct(int key_bytes, const void *keys, uint64_t n, void *set, uint32_t slots, uint64_t *out,
                                 unsigned grid, hipStream_t st) {
  if (slots == 0 || (slots & (slots - 1)) || grid == 0) return hipErrorInvalidValue;
  hipError_t e = launch_fill_u64(out, 2, 0, st);
  if (e != hipSuccess) return e;
  const bool vec = ((uintptr_t)keys & 15) == 0;
  if (key_bytes == 4) {
    if ((e = launch_fill_u32((uint32_t *)set, slots, 0xFFFFFFFFu, st)) != hipSuccess) return e;
    auto *fn = vec ? &k_slice_distinct<uint32_t, true> : &k_slice_distinct<uint32_t, false>;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(1024), 0, st, (const uint32_t *)keys, n, (uint32_t *)set, slots,
                       (unsigned long long *)out);
  } else {
    if ((e = launch_fill_u64((uint64_t *)set, slots, ~0ULL, st)) != hipSuccess) return e;
    auto *fn = vec ? &k_slice_distinct<uint64_t, true> : &k_slice_distinct<uint64_t, false>;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(1024), 0, st, (const uint64_t *)keys, n, (uint64_t *)set, slots,
                       (unsigned long long *)out);
  }
  return hipGetLastError();
}

hipError_t launch_slice_records(int key_bytes, const void *rec, uint32_t rd, const uint64_t *list,
                                const uint64_t *pstart, void *set, uint32_t slots, uint64_t *out, unsigned grid,
                                hipStream_t st) {
  if (slots == 0 || (slots & (slots - 1)) || grid == 0 || rd < (uint32_t)key_bytes / 4) return hipErrorInvalidValue;
  hipError_t e = launch_fill_u64(out, 2, 0, st);
  if (e != hipSuccess) return e;
  if (key_bytes == 4) {
    if ((e = launch_fill_u32((uint32_t *)set, slots, 0xFFFFFFFFu, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_slice_records<uint32_t>, dim3(grid), dim3(1024), 0, st, (const uint32_t *)rec, rd, list,
                       pstart, (uint32_t *)set, slots, (unsigned long long *)out);
  } else {
    if ((e = launch_fill_u64((uint64_t *)set, slots, ~0ULL, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_slice_records<uint64_t>, dim3(grid), dim3(1024), 0, st, (const uint32_t *)rec, rd, list,
                       pstart, (uint64_t *)set, slots, (unsigned long long *)out);
  }
  return hipGetLastError();
}

hipError_t launch_gen(int kind, void *out, uint64_t n, uint64_t a, uint64_t b, double x, double y, uint64_t seed,
                      uint64_t salt, hipStream_t st) {
  const unsigned grid = grid_for(n, 256, 8192);
  switch (kind) {
    case GEN_KEYS_U32: hipLaunchKernelGGL(k_gen_keys_u32, dim3(grid), dim3(256), 0, st, (uint32_t *)out, n, a, seed, salt); break;
    case GEN_KEYS_U64: hipLaunchKernelGGL(k_gen_keys_u64, dim3(grid), dim3(256), 0, st, (uint64_t *)out, n, a, seed, salt); break;
    case GEN_VALS_I32: hipLaunchKernelGGL(k_gen_vals_i32, dim3(grid), dim3(256), 0, st, (int32_t *)out, n, (int64_t)a, b, seed); break;
    case GEN_ZIPF_U64: {
      ZipfConst z;
      z.s = x; z.D = a;
      // host-side constants of the rejection-inversion sampler
      auto hh1 = [](double v) { return std::fabs(v) > 1e-8 ? std::log1p(v) / v : 1.0 - v * (0.5 - v * (1.0 / 3.0)); };
      auto hh2 = [](double v) { return std::fabs(v) > 1e-8 ? std::expm1(v) / v : 1.0 + v * 0.5 * (1.0 + v * (1.0 / 3.0)); };
      auto H = [&](double v) { const double lv = std::log(v); return hh2((1.0 - x) * lv) * lv; };
      auto hfn = [&](double v) { return std::exp(-x * std::log(v)); };
      auto Hinv = [&](double v) { double t = v * (1.0 - x); if (t < -1.0) t = -1.0; return std::exp(hh1(t) * v); };
      z.hx1 = H(1.5) - 1.0;
      z.hn = H((double)a + 0.5);
      z.sh = 2.0 - Hinv(H(2.5) - hfn(2.0));
      hipLaunchKernelGGL(k_gen_zipf_u64, dim3(grid), dim3(256), 0, st, (uint64_t *)out, n, z, seed, salt);
      break;
    }
    case GEN_MOD_U32: hipLaunchKernelGGL(k_gen_mod_u32, dim3(grid), dim3(256), 0, st, (uint32_t *)out, n, a, salt); break;
    case GEN_MOD_U64: hipLaunchKernelGGL(k_gen_mod_u64, dim3(grid), dim3(256), 0, st, (uint64_t *)out, n, a, salt); break;
    case GEN_SEQ_I32: hipLaunchKernelGGL(k_gen_seq_i32, dim3(grid), dim3(256), 0, st, (int32_t *)out, n, a); break;
    case GEN_NORMAL_F64: hipLaunchKernelGGL(k_gen_normal_f64, dim3(grid), dim3(256), 0, st, (double *)out, n, x, y, seed); break;
    case GEN_REF_U32: {
      // a = cardinality (1 <= a <= n, a < 2^32 - 1), b = dist (0 uniform, 1 Zipf, 2 self-similar), x = theta or h
      if (a < 1 || a > n || a >= 0xffffffffULL || b > 2 || n >= (1ULL << 32)) return hipErrorInvalidValue;
      if (b != 0 && !(x > 0.0 && x < 1.0)) return hipErrorInvalidValue;
      RefGenConst c{};
      c.n = n;
      c.card = a;
      c.seed = seed ^ (salt * 0x9e3779b97f4a7c15ULL);
      c.dist = (uint32_t)b;
      uint32_t w = 2;
      while ((1ULL << w) < n) w += 2;
      c.half = w / 2;
      if (b == 1) {
        double z = 0.0;  // zeta(card, theta), summed as the reference does
        for (uint64_t i = 1; i <= a; ++i) z += std::pow(1.0 / (double)i, x);
        const double z2 = 1.0 + std::pow(0.5, x);
        c.zetan = z;
        c.alpha = 1.0 / (1.0 - x);
        c.eta = (1.0 - std::pow(2.0 / (double)a, 1.0 - x)) / (1.0 - z2 / z);
        c.zcut2 = 1.0 + std::pow(0.5, x);
      } else if (b == 2) {
        c.ss_exp = std::log(x) / std::log(1.0 - x);
      }
      hipLaunchKernelGGL(k_gen_ref_u32, dim3(grid), dim3(256), 0, st, (uint32_t *)out, c);
      break;
    }
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tga
