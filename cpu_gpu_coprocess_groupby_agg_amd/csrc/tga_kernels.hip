// tga_kernels.hip -- gfx950 kernels of the plain group-by operator.
//
//   (a) k_agg           LDS open-addressing pre-aggregation with wave64 match
//                       folding; flushes each workgroup table into the global
//                       table or emits whole partitions.  Used directly on the
//                       input when the groups fit one LDS table, and per
//                       partition after (b).
//   (b) k_part_bucket   count-free partitioning: LDS counting sort of a tile,
//                       then each partition's run appended to the workgroup's
//                       open 256-row bucket of that partition
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

// Plain 16-byte streaming copy (bandwidth reference for the counters).
__global__ void k_copy16(const uint4 *__restrict__ a, uint4 *__restrict__ b, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

template <typename K>
__global__ void k_gtable_init(GTable g, Schema s) {
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
}

// ---------------------------------------------------------------------------
// (a)/(c) LDS aggregation
// ---------------------------------------------------------------------------

// Persistent: workgroups pull work items from plan.ticket until none remain.
// Item -> (segment p, slice j of nch) through plan.cstart; rows of segment p
// are [pstart[p], pstart[p+1]).
template <typename K, bool FX, int NC, int VW, bool FOLD>
__global__ void __launch_bounds__(1024) k_agg(const K *__restrict__ keys, Schema s, AggPlan plan, Layout L,
                                              GTable g, AccProg ap, int nbits, int max_probe, int emit, OutDesc od,
                                              uint64_t *emit_count, uint32_t *ovf) {
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t s_item, s_sentinel;
  __shared__ uint32_t scratch[1024 / kWave + 1];
  __shared__ uint64_t s_base;
  // descriptors the out-of-line global-table path reads
  __shared__ GTable s_g;
  __shared__ Schema s_s;
  __shared__ uint32_t s_abort[2];  // table full / rows consumed (abort mode)
  if (threadIdx.x == 0) {
    s_g = g;
    s_s = s;
  }
  LdsTable<K> t{smem, L, &s_sentinel};
  const uint32_t total = plan.cstart[plan.P];
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(plan.ticket, 1u);
    __syncthreads();
    const uint32_t item = wave_uniform(s_item);
    if (item >= total) break;
    uint32_t lo = 0, hi = plan.P;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (plan.cstart[mid] <= item) lo = mid; else hi = mid;
    }
    const uint32_t p = lo;
    const uint32_t j = item - plan.cstart[p], nch = plan.cstart[p + 1] - plan.cstart[p];
    const uint64_t sb = plan.pstart[p], se = plan.pstart[p + 1], len = se - sb;
    const uint64_t b = sb + len * j / nch, e = sb + len * (j + 1) / nch;
    // with ovf set, a work item whose groups overflow the LDS table is
    // abandoned and recorded in ovf = {count, capacity, (rows consumed, rows)...}
    // instead of spilling: the host then plans a finer partitioning
    const bool whole = emit && nch == 1;
    const bool abandon = !FX && ovf != nullptr;
    lds_init(t, s);
    if (threadIdx.x == 0) {
      s_abort[0] = 0;
      s_abort[1] = 0;
    }
    __syncthreads();
    const BucketList bl{plan.blist};
    if constexpr (FX) {
      agg_rows<K, FX, NC>(keys, s, b, e, t, g, nbits, max_probe, bl);
    } else {
      const void *cols[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) cols[c] = s.cols[c];
      agg_rows_int<K, NC, VW, FOLD>(keys, cols, b, e, t, ap, nbits, max_probe, &s_g, &s_s,
                                    abandon ? s_abort : nullptr, nullptr, bl);
    }
    __syncthreads();
    if (abandon && s_abort[0]) {
      if (threadIdx.x == 0) {
        const uint32_t i = atomicAdd(ovf, 1u);
        if (i < ovf[1]) {
          // rows of the item (buckets counted full)
          const uint64_t rows = plan.blist ? (e - b) << kBucketShift : e - b;
          ovf[2 + 2 * i] = s_abort[1];
          ovf[3 + 2 * i] = (uint32_t)min(rows, (uint64_t)UINT32_MAX);
        }
      }
    } else if (whole) {
      lds_emit<K>(t, s, od, emit_count, g.flags, scratch, &s_base);
    } else {
      lds_flush<K>(t, s, g);
    }
    __syncthreads();
  }
}

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
// With one value column the next tile's rows are loaded into a second
// register set right after this tile is ranked, so they have this tile's
// remaining phases to arrive; with more columns (whose second register set
// would spill) they are loaded after this tile is staged, during its writes.
// VW: width of every value cell when NC == 1 (4 or 8), 0 = per column.
constexpr uint32_t kReserve = 64;

__host__ __device__ inline uint32_t bucket_of(uint32_t k, uint32_t first, uint32_t delta) {
  return k == 0 ? first : delta + k;
}

template <typename K, int ITEMS, int NC, int VW>
__global__ void __launch_bounds__(1024) k_part_bucket(const K *__restrict__ keys, Schema s, uint64_t n, int pbits,
                                                      uint32_t ntiles, K *okeys, PartOut o, uint32_t *bpart,
                                                      uint32_t *bcnt, uint32_t *pool, uint32_t cap) {
  using CellT = typename std::conditional<VW == 4, uint32_t, uint64_t>::type;
  constexpr bool kEarly = NC == 1;
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t scratch[17];
  const uint32_t P = 1u << pbits;
  const uint32_t TR = blockDim.x * ITEMS;
  const int lane = (int)__lane_id();
  // wdesc[p] = {q0, first, delta, -}: row j of the tile's run of p goes to
  // position q = j + q0 of the bucket chain bucket_of(q >> shift, first, delta)
  uint4 *wdesc = (uint4 *)smem;
  uint32_t *hist = (uint32_t *)(wdesc + P);  // tile counts, then exclusive offsets
  uint32_t *fill = hist + P;                 // rows in the open bucket (0: none open)
  uint32_t *cur = fill + P;                  // the open bucket
  char *stage = (char *)(cur + P);
  // stage layout: K keys[TR] | per column [TR] cells | u16 pid[TR]
  K *skey = (K *)stage;
  char *scol = stage + (size_t)TR * sizeof(K);
  const int ncol = NC == 1 ? 1 : s.ncols;
  uint32_t colb[NC];
  size_t coff[NC];
  size_t acc_off = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    colb[c] = VW ? (uint32_t)VW : c < ncol ? (uint32_t)dtype_bytes(s.col_dtype[c]) : 0u;
    coff[c] = acc_off;
    acc_off += (size_t)TR * colb[c];
  }
  uint16_t *spid = (uint16_t *)(scol + acc_off);

  K kk[ITEMS], nkk[ITEMS];
  CellT cell[ITEMS][NC], ncell[ITEMS][NC];
  uint32_t rlo = 0, rhi = 0;  // this wave's reserve of bucket ids (wave-uniform)
  // chunk of consecutive tiles; chunks renumbered so the chunks of one XCD
  // are neighbours too (the grid is a multiple of the XCD count)
  const uint32_t per = (ntiles + gridDim.x - 1) / gridDim.x;
  uint32_t tile = xcd_local_index(blockIdx.x, gridDim.x) * per;
  const uint32_t tend = min(ntiles, tile + per);
  if (tile >= tend) return;  // uniform per workgroup
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    fill[p] = 0;
    cur[p] = 0;
  }
  // rows past the end re-read the last row (in bounds); they are not staged
#define TGA_SCATTER_LOAD(T, KK, CELL)                                                           \
  {                                                                                             \
    const uint64_t b0 = (uint64_t)(T) * TR;                                                     \
    _Pragma("unroll") for (int i = 0; i < ITEMS; ++i) {                                         \
      const uint64_t row = b0 + (uint64_t)i * blockDim.x + threadIdx.x;                         \
      const uint64_t r = row < n ? row : n - 1;                                                 \
      KK[i] = keys[r];                                                                          \
      _Pragma("unroll") for (int c = 0; c < NC; ++c) {                                          \
        if (c >= ncol) { CELL[i][c] = 0; continue; }                                            \
        CELL[i][c] = colb[c] == 4 ? (CellT)((const uint32_t *)s.cols[c])[r]                     \
                                  : (CellT)((const uint64_t *)s.cols[c])[r];                    \
      }                                                                                         \
    }                                                                                           \
  }
  TGA_SCATTER_LOAD(tile, kk, cell)
  for (;;) {
    const uint64_t b = (uint64_t)tile * TR;
    for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) hist[p] = 0;
    __syncthreads();
    uint32_t pp[ITEMS], rk[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint64_t row = b + (uint64_t)i * blockDim.x + threadIdx.x;
      pp[i] = key_partition((uint64_t)kk[i], pbits);
      rk[i] = row < n ? atomicAdd(&hist[pp[i]], 1u) : 0u;
    }
    const uint32_t next = tile + 1;
    if constexpr (kEarly) {
      if (next < tend) TGA_SCATTER_LOAD(next, nkk, ncell)
    }
    __syncthreads();
    const uint32_t rows = block_exclusive_scan_lds(hist, P, scratch);
    // extend every partition's bucket chain by its run of this tile
    for (uint32_t p0 = 0; p0 < P; p0 += blockDim.x) {
      const uint32_t p = p0 + threadIdx.x;
      const bool on = p < P;
      uint32_t off = 0, c = 0, f = 0, first = 0, fresh = 0;
      if (on) {
        off = hist[p];
        c = (p + 1 < P ? hist[p + 1] : rows) - off;
        f = fill[p];
        first = cur[p];
        if (c) fresh = ((f + c + kBucketRows - 1) >> kBucketShift) - (f ? 1u : 0u);
      }
      // single fresh buckets: from the wave's reserve
      const bool one = fresh == 1;
      const uint64_t m1 = __ballot(one);
      const uint32_t n1 = (uint32_t)__builtin_popcountll(m1);
      const uint32_t j = (uint32_t)__builtin_popcountll(m1 & lanemask_lt());
      uint32_t base = 0;
      if (n1) {
        const uint32_t have = rhi - rlo;
        uint32_t got = 0;
        if (have < n1) {  // refill (wave-uniform branch)
          if (lane == 0) got = atomicAdd(pool, kReserve);
          got = wave_uniform((uint32_t)__shfl((int)got, 0, kWave));
        }
        if (one) base = j < have ? rlo + j : got + (j - have);
        if (have < n1) {
          rlo = got + (n1 - have);
          rhi = got + kReserve;
        } else {
          rlo += n1;
        }
      }
      if (fresh > 1) base = atomicAdd(pool, fresh);  // several: contiguous, own atomic
      uint32_t delta = 0;
      if (on && c) {
        for (uint32_t k = 0; k < fresh; ++k)
          if (base + k < cap) bpart[base + k] = p;
        delta = f ? base - 1 : base;
        if (!f) first = base;
        const uint32_t end = f + c;
        const uint32_t full = end >> kBucketShift;  // buckets this run completes
        for (uint32_t k = 0; k < full; ++k) {
          const uint32_t id = bucket_of(k, first, delta);
          if (id < cap) bcnt[id] = kBucketRows;
        }
        fill[p] = end & (kBucketRows - 1);
        cur[p] = bucket_of((end - 1) >> kBucketShift, first, delta);
      }
      if (on) wdesc[p] = make_uint4(f - off, first, delta, 0u);
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint64_t row = b + (uint64_t)i * blockDim.x + threadIdx.x;
      if (row < n) {
        const uint32_t lp = hist[pp[i]] + rk[i];
        skey[lp] = kk[i];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          if (c >= ncol) break;
          if (colb[c] == 4) ((uint32_t *)(scol + coff[c]))[lp] = (uint32_t)cell[i][c];
          else ((uint64_t *)(scol + coff[c]))[lp] = (uint64_t)cell[i][c];
        }
        spid[lp] = (uint16_t)pp[i];
      }
    }
    __syncthreads();
    if constexpr (!kEarly) {
      if (next < tend) TGA_SCATTER_LOAD(next, kk, cell)
    }
    for (uint32_t jr = threadIdx.x; jr < rows; jr += blockDim.x) {
      const uint4 d = wdesc[spid[jr]];
      const uint32_t q = jr + d.x;
      const uint32_t id = bucket_of(q >> kBucketShift, d.y, d.z);
      if (id >= cap) continue;
      const uint64_t dst = ((uint64_t)id << kBucketShift) | (q & (kBucketRows - 1));
      okeys[dst] = skey[jr];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (c >= ncol) break;
        if (colb[c] == 4) ((uint32_t *)o.cols[c])[dst] = ((const uint32_t *)(scol + coff[c]))[jr];
        else ((uint64_t *)o.cols[c])[dst] = ((const uint64_t *)(scol + coff[c]))[jr];
      }
    }
    __syncthreads();
    if (next >= tend) break;
    tile = next;
    if constexpr (kEarly) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        kk[i] = nkk[i];
#pragma unroll
        for (int c = 0; c < NC; ++c) cell[i][c] = ncell[i][c];
      }
    }
  }
#undef TGA_SCATTER_LOAD
  // the open buckets are the last of their chains
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x)
    if (fill[p] && cur[p] < cap) bcnt[cur[p]] = fill[p];
  // unused reserve ids
  for (uint32_t id = rlo + lane; id < rhi; id += kWave)
    if (id < cap) bpart[id] = ~0u;
}

// Per-partition bucket and row totals from the bucket headers (pbk, prow
// zeroed by the caller).
__global__ void __launch_bounds__(1024) k_bucket_count(const uint32_t *__restrict__ bpart,
                                                       const uint32_t *__restrict__ bcnt, const uint32_t *pool,
                                                       uint32_t cap, uint32_t P, uint32_t *pbk, uint64_t *prow) {
  extern __shared__ __align__(16) char smem[];
  uint64_t *lrow = (uint64_t *)smem;
  uint32_t *lbk = (uint32_t *)(lrow + P);
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
    atomicAdd((unsigned long long *)&lrow[p], (unsigned long long)bcnt[id]);
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
                                                      uint64_t chmax, uint32_t *bcur, AggPlan plan) {
  extern __shared__ uint32_t cnt[];  // [2P]
  __shared__ uint32_t scratch[17];
  uint32_t *nbk = cnt, *nit = cnt + P;
  if (threadIdx.x == 0 && plan.split_rows) *plan.split_rows = 0;
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    const uint64_t rows = prow[p];
    nbk[p] = pbk[p];
    nit[p] = (uint32_t)((rows + chmax - 1) / chmax);
    if (nit[p] > 1 && plan.split_rows) atomicAdd((unsigned long long *)plan.split_rows, (unsigned long long)rows);
  }
  __syncthreads();
  const uint32_t buckets = block_exclusive_scan_lds(nbk, P, scratch);
  const uint32_t items = block_exclusive_scan_lds(nit, P, scratch);
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    plan.pstart[p] = nbk[p];
    bcur[p] = nbk[p];
    plan.cstart[p] = nit[p];
  }
  if (threadIdx.x == 0) {
    plan.pstart[P] = buckets;
    plan.cstart[P] = items;
    *plan.ticket = 0;
  }
}

// Buckets grouped by partition: blist[pstart[p] ..] = the entries
// rows << 32 | id of the buckets of p (any order).  Each workgroup counts its
// range of ids per partition in LDS, reserves its share of every list with
// one atomic, then places its entries.
__global__ void __launch_bounds__(1024) k_bucket_list(const uint32_t *__restrict__ bpart,
                                                      const uint32_t *__restrict__ bcnt, const uint32_t *pool,
                                                      uint32_t cap, uint32_t P, uint32_t *bcur, uint64_t *blist) {
  extern __shared__ uint32_t cnt[];  // [2P]
  uint32_t *lc = cnt, *lb = cnt + P;
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) lc[p] = 0;
  __syncthreads();
  const uint32_t nb = min(*pool, cap);
  const uint32_t per = (nb + gridDim.x - 1) / gridDim.x;
  const uint32_t lo = blockIdx.x * per, hi = min(nb, lo + per);
  for (uint32_t id = lo + threadIdx.x; id < hi; id += blockDim.x) {
    const uint32_t p = bpart[id];
    if (p < P) atomicAdd(&lc[p], 1u);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < P; p += blockDim.x) {
    lb[p] = lc[p] ? atomicAdd(&bcur[p], lc[p]) : 0u;
    lc[p] = 0;
  }
  __syncthreads();
  for (uint32_t id = lo + threadIdx.x; id < hi; id += blockDim.x) {
    const uint32_t p = bpart[id];
    if (p >= P) continue;
    blist[lb[p] + atomicAdd(&lc[p], 1u)] = ((uint64_t)bcnt[id] << 32) | id;
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
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint64_t tile = s_tile;
  const uint64_t n = g.S + 1;
  const uint64_t TS = (uint64_t)kCompactThreads * kCompactItems;
  const uint64_t tile_base = tile * TS;
  const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
  const K *keys = (const K *)g.keys;
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

// strided key sample for the cardinality estimate
template <typename K>
__global__ void k_sample_keys(const K *keys, uint64_t n, uint64_t m, uint64_t *out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint64_t)keys[(uint64_t)(((unsigned __int128)i * n) / m)];
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
  hipLaunchKernelGGL(k_copy16, dim3(grid_for(n16, 256, 8192)), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst, n16);
  return hipGetLastError();
}

hipError_t launch_gtable_init(int key_bytes, const GTable &g, const Schema &s, hipStream_t st) {
  const uint64_t n = (g.S + 1) * kLimbs;
  if (key_bytes == 4) hipLaunchKernelGGL(k_gtable_init<uint32_t>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, g, s);
  else hipLaunchKernelGGL(k_gtable_init<uint64_t>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, g, s);
  return hipGetLastError();
}

template <typename K>
static hipError_t set_lds_limit(const void *fn, size_t bytes) {
  (void)sizeof(K);
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename K, bool FX, int NC, int VW, bool FOLD>
static hipError_t launch_agg_t(const void *keys, const Schema &s, const AggPlan &plan, const Layout &L, const GTable &g,
                               int nbits, int max_probe, unsigned grid, unsigned block, hipStream_t st, int emit,
                               const OutDesc &o, uint64_t *emit_count, uint32_t *ovf) {
  const size_t lds = L.bytes;
  hipError_t e = set_lds_limit<K>((const void *)&k_agg<K, FX, NC, VW, FOLD>, lds);
  if (e != hipSuccess) return e;
  const AccProg ap = make_acc_prog(s, L);
  hipLaunchKernelGGL((k_agg<K, FX, NC, VW, FOLD>), dim3(grid), dim3(block), lds, st, (const K *)keys, s, plan, L, g,
                     ap, nbits, max_probe, emit, o, emit_count, ovf);
  return hipGetLastError();
}

// One kernel per (key width, fp64 SUM present, value columns, wave folding):
// each variant gets its own register allocation, and a single value column
// of a known width is read without branches.  fp64 SUM accumulators always
// fold (their limb vectors make per-row atomics costly).
hipError_t launch_agg(int key_bytes, const void *keys, const Schema &s, const AggPlan &plan, const Layout &L,
                      const GTable &g, int nbits, int max_probe, unsigned grid, unsigned block, hipStream_t st,
                      int emit, const OutDesc *od, uint64_t *emit_count, int fold, uint32_t *ovf) {
  OutDesc o{};
  if (od) o = *od;
  const bool fx = s.nfx > 0;
  const bool one = s.ncols <= 1;
  const bool w8 = s.ncols >= 1 && dtype_bytes(s.col_dtype[0]) == 8;
#define TGA_AGG_ARGS keys, s, plan, L, g, nbits, max_probe, grid, block, st, emit, o, emit_count, ovf
#define TGA_AGG_CASE(K)                                                                         \
  if (fx) return launch_agg_t<K, true, kMaxCols, 0, true>(TGA_AGG_ARGS);                        \
  if (one) {                                                                                    \
    if (w8) return fold ? launch_agg_t<K, false, 1, 8, true>(TGA_AGG_ARGS)                      \
                        : launch_agg_t<K, false, 1, 8, false>(TGA_AGG_ARGS);                    \
    return fold ? launch_agg_t<K, false, 1, 4, true>(TGA_AGG_ARGS)                              \
                : launch_agg_t<K, false, 1, 4, false>(TGA_AGG_ARGS);                            \
  }                                                                                             \
  return fold ? launch_agg_t<K, false, kMaxCols, 0, true>(TGA_AGG_ARGS)                         \
              : launch_agg_t<K, false, kMaxCols, 0, false>(TGA_AGG_ARGS);
  if (key_bytes == 4) { TGA_AGG_CASE(uint32_t) }
  TGA_AGG_CASE(uint64_t)
#undef TGA_AGG_CASE
#undef TGA_AGG_ARGS
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
  size_t b = (size_t)28 << pbits;  // wdesc + hist + fill + cur
  size_t row = (size_t)key_bytes + 2;
  for (int c = 0; c < s.ncols; ++c) row += (size_t)dtype_bytes(s.col_dtype[c]);
  return b + row * tile_rows;
}

template <typename K, int NC, int VW>
static hipError_t launch_bucket_t(const void *keys, const Schema &s, uint64_t n, int pbits, uint32_t tile_rows,
                                  uint32_t ntiles, void *okeys, const PartOut &o, const BucketBufs &bb, unsigned grid,
                                  hipStream_t st) {
  constexpr int ITEMS = 8;
  const unsigned block = tile_rows / ITEMS;
  const size_t lds = scatter_lds_bytes(s, sizeof(K), pbits, tile_rows);
  hipError_t e = set_lds_limit<K>((const void *)&k_part_bucket<K, ITEMS, NC, VW>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_part_bucket<K, ITEMS, NC, VW>), dim3(grid), dim3(block), lds, st, (const K *)keys, s, n,
                     pbits, ntiles, (K *)okeys, o, bb.bpart, bb.bcnt, bb.pool, bb.cap);
  return hipGetLastError();
}

unsigned bucket_grid(uint32_t ntiles, unsigned max_grid) {
  // at least 8 tiles per workgroup, so its buckets fill up
  unsigned g = std::max<unsigned>(1, std::min<unsigned>(max_grid, ntiles / 8));
  return xcd_grid(g);
}

uint64_t bucket_cap(uint64_t n, unsigned grid, int pbits, uint32_t tile_rows) {
  const uint64_t block = tile_rows / 8;  // the scatter's ITEMS
  return (n + kBucketRows - 1) / kBucketRows + (uint64_t)grid * ((1ULL << pbits) + block) + 1;
}

hipError_t launch_part_bucket(int key_bytes, const void *keys, const Schema &s, uint64_t n, int pbits,
                              uint32_t tile_rows, void *okeys, const PartOut &o, const BucketBufs &bb, unsigned grid,
                              hipStream_t st) {
  const uint32_t ntiles = (uint32_t)((n + tile_rows - 1) / tile_rows);
  if (ntiles == 0) return hipSuccess;
  const bool one = s.ncols <= 1;
  const bool w8 = s.ncols >= 1 && dtype_bytes(s.col_dtype[0]) == 8;
#define TGA_SC_ARGS keys, s, n, pbits, tile_rows, ntiles, okeys, o, bb, grid, st
#define TGA_SC_CASE(K)                                                                              \
  if (one) return w8 ? launch_bucket_t<K, 1, 8>(TGA_SC_ARGS) : launch_bucket_t<K, 1, 4>(TGA_SC_ARGS); \
  return launch_bucket_t<K, kMaxCols, 0>(TGA_SC_ARGS);
  if (key_bytes == 4) { TGA_SC_CASE(uint32_t) }
  TGA_SC_CASE(uint64_t)
#undef TGA_SC_CASE
#undef TGA_SC_ARGS
}

static hipError_t lds_attr(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_bucket_lists(const BucketBufs &bb, uint32_t P, uint64_t chmax, const AggPlan &plan,
                               hipStream_t st) {
  const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(256, bb.cap / 4096));
  size_t lds = (size_t)P * 12;
  hipError_t e = lds_attr((const void *)&k_bucket_count, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_count, dim3(grid), dim3(1024), lds, st, bb.bpart, bb.bcnt, bb.pool, bb.cap, P, bb.pbk,
                     bb.prow);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  lds = (size_t)P * 8;
  if ((e = lds_attr((const void *)&k_bucket_plan, lds)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_plan, dim3(1), dim3(1024), lds, st, bb.pbk, bb.prow, P, chmax, bb.bcur, plan);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = lds_attr((const void *)&k_bucket_list, lds)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_bucket_list, dim3(grid), dim3(1024), lds, st, bb.bpart, bb.bcnt, bb.pool, bb.cap, P, bb.bcur,
                     bb.blist);
  return hipGetLastError();
}

uint64_t compact_tiles(uint64_t slots_plus_one) {
  const uint64_t TS = (uint64_t)kCompactThreads * kCompactItems;
  return (slots_plus_one + TS - 1) / TS;
}

hipError_t launch_compact(int key_bytes, const GTable &g, const Schema &s, const OutDesc &o, uint64_t *desc,
                          uint32_t *ticket, uint64_t *count_out, hipStream_t st, const uint64_t *base) {
  const uint64_t tiles = compact_tiles(g.S + 1);
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_compact<uint32_t>, dim3((unsigned)tiles), dim3(kCompactThreads), 0, st, g, s, o, desc, ticket,
                       count_out, base);
  else
    hipLaunchKernelGGL(k_compact<uint64_t>, dim3((unsigned)tiles), dim3(kCompactThreads), 0, st, g, s, o, desc, ticket,
                       count_out, base);
  return hipGetLastError();
}

hipError_t launch_sample_keys(int key_bytes, const void *keys, uint64_t n, uint64_t m, uint64_t *out,
                              hipStream_t st) {
  if (key_bytes == 4)
    hipLaunchKernelGGL(k_sample_keys<uint32_t>, dim3(grid_for(m, 256)), dim3(256), 0, st, (const uint32_t *)keys, n, m, out);
  else
    hipLaunchKernelGGL(k_sample_keys<uint64_t>, dim3(grid_for(m, 256)), dim3(256), 0, st, (const uint64_t *)keys, n, m, out);
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
    case GEN_NORMAL_F64: hipLaunchKernelGGL(k_gen_normal_f64, dim3(grid), dim3(256), 0, st, (double *)out, n, x, y, seed); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tga
