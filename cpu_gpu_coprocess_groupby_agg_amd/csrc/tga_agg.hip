// tga_agg.hip -- k_agg, the LDS aggregation of the plain operator (direct
// path: row ranges of the input; partitioned path: the work items of the
// partitioned copy), and its launcher.  Its many variants (key width, fp64
// SUM, value columns and their widths, records, wave folding, fixed
// accumulator programs) are compiled in their own translation unit.
//
// Reference counterparts (T/ = tiled_groupby_agg_partition/):
//   K1-K5 pre-aggregation      T/src/kernel.cu:9-186     -> k_agg (direct)
//   K8/K9 intra-partition agg   T/src/kernel.cu:294-378  -> k_agg (partitions)
#include "tga_device.hpp"
#include "tga_internal.hpp"

namespace tga {

// ---------------------------------------------------------------------------
// (a)/(c) LDS aggregation
// ---------------------------------------------------------------------------

// Segment of work item `item`: the largest p with cstart[p] <= item (cstart
// is the exclusive prefix of the items per segment).  While no segment
// before it was split or empty, item i is segment i: two independent loads
// settle that, where the search takes log2(P) dependent ones.
__device__ __forceinline__ uint32_t plan_segment(const AggPlan &plan, uint32_t item) {
  if (item < plan.P && plan.cstart[item] == item && plan.cstart[item + 1] == item + 1) return item;
  uint32_t lo = 0, hi = plan.P;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (plan.cstart[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

// Persistent: workgroups pull work items from plan.ticket until none remain.
// Item -> (segment p, slice j of nch) through plan.cstart; rows of segment p
// are [pstart[p], pstart[p+1]).  Thread 0 takes the next item's ticket and
// finds its segment while the other waves finish the current item's rows.
template <typename K, bool FX, int NC, int VW, bool FOLD, uint32_t SIG>
__global__ void __launch_bounds__(1024) k_agg(const K *__restrict__ keys, Schema s, AggPlan plan, Layout L,
                                              GTable g, AccProg ap, int nbits, int max_probe, int emit, OutDesc od,
                                              uint64_t *emit_count, uint32_t *ovf) {
  extern __shared__ __align__(16) char smem[];
  __shared__ uint32_t s_item, s_seg, s_sentinel;
  __shared__ uint32_t scratch[1024 / kWave + 1];
  __shared__ uint64_t s_base;
  // Descriptors read outside the row loop (table init, flush, emit, the
  // out-of-line global-table path) are read from LDS copies: taken from the
  // kernel arguments, they would stay in scalar registers across the row
  // loop, which then spills them into vector registers.
  __shared__ GTable s_g;
  __shared__ Schema s_s;
  __shared__ OutDesc s_od;
  __shared__ uint32_t s_abort[2];  // table full / rows consumed (abort mode)
  if (threadIdx.x == 0) {
    if (g.S_dev) g.S = *g.S_dev;
    s_g = g;
    s_s = s;
    s_od = od;
  }
  LdsTable<K> t{smem, L, &s_sentinel};
  const uint32_t total = plan.cstart[plan.P];
  auto take_item = [&] {  // thread 0
    const uint32_t it = atomicAdd(plan.ticket, 1u);
    s_item = it;
    s_seg = it < total ? plan_segment(plan, it) : 0u;
  };
  if (threadIdx.x == 0) take_item();
  for (;;) {
    __syncthreads();
    const uint32_t item = wave_uniform(s_item);
    if (item >= total) break;
    const uint32_t p = wave_uniform(s_seg);
    const uint32_t j = item - plan.cstart[p], nch = plan.cstart[p + 1] - plan.cstart[p];
    const uint64_t sb = plan.pstart[p], se = plan.pstart[p + 1], len = se - sb;
    const uint64_t b = sb + len * j / nch, e = sb + len * (j + 1) / nch;
    // with ovf set, a work item whose groups overflow the LDS table is
    // abandoned and recorded in ovf = {count, capacity, (rows consumed, rows)...}
    // instead of spilling: the host then plans a finer partitioning
    const bool whole = emit && nch == 1;
    const bool abandon = ovf != nullptr;
    lds_init(t, s_s);
    if (threadIdx.x == 0) {
      s_abort[0] = 0;
      s_abort[1] = 0;
    }
    __syncthreads();
    const BucketList bl{plan.blist, plan.blist2};
    {
      const void *cols[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) cols[c] = s.cols[c];
      agg_rows_int<K, NC, VW, FX ? false : FOLD, SIG, FX>(keys, cols, b, e, t, ap, nbits, max_probe, &s_g, &s_s,
                                                          abandon ? s_abort : nullptr, nullptr, bl);
    }
    // every wave has read s_item / s_seg (before the barrier after lds_init)
    if (threadIdx.x == 0) take_item();
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
      lds_emit<K>(t, s_s, s_od, emit_count, s_g.flags, scratch, &s_base);
    } else {
      lds_flush<K>(t, s_s, s_g);
    }
    // the barrier at the top of the loop keeps the next lds_init behind
    // every read of this table
  }
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------

template <typename K>
static hipError_t set_lds_limit(const void *fn, size_t bytes) {
  (void)sizeof(K);
  return raise_lds_limit(fn, bytes);
}

template <typename K, bool FX, int NC, int VW, bool FOLD, uint32_t SIG = 0>
static hipError_t launch_agg_t(const void *keys, const Schema &s, const AggPlan &plan, const Layout &L, const GTable &g,
                               int nbits, int max_probe, unsigned grid, unsigned block, hipStream_t st, int emit,
                               const OutDesc &o, uint64_t *emit_count, uint32_t *ovf) {
  const size_t lds = L.bytes;
  hipError_t e = set_lds_limit<K>((const void *)&k_agg<K, FX, NC, VW, FOLD, SIG>, lds);
  if (e != hipSuccess) return e;
  const AccProg ap = make_acc_prog(s, L);
  hipLaunchKernelGGL((k_agg<K, FX, NC, VW, FOLD, SIG>), dim3(grid), dim3(block), lds, st, (const K *)keys, s, plan, L, g,
                     ap, nbits, max_probe, emit, o, emit_count, ovf);
  return hipGetLastError();
}

// The compile-time accumulator signature a schema matches (see prog_kind),
// or 0: one 4-byte value column, the accumulators in one of the fixed orders.
static uint32_t fixed_signature(const Schema &s) {
  if (s.nfx > 0 || s.ncols != 1 || dtype_bytes(s.col_dtype[0]) != 4 || s.nacc > 4) return 0;
  uint32_t sig = 0;
  for (int a = 0; a < s.nacc; ++a) {
    if (s.acc[a].kind != ACC_COUNT && s.acc[a].col != 0) return 0;
    sig |= (uint32_t)(s.acc[a].kind + 1) << (4 * a);
  }
  return sig == kSigCountSum || sig == kSigCountSumMinMax ? sig : 0;
}

// One kernel per (key width, fp64 SUM present, value columns, wave folding):
// each variant gets its own register allocation, and a single value column
// of a known width is read without branches.  fp64 SUM accumulators never
// fold: a row adds at most three 32-bit chunks to its limbs with return-less
// LDS atomics (folding the limb vectors across a wave -- ten 64-bit shuffles
// per duplicate lane -- cost ~6000 VALU instructions per 64 rows under
// Zipf(1.1) keys).
hipError_t launch_agg(int key_bytes, const void *keys, const Schema &s, const AggPlan &plan, const Layout &L,
                      const GTable &g, int nbits, int max_probe, unsigned grid, unsigned block, hipStream_t st,
                      int emit, const OutDesc *od, uint64_t *emit_count, int fold, uint32_t *ovf, int rec) {
  OutDesc o{};
  if (od) o = *od;
  const bool fx = s.nfx > 0;
  const bool one = s.ncols <= 1;
  const bool w8 = s.ncols >= 1 && dtype_bytes(s.col_dtype[0]) == 8;
  const uint32_t sig = fixed_signature(s);
#define TGA_AGG_ARGS keys, s, plan, L, g, nbits, max_probe, grid, block, st, emit, o, emit_count, ovf
#define TGA_AGG_CASE(K)                                                                         \
  if (fx) return launch_agg_t<K, true, kMaxCols, 0, true>(TGA_AGG_ARGS);                        \
  if (!fold && sig == kSigCountSumMinMax)                                                       \
    return launch_agg_t<K, false, 1, 4, false, kSigCountSumMinMax>(TGA_AGG_ARGS);               \
  if (!fold && sig == kSigCountSum) return launch_agg_t<K, false, 1, 4, false, kSigCountSum>(TGA_AGG_ARGS); \
  if (one) {                                                                                    \
    if (w8) return fold ? launch_agg_t<K, false, 1, 8, true>(TGA_AGG_ARGS)                      \
                        : launch_agg_t<K, false, 1, 8, false>(TGA_AGG_ARGS);                    \
    return fold ? launch_agg_t<K, false, 1, 4, true>(TGA_AGG_ARGS)                              \
                : launch_agg_t<K, false, 1, 4, false>(TGA_AGG_ARGS);                            \
  }                                                                                             \
  return fold ? launch_agg_t<K, false, kMaxCols, 0, true>(TGA_AGG_ARGS)                         \
              : launch_agg_t<K, false, kMaxCols, 0, false>(TGA_AGG_ARGS);
#define TGA_AGG_REC(K)                                                                          \
  if (fold) return launch_agg_t<K, false, 1, kVwRec4, true>(TGA_AGG_ARGS);                      \
  if (sig == kSigCountSumMinMax) return launch_agg_t<K, false, 1, kVwRec4, false, kSigCountSumMinMax>(TGA_AGG_ARGS); \
  if (sig == kSigCountSum) return launch_agg_t<K, false, 1, kVwRec4, false, kSigCountSum>(TGA_AGG_ARGS); \
  return launch_agg_t<K, false, 1, kVwRec4, false>(TGA_AGG_ARGS);
// records of VD value words hold at most VD columns: one cell per word
#define TGA_AGG_RECW(K, VD)                                                                     \
  return fx ? launch_agg_t<K, true, VD, kVwRecBase + VD, false>(TGA_AGG_ARGS)                    \
            : launch_agg_t<K, false, VD, kVwRecBase + VD, false>(TGA_AGG_ARGS);
#define TGA_AGG_RECW_VD(K)                                                                      \
  switch (rec) {                                                                                \
    case 2: TGA_AGG_RECW(K, 2)                                                                  \
    case 3: TGA_AGG_RECW(K, 3)                                                                  \
    case 4: TGA_AGG_RECW(K, 4)                                                                  \
    default: return hipErrorInvalidValue;                                                       \
  }
  if (rec == 1) {  // keys = (key, 4-byte value) records
    if (fx || !one || w8) return hipErrorInvalidValue;
    if (key_bytes == 4) { TGA_AGG_REC(uint32_t) }
    TGA_AGG_REC(uint64_t)
  }
  if (rec > 1) {  // keys = records of `rec` value words (no wave folding)
    if (rec != record_value_dwords(s)) return hipErrorInvalidValue;
    if (key_bytes == 4) { TGA_AGG_RECW_VD(uint32_t) }
    TGA_AGG_RECW_VD(uint64_t)
  }
#undef TGA_AGG_RECW_VD
#undef TGA_AGG_RECW
#undef TGA_AGG_REC
  if (key_bytes == 4) { TGA_AGG_CASE(uint32_t) }
  TGA_AGG_CASE(uint64_t)
#undef TGA_AGG_CASE
#undef TGA_AGG_ARGS
}

}  // namespace tga
