// tga_internal.hpp -- host-side launch interface between the C ABI layer and
// the kernels (each kernel is launched from the translation unit defining it).
#pragma once

#include <cmath>

#include "tga_common.hpp"

namespace tga {

// Work plan of k_agg: segment p covers rows [pstart[p], pstart[p+1]) and is
// cut into cstart[p+1]-cstart[p] slices; ticket hands out slice ids.  With a
// bucket list (blist != null) pstart indexes blist instead: segment p is the
// buckets blist[pstart[p] .. pstart[p+1]), each entry rows << 32 | id for the
// bucket holding `rows` rows at rows [id * kBucketRows, ...).  blist2 (set
// with blist for k_agg) holds a second bucket of the same block, entry 0 for
// none: two partly filled buckets whose rows fit one block are one block.
// An entry of 0 rows ends the segment (paired lists end in such entries).
struct AggPlan {
  uint64_t *pstart;  // [P+1]
  uint32_t *cstart;  // [P+1]
  uint32_t P;
  uint32_t *ticket;
  uint64_t *split_rows;  // rows of partitions cut into more than one item
  const uint64_t *blist;
  const uint64_t *blist2;
  // Global-table sizing by the plan kernel (gslots null: none).  Only the
  // partitions cut into several items aggregate through the global table,
  // so it takes the power of two >= min(2 * their rows, gs_per_split * their
  // count), at least 4096 slots and at most gs_cap.
  uint64_t *gslots;
  uint64_t gs_cap;
  uint32_t gs_per_split;
};

struct PartOut {
  void *cols[kMaxCols];
};

enum GenKind : int { GEN_KEYS_U32 = 1, GEN_KEYS_U64 = 2, GEN_VALS_I32 = 3, GEN_ZIPF_U64 = 4, GEN_NORMAL_F64 = 5,
                     GEN_REF_U32 = 6, GEN_MOD_U32 = 7, GEN_MOD_U64 = 8, GEN_SEQ_I32 = 9 };

hipError_t launch_fill_u64(uint64_t *p, uint64_t n, uint64_t v, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t st);
hipError_t launch_copy16(const void *src, void *dst, uint64_t n16, hipStream_t st);
// Raises a kernel's dynamic-LDS limit to `bytes` (past 64 KB) on the current
// device; the attribute is set once per kernel and size.
hipError_t raise_lds_limit(const void *fn, size_t bytes);

// Up to two counter arrays the table's init launch also sets: p[k][0, n[k])
// = v[k] (p[k] null: none).
struct WordInit {
  uint64_t *p[2];
  uint64_t v[2];
  uint32_t n[2];
};
hipError_t launch_gtable_init(int key_bytes, const GTable &g, const Schema &s, hipStream_t st,
                              const WordInit &wi = WordInit{});
// emit != 0: work items that hold a whole partition write their groups
// straight to `od` (rows reserved on *emit_count); the others flush into g.
hipError_t launch_agg(int key_bytes, const void *keys, const Schema &s, const AggPlan &plan, const Layout &L,
                      const GTable &g, int nbits, int max_probe, unsigned grid, unsigned block, hipStream_t st,
                      int emit = 0, const OutDesc *od = nullptr, uint64_t *emit_count = nullptr, int fold = 0,
                      uint32_t *ovf = nullptr, int rec = 0);
// rec: `keys` holds records, not keys and s.cols -- 1: (key, 4-byte value)
// records (a partitioned copy of a one-column schema with a 4-byte value
// column); 2..kMaxRecValueDwords: the key followed by that many value words,
// the columns in order at their own widths (record_value_dwords(s))
bool copy_as_records(const Schema &s);
constexpr int kMaxRecValueDwords = 4;
// value words of a record of this schema's columns, 0 if they take more
// than kMaxRecValueDwords (or there are none)
int record_value_dwords(const Schema &s);
hipError_t launch_plan_direct(const AggPlan &plan, uint64_t n, uint32_t nchunks, hipStream_t st);
uint64_t scan_tiles(uint64_t n);
hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *desc, uint32_t *ticket,
                           hipStream_t st);
size_t scatter_lds_bytes(const Schema &s, int key_bytes, int pbits, uint32_t tile_rows);

// Buckets of the count-free partitioning: bpart/bcnt [cap] u32, blist and
// blist2 [cap] u64 entries; pool (the allocation counter), pbk [P] u32, prow
// [P] u64 and bcur [P] u32 start zeroed; bend [P] u32 (the back cursor of
// the partly filled buckets) is set by the list plan.
struct BucketBufs {
  uint32_t *bpart, *bcnt, *pool, *pbk, *bcur, *bend;
  uint64_t *blist, *blist2, *prow;
  uint32_t cap;
};
// persistent workgroups of the scatter for ntiles tiles (at most max_grid)
unsigned bucket_grid(uint32_t ntiles, unsigned max_grid);
// workgroup size of the scatter (wc: the write-combining one)
unsigned scatter_block(uint32_t tile_rows, int wc);
// buckets the scatter can take: ceil(n / kBucketRows) + grid * (2^pbits +
// its workgroup size)
uint64_t bucket_cap(uint64_t n, unsigned grid, int pbits, unsigned block);
// Two-pass partitioning (fanouts above the write-combining scatter's 2^9):
// the first pass's bucket lists (pstart: list offsets of its P1 partitions)
// are cut into work items of items_per list entries (launch_pass2_plan), and
// launch_part_scatter2 splits each into 2^pbits sub-partitions by hash bits
// below the first pass's, writing records into the buckets of bb (8192-row
// tiles; headers record the partition p1 * 2^pbits + sub-partition).
hipError_t launch_pass2_plan(const uint64_t *pstart, uint32_t P1, uint32_t items_per, uint32_t *istart,
                             uint32_t *ticket, hipStream_t st);
hipError_t launch_part_scatter2(int key_bytes, int vd, const void *rec, const uint64_t *list, const uint64_t *pstart,
                                const uint32_t *istart, uint32_t *ticket, uint32_t P1, int hash_bits,
                                uint32_t items_per, int pbits, void *orec, const BucketBufs &bb, unsigned grid,
                                hipStream_t st);
// tile rows of launch_part_scatter2 for records of vd value words
uint32_t scatter2_tile(int vd);
// LDS of the write-combining scatter over records of vd value words, and its
// largest tile (4096, 6144 or 8192 rows; 4096 only for vd > 1) whose LDS fits
// lds_max, 0 if none does
size_t scatter_wc_lds_bytes(int key_bytes, int vd, int pbits, uint32_t tile_rows);
uint32_t scatter_wc_tile(int key_bytes, int vd, int pbits, size_t lds_max);
// rows [0, n) into the partitioned copy okeys / o.cols: into buckets with
// bb, else partition-contiguous at pbase[p] + offs[tile][p] (from
// launch_part_count + launch_colscan over the same tiles).  wc: buckets of
// (key, value) records written as whole cache lines (tile_rows from
// scatter_wc_tile)
hipError_t launch_part_scatter(int key_bytes, const void *keys, const Schema &s, uint64_t n, int pbits,
                               uint32_t tile_rows, void *okeys, const PartOut &o, const BucketBufs *bb,
                               const uint32_t *offs, const uint32_t *pbase, unsigned grid, hipStream_t st,
                               int wc = 0);
// Partition-contiguous layout: per-tile histogram cnt[ntiles][P]; its
// column-wise scan offs (same shape, exclusive per partition), total[P] and
// pbase[P+1] (partition starts); desc: colscan_desc_words() zeroed words
// followed by a zeroed ticket word; work plan over pbase.
hipError_t launch_part_count(int key_bytes, const void *keys, uint64_t n, int pbits, uint32_t tile_rows,
                             uint32_t ntiles, uint32_t *cnt, hipStream_t st);
uint64_t colscan_desc_words(uint32_t ntiles, uint32_t P);
hipError_t launch_colscan(const uint32_t *cnt, uint32_t ntiles, uint32_t P, uint64_t *desc, uint32_t *ticket,
                          uint32_t *offs, uint32_t *total, uint32_t *pbase, hipStream_t st);
hipError_t launch_plan_chunks(const uint32_t *pbase, uint32_t P, uint64_t n, uint64_t chmax, const AggPlan &plan,
                              hipStream_t st);
// bucket headers -> per-partition bucket lists and the k_agg work plan
// (items of at most about chmax rows); plan.blist/bcnt must point at bb's.
// for_agg: the lists feed k_agg (plan.blist2 = bb.blist2 is written); with
// pair, partly filled buckets of a partition are paired into blocks of at
// most kBucketRows rows (k_agg reads such blocks for 8-byte keys only).
hipError_t launch_bucket_lists(const BucketBufs &bb, uint32_t P, uint64_t chmax, const AggPlan &plan,
                               hipStream_t st, bool for_agg, bool pair);
// The k_agg part alone, over lists launch_bucket_lists made for another use.
hipError_t launch_bucket_pairs(const BucketBufs &bb, uint32_t P, const AggPlan &plan, hipStream_t st, bool pair);
uint64_t compact_tiles(uint64_t slots_plus_one);
// Output rows start at *base when base is non-null (after directly emitted groups).
// adds the group count of g (occupied slots + the empty-key slot) to *out
hipError_t launch_count_groups(int key_bytes, const GTable &g, uint64_t *out, hipStream_t st);
// Exact distinct count of the keys whose partition hash has its top
// kSliceBits bits 0 (out[0]; out[1] != 0: the set overflowed, a lower bound).
// set: slice_set_slots(n) keys of workspace (filled here).
constexpr int kSliceBits = 10;
uint32_t slice_set_slots(uint64_t n);
hipError_t launch_slice_distinct(int key_bytes, const void *keys, uint64_t n, void *set, uint32_t slots, uint64_t *out,
                                 unsigned grid, hipStream_t st);
// The same count over the records (rd words, key first) of partition 0 of a
// bucket-layout copy of fanout <= 2^kSliceBits (list entries [pstart[0],
// pstart[1]) of `list`): the slice lies inside that partition.
hipError_t launch_slice_records(int key_bytes, const void *rec, uint32_t rd, const uint64_t *list,
                                const uint64_t *pstart, void *set, uint32_t slots, uint64_t *out, unsigned grid,
                                hipStream_t st);
hipError_t launch_compact(int key_bytes, const GTable &g, const Schema &s, const OutDesc &o, uint64_t *desc,
                          uint32_t *ticket, uint64_t *count_out, hipStream_t st, const uint64_t *base = nullptr);
// out[0..3] = {distinct keys, keys seen once, largest count of one key,
// overflow} of the m <= 16384 keys sampled at random rows (all rows when m >= n; gathered into
// smp[m], then counted in LDS hash tables by 16 workgroups, each taking a
// sixteenth of the key hash; overflow 1: the counts are incomplete, smp
// still holds the sample)
hipError_t launch_sample_stats(int key_bytes, const void *keys, uint64_t n, uint32_t m, uint64_t *smp, uint64_t *out,
                               hipStream_t st);
// Copies n[i] 4-byte words from src[i] (i < count <= 8) one after another to
// dst (one workgroup), so a call's small results come back in one copy.
constexpr int kStatusWords = 64;
struct WordGather {
  const uint32_t *src[8];
  uint32_t n[8];
  uint32_t count;
};
hipError_t launch_gather_words(const WordGather &wg, uint32_t *dst, hipStream_t st);
// m keys at random rows (all rows when m >= n) into out[m], widened to 64
// bits; stats (optional): four counters zeroed for launch_sample_stats
hipError_t launch_sample_keys(int key_bytes, const void *keys, uint64_t n, uint64_t m, uint64_t *out,
                              hipStream_t st, uint64_t *stats = nullptr);
hipError_t launch_gen(int kind, void *out, uint64_t n, uint64_t a, uint64_t b, double x, double y, uint64_t seed,
                      uint64_t salt, hipStream_t st);

}  // namespace tga
