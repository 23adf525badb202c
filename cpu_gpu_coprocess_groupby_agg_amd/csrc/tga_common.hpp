// tga_common.hpp -- definitions shared by host and device code of the
// MI355X (gfx950) tiled hash group-by.
//
// Keys are 32- or 64-bit bit patterns compared for equality (the reference
// keys are u32, T/include/util.cuh:6).  Values come from up to kMaxCols
// columns; each aggregate becomes one accumulator ("acc") whose state lives in
// a structure-of-arrays hash table, first in LDS per workgroup, then in a
// global table in HBM.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TGA_HD __host__ __device__ __forceinline__

namespace tga {

constexpr int kMaxCols = 4;
constexpr int kMaxAccs = 8;
// rows per bucket of the partitioned copy (one wave step of the aggregation:
// 64 lanes x 4 rows)
constexpr int kBucketShift = 8;
constexpr uint32_t kBucketRows = 1u << kBucketShift;
constexpr int kMaxFx = 2;  // exact fp64 SUM accumulators per query

// Exact fp64 SUM: a double is split into 32-bit chunks of the fixed-point
// integer round(x * 2^kFxShift) and each chunk is added into an int64 limb.
// Limb i carries weight 2^(32 i - kFxShift).  Five limbs cover [2^-96, 2^64);
// bits below 2^-96 are truncated, |x| >= 2^64, Inf and NaN raise a range flag.
// Each add is < 2^32 in magnitude, so a limb is exact for < 2^31 rows.
constexpr int kLimbs = 5;
constexpr int kFxShift = 96;

enum Dtype : int32_t { DT_U32 = 1, DT_I32 = 2, DT_U64 = 3, DT_I64 = 4, DT_F64 = 5 };
enum AggOp : int32_t { OP_COUNT = 0, OP_SUM = 1, OP_MIN = 2, OP_MAX = 3 };
enum AccKind : int32_t {
  ACC_COUNT = 0,    // LDS u32, global u64
  ACC_SUM_I = 1,    // int64 (wraps mod 2^64)
  ACC_MIN_I32 = 2,  // int32 column only
  ACC_MAX_I32 = 3,
  ACC_MIN_I64 = 4,  // every other column through an order-preserving map to int64
  ACC_MAX_I64 = 5,
  ACC_SUM_F = 6,    // exact fixed-point limbs
};

// global-table status words
enum : int { FLAG_SENTINEL = 0, FLAG_OVERFLOW = 1, FLAG_RANGE = 2, kFlagWords = 3 };

struct Acc {
  int32_t kind;
  int32_t col;
  int32_t col_dtype;  // dtype of column `col` (copied here so kernels index no arrays)
  int32_t pad;
};

struct Schema {
  const void *cols[kMaxCols];
  int32_t col_dtype[kMaxCols];
  int32_t ncols;
  int32_t nacc;
  Acc acc[kMaxAccs];
  int32_t fx_slot[kMaxAccs];  // ACC_SUM_F accumulators: index among them
  int32_t nfx;                // number of ACC_SUM_F accumulators
};

// Global table in HBM: the LDS table's SoA shape with S regular slots.
// With S_dev set, the kernels use the slot count stored there (a power of
// two <= S, written on the GPU by the work plan) in place of S: the arrays
// are allocated for S and only the first *S_dev slots (+1) are used.
struct GTable {
  void *keys;               // [S+1]
  void *acc[kMaxAccs];      // [S+1] each (ACC_SUM_F: [kLimbs][S+1])
  uint64_t S;
  uint32_t *flags;          // kFlagWords status words
  const uint64_t *S_dev;    // null: S slots
};

// Dense output: keys plus one column per aggregate, `capacity` rows.
struct OutDesc {
  void *keys;
  void *cols[kMaxAccs];
  uint64_t capacity;
};

// Byte offsets of the SoA arrays of one table with `slots` entries plus one
// extra entry (index `slots`) reserved for the key whose bits equal the empty
// marker, so every key value is supported.
struct Layout {
  uint32_t slots;
  uint32_t key_off;
  uint32_t acc_off[kMaxAccs];
  uint32_t bytes;
};


// The accumulator program of one table in a few words: all the row loop of
// the aggregation kernels needs, so that loop keeps a handful of scalar
// registers live instead of the whole Schema / table descriptors.
enum : uint32_t { MAP_RAW = 0, MAP_U64 = 1, MAP_F64 = 2 };
struct AccProg {
  uint32_t nacc;
  uint32_t kinds;          // 4 bits per accumulator: AccKind
  uint32_t cols;           // 2 bits per accumulator: its value column
  uint32_t maps;           // 2 bits per accumulator: operand map (MAP_*)
  uint32_t ncols;
  uint32_t coldt;          // 4 bits per value column: Dtype
  uint32_t off[kMaxAccs];  // LDS byte offset of each accumulator array
};

TGA_HD int dtype_bytes(int dt) { return (dt == DT_U32 || dt == DT_I32) ? 4 : 8; }

TGA_HD int acc_lds_bytes(int kind) {
  switch (kind) {
    case ACC_COUNT: case ACC_MIN_I32: case ACC_MAX_I32: return 4;
    case ACC_SUM_F: return 8 * kLimbs;
    default: return 8;
  }
}
TGA_HD int acc_global_bytes(int kind) {
  switch (kind) {
    case ACC_MIN_I32: case ACC_MAX_I32: return 4;
    case ACC_SUM_F: return 8 * kLimbs;
    default: return 8;
  }
}

// SplitMix64 finalizer (Steele, Lea & Flood 2014): a bijection on 64 bits.
TGA_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// Row of the i-th of m sample keys over n rows: every row when m >= n,
// else rows drawn uniformly at random (with replacement, a fixed SplitMix64
// stream).  Evenly strided rows alias with periodic inputs: keys repeating
// every 2^20 rows, sampled every 2^14 rows, show only 64 keys.
TGA_HD uint64_t sample_row(uint64_t i, uint64_t m, uint64_t n) {
  if (m >= n) return i;
  return (uint64_t)(((unsigned __int128)mix64(0x6a09e667f3bcc909ULL + (i + 1) * 0x9e3779b97f4a7c15ULL) * n) >> 64);
}
// MurmurHash3 fmix32: a bijection on 32 bits.
TGA_HD uint32_t fmix32(uint32_t h) {
  h ^= h >> 16; h *= 0x85ebca6bu; h ^= h >> 13; h *= 0xc2b2ae35u; h ^= h >> 16;
  return h;
}

// Row hash.  Independent bit fields feed the three placements, so keys of
// one partition still spread over all slots of the tables (the reference uses
// one hash modulo both P and the capacity, T/src/kernel.cu:225,327):
//   partition  = top `pbits` bits
//   LDS slot   = bits 0..31 (multiply-shift onto [0, slots))
//   global slot= bits 16..47
TGA_HD uint64_t key_hash(uint64_t k) { return mix64(k ^ 0x5851f42d4c957f2dULL); }
TGA_HD uint32_t hash_partition(uint64_t h, int pbits) { return pbits ? (uint32_t)(h >> (64 - pbits)) : 0u; }
// Partition of a key in the plain operator's partitioned path (pbits <= 32):
// the top bits of k * 2^64/phi mod 2^64 (Fibonacci hashing, Knuth TAOCP
// vol. 3, 6.4).  Only the high product word is formed: three 32-bit
// multiplies instead of key_hash's two 64-bit ones, and a function unrelated
// to the key_hash bits that place a key inside its partition's tables.
TGA_HD uint32_t key_partition(uint64_t k, int pbits) {
  return pbits ? (uint32_t)((k * 0x9e3779b97f4a7c15ULL) >> 32) >> (32 - pbits) : 0u;
}
// LDS-slot hash of the integer row loop: the high word of k * M mod 2^64,
// M odd (three 32-bit multiplies where key_hash needs two 64-bit ones and
// the shifts between them).  M is unrelated to key_partition's multiplier,
// so the keys of one partition still spread over the whole table; every
// insert into and lookup in one LDS table uses the same hash.
TGA_HD uint32_t key_slot_hash(uint64_t k) { return (uint32_t)((k * 0xd6e8feb86659fd93ULL) >> 32); }
TGA_HD uint32_t hash_lds_slot(uint64_t h, uint32_t slots) {
  return (uint32_t)(((uint64_t)(uint32_t)h * slots) >> 32);
}
TGA_HD uint64_t hash_global_slot(uint64_t h, uint64_t slots) {
  return ((h >> 16) & 0xffffffffULL) * slots >> 32;
}

// Order-preserving maps onto int64 for MIN/MAX.
TGA_HD int64_t f64_bits_to_ordered(uint64_t b) { return (int64_t)(b ^ ((uint64_t)((int64_t)b >> 63) & 0x7fffffffffffffffULL)); }
TGA_HD uint64_t ordered_to_f64_bits(int64_t o) { return (uint64_t)o ^ ((uint64_t)(o >> 63) & 0x7fffffffffffffffULL); }

// Builds the accumulator program of a schema over one table layout.
TGA_HD AccProg make_acc_prog(const Schema &s, const Layout &L) {
  AccProg p{};
  p.nacc = (uint32_t)s.nacc;
  p.ncols = (uint32_t)s.ncols;
  for (int c = 0; c < s.ncols && c < kMaxCols; ++c) p.coldt |= (uint32_t)s.col_dtype[c] << (4 * c);
  for (int a = 0; a < s.nacc && a < kMaxAccs; ++a) {
    const Acc &ac = s.acc[a];
    p.kinds |= (uint32_t)ac.kind << (4 * a);
    p.cols |= (uint32_t)(ac.col & 3) << (2 * a);
    uint32_t m = MAP_RAW;
    if (ac.kind == ACC_MIN_I64 || ac.kind == ACC_MAX_I64) m = ac.col_dtype == DT_U64 ? MAP_U64 : ac.col_dtype == DT_F64 ? MAP_F64 : MAP_RAW;
    p.maps |= m << (2 * a);
    p.off[a] = L.acc_off[a];
  }
  return p;
}

// Splits a double into at most three signed 32-bit chunks of its fixed-point
// image; returns the first limb index, or -1 when nothing is added.  Sets
// *range when x cannot be represented (|x| >= 2^64, Inf, NaN).
TGA_HD int fx_split(uint64_t b, int64_t c[3], int *range) {
  uint32_t e = (uint32_t)(b >> 52) & 0x7ffu;
  uint64_t m = b & ((1ULL << 52) - 1);
  if (e == 0x7ffu) { *range = 1; return -1; }
  int ex;
  if (e == 0) { if (m == 0) return -1; ex = -1074; }
  else { m |= 1ULL << 52; ex = (int)e - 1075; }
  int sh = ex + kFxShift;  // bit position of m's bit 0 in the fixed-point integer
  if (sh < 0) {
    if (sh <= -53) return -1;  // entirely below the window
    m >>= -sh;
    sh = 0;
    if (m == 0) return -1;
  }
  if (sh + 53 > 32 * kLimbs) { *range = 1; return -1; }
  int i0 = sh >> 5, lo = sh & 31;
  uint32_t c0, c1, c2;
  if (lo == 0) { c0 = (uint32_t)m; c1 = (uint32_t)(m >> 32); c2 = 0; }
  else {
    c0 = (uint32_t)(m << lo);
    c1 = (uint32_t)(m >> (32 - lo));
    c2 = lo > 11 ? (uint32_t)(m >> (64 - lo)) : 0u;
  }
  bool neg = (b >> 63) != 0;
  c[0] = neg ? -(int64_t)c0 : (int64_t)c0;
  c[1] = neg ? -(int64_t)c1 : (int64_t)c1;
  c[2] = neg ? -(int64_t)c2 : (int64_t)c2;
  return i0;
}

TGA_HD int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// Exact value of sum_i L[i] * 2^(32 i - kFxShift), rounded to nearest even.
// The limbs are combined into a 192-bit two's-complement integer (hi:lo),
// whose magnitude is rounded once; only fixed-index operations, so the
// device code keeps everything in registers.
TGA_HD double fx_to_double(const int64_t *L) {
  static_assert(kLimbs == 5, "fx_to_double combines exactly five limbs");
  const __int128 t = (__int128)L[0] + ((__int128)L[1] << 32);
  __int128 u = (__int128)L[2] + ((__int128)L[3] << 32) + ((__int128)L[4] << 64);
  u += t >> 64;                 // V = u * 2^64 + lo
  uint64_t lo = (uint64_t)t;
  const bool neg = u < 0;
  unsigned __int128 hi;
  if (neg) {                    // magnitude of the 192-bit value
    lo = ~lo + 1;
    hi = ~(unsigned __int128)u + (lo == 0 ? 1 : 0);
  } else {
    hi = (unsigned __int128)u;
  }
  if (hi == 0 && lo == 0) return 0.0;
  int hb;                       // highest set bit of hi:lo
  uint64_t mant;                // the 64 bits from hb down
  bool sticky;                  // any bit below them
  if (hi != 0) {
    const uint64_t h1 = (uint64_t)(hi >> 64), h0 = (uint64_t)hi;
    hb = h1 ? 191 - clz64(h1) : 127 - clz64(h0);
    const int sh = hb - 63;     // 1 .. 128
    if (sh >= 64) {
      const int r = sh - 64;    // 0 .. 64
      mant = (uint64_t)(hi >> r);
      const unsigned __int128 mask = r ? ((((unsigned __int128)1) << r) - 1) : 0;
      sticky = lo != 0 || (hi & mask) != 0;
    } else {
      mant = (uint64_t)((hi << (64 - sh)) | (unsigned __int128)(lo >> sh));
      sticky = (lo & ((1ULL << sh) - 1)) != 0;
    }
  } else {
    hb = 63 - clz64(lo);
    mant = lo << (63 - hb);
    sticky = false;
  }
  uint64_t q = mant >> 11;
  const uint64_t rem = mant & 0x7ff;
  if (rem > 0x400 || (rem == 0x400 && (sticky || (q & 1)))) {
    ++q;
    if (q == (1ULL << 53)) { q >>= 1; ++hb; }
  }
  // value = q * 2^(hb - 52 - kFxShift), a normal double for every window value
  int ex = hb - 52 - kFxShift;
  double r = (double)q;
  while (ex > 0) { const int k = ex > 60 ? 60 : ex; r *= (double)(1ULL << k); ex -= k; }
  while (ex < 0) { const int k = -ex > 60 ? 60 : -ex; r /= (double)(1ULL << k); ex += k; }
  return neg ? -r : r;
}

template <typename K> struct KeyTraits;
template <> struct KeyTraits<uint32_t> { static constexpr uint32_t kEmpty = 0xffffffffu; };
template <> struct KeyTraits<uint64_t> { static constexpr uint64_t kEmpty = ~0ULL; };

}  // namespace tga
