// tga_device.hpp -- device building blocks of the gfx950 group-by kernels:
// wave64 match folding, the LDS open-addressing table, the global table in
// HBM, and the decoupled look-back prefix used by every scan/compaction.
#pragma once

#include "tga_common.hpp"

namespace tga {

constexpr int kWave = 64;
constexpr uint64_t kMaxGlobalProbe = 256;

// A value every work-item read from the same LDS word after a barrier, as a
// wave-uniform scalar.  Work-queue loops exit on such a value; read from LDS
// it is a per-lane register, and a loop with a per-lane exit may be compiled
// so that a wave re-enters the loop's barrier while some of its lanes are
// still parked at the exit (e.g. lane 0 before it fetches the next ticket).
// A scalar exit condition keeps every lane of the wave on one path.
__device__ __forceinline__ uint32_t wave_uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t wave_uniform64(uint64_t v) {
  return (uint64_t)wave_uniform((uint32_t)v) | (uint64_t)wave_uniform((uint32_t)(v >> 32)) << 32;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = (int)__lane_id();
  return l == 0 ? 0ULL : (~0ULL >> (64 - l));
}

// Mask of the active lanes whose `v` equals this lane's `v`, from one ballot
// per bit (v < 2^nbits).  HIP's __match_any is a serial loop over lanes.
__device__ __forceinline__ uint64_t match_mask(uint32_t v, int nbits, bool active) {
  uint64_t m = __ballot(active);
  for (int b = 0; b < nbits; ++b) {
    const bool bit = (v >> b) & 1u;
    const uint64_t bal = __ballot(active && bit);
    m &= bit ? bal : ~bal;
  }
  return active ? m : 0ULL;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, kWave);
  const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) {
    if constexpr (sizeof(T) == 8) v += (T)shfl64((uint64_t)v, (int)(__lane_id() ^ o));
    else v += (T)__shfl((int)v, (int)(__lane_id() ^ o), kWave);
  }
  return v;
}

// ---------------------------------------------------------------------------
// column access
// ---------------------------------------------------------------------------

// Raw 64-bit image of one cell: 32-bit columns are sign/zero extended.
__device__ __forceinline__ uint64_t load_cell(const Schema &s, int c, int64_t row) {
  const void *p = s.cols[c];
  switch (s.col_dtype[c]) {
    case DT_U32: return (uint64_t)((const uint32_t *)p)[row];
    case DT_I32: return (uint64_t)(int64_t)((const int32_t *)p)[row];
    default: return ((const uint64_t *)p)[row];
  }
}

// Operand an accumulator folds in for one row.
__device__ __forceinline__ uint64_t acc_operand(const Schema &s, int a, const uint64_t *cell) {
  const Acc ac = s.acc[a];
  if (ac.kind == ACC_COUNT) return 1;
  // fixed-index select keeps `cell` in registers
  const uint64_t raw = ac.col == 0 ? cell[0] : ac.col == 1 ? cell[1] : ac.col == 2 ? cell[2] : cell[3];
  if (ac.kind == ACC_MIN_I64 || ac.kind == ACC_MAX_I64) {
    const int dt = ac.col_dtype;
    if (dt == DT_U64) return raw ^ (1ULL << 63);
    if (dt == DT_F64) return (uint64_t)f64_bits_to_ordered(raw);
  }
  return raw;
}

// ---------------------------------------------------------------------------
// identity values and combine rules
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t acc_identity(int kind) {
  switch (kind) {
    case ACC_MIN_I32: return (uint64_t)(uint32_t)INT32_MAX;
    case ACC_MAX_I32: return (uint64_t)(uint32_t)INT32_MIN;
    case ACC_MIN_I64: return (uint64_t)INT64_MAX;
    case ACC_MAX_I64: return (uint64_t)INT64_MIN;
    default: return 0;
  }
}

// Fold operand b into a (registers), for every kind except ACC_SUM_F.
__device__ __forceinline__ uint64_t acc_fold(int kind, uint64_t a, uint64_t b) {
  switch (kind) {
    case ACC_COUNT: case ACC_SUM_I: return a + b;
    case ACC_MIN_I32: return (uint64_t)(int64_t)min((int32_t)a, (int32_t)b);
    case ACC_MAX_I32: return (uint64_t)(int64_t)max((int32_t)a, (int32_t)b);
    case ACC_MIN_I64: return (uint64_t)min((int64_t)a, (int64_t)b);
    case ACC_MAX_I64: return (uint64_t)max((int64_t)a, (int64_t)b);
    default: return a;
  }
}

// ---------------------------------------------------------------------------
// LDS table: keys[slots+2] plus one SoA array per accumulator (fp64 SUM limbs
// at a stride of slots+1); index `slots` holds the key equal to the empty
// marker, index slots+1 is a discard entry for rows that have no slot.
// ---------------------------------------------------------------------------

template <typename K>
struct LdsTable {
  char *base;
  Layout L;
  uint32_t *sentinel;  // nonzero once the empty-marker key was inserted

  __device__ __forceinline__ K *keys() const { return (K *)(base + L.key_off); }
  template <typename T>
  __device__ __forceinline__ T *arr(int a) const { return (T *)(base + L.acc_off[a]); }
};

template <typename K>
__device__ __forceinline__ void lds_init(const LdsTable<K> &t, const Schema &s) {
  const uint32_t n = t.L.slots + 1;
  K *k = t.keys();
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) k[i] = KeyTraits<K>::kEmpty;
  #pragma unroll
  for (int a = 0; a < kMaxAccs; ++a) {
    if (a >= s.nacc) break;
    const int kind = s.acc[a].kind;
    if (kind == ACC_SUM_F) {
      int64_t *p = t.template arr<int64_t>(a);
      for (uint32_t i = threadIdx.x; i < n * kLimbs; i += blockDim.x) p[i] = 0;
    } else if (acc_lds_bytes(kind) == 4) {
      uint32_t *p = t.template arr<uint32_t>(a);
      const uint32_t id = (uint32_t)acc_identity(kind);
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = id;
    } else {
      uint64_t *p = t.template arr<uint64_t>(a);
      const uint64_t id = acc_identity(kind);
      for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = id;
    }
  }
  if (threadIdx.x == 0) *t.sentinel = 0;
}

// Slot of `key`, inserting it if absent; -1 when `max_probe` buckets were
// full.  The table is probed in buckets of 4 slots read with one 16-byte
// (u32 keys) or two 16-byte (u64 keys) LDS loads, so a lane needs about one
// LDS round trip per lookup even at 3/4 load (one slot at a time, the
// longest probe among the 64 lanes of a wave sets the wave's cost).  Empty
// slots of a bucket are claimed in index order with CAS, so a key never sits
// behind an empty slot of its probe sequence.
template <typename K>
__device__ __forceinline__ void lds_bucket(const K *keys, uint32_t s0, K *k4) {
  if constexpr (sizeof(K) == 8) {
    const uint4 a = *(const uint4 *)(keys + s0);
    const uint4 b = *(const uint4 *)(keys + s0 + 2);
    k4[0] = ((uint64_t)a.y << 32) | a.x;
    k4[1] = ((uint64_t)a.w << 32) | a.z;
    k4[2] = ((uint64_t)b.y << 32) | b.x;
    k4[3] = ((uint64_t)b.w << 32) | b.z;
  } else {
    const uint4 a = *(const uint4 *)(keys + s0);
    k4[0] = a.x; k4[1] = a.y; k4[2] = a.z; k4[3] = a.w;
  }
}

template <typename K>
__device__ __forceinline__ int lds_find_or_insert(const LdsTable<K> &t, K key, uint64_t h, int max_probe) {
  const uint32_t S = t.L.slots;  // a multiple of 4
  if (key == KeyTraits<K>::kEmpty) {
    *t.sentinel = 1;
    return (int)S;
  }
  K *keys = t.keys();
  const uint32_t nb = S >> 2;
  uint32_t b = hash_lds_slot(h, nb);
  for (int i = 0; i < max_probe; ++i) {
    const uint32_t s0 = b << 2;
    K k4[4];
    lds_bucket<K>(keys, s0, k4);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (k4[j] == key) return (int)(s0 + j);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (k4[j] != KeyTraits<K>::kEmpty) continue;
      const K prev = atomicCAS(&keys[s0 + j], KeyTraits<K>::kEmpty, key);
      if (prev == KeyTraits<K>::kEmpty || prev == key) return (int)(s0 + j);
    }
    b = (b + 1 == nb) ? 0 : b + 1;
  }
  return -1;
}

// Apply one folded operand to LDS slot `s`.
template <typename K>
__device__ __forceinline__ void lds_apply(const LdsTable<K> &t, int a, int kind, int s, uint64_t v) {
  switch (kind) {
    case ACC_COUNT: atomicAdd(t.template arr<uint32_t>(a) + s, (uint32_t)v); break;
    case ACC_SUM_I: atomicAdd(t.template arr<unsigned long long>(a) + s, (unsigned long long)v); break;
    case ACC_MIN_I32: atomicMin(t.template arr<int>(a) + s, (int)(int32_t)v); break;
    case ACC_MAX_I32: atomicMax(t.template arr<int>(a) + s, (int)(int32_t)v); break;
    case ACC_MIN_I64: atomicMin(t.template arr<long long>(a) + s, (long long)v); break;
    case ACC_MAX_I64: atomicMax(t.template arr<long long>(a) + s, (long long)v); break;
    default: break;
  }
}

template <typename K>
__device__ __forceinline__ void lds_apply_limbs(const LdsTable<K> &t, int a, int s, const int64_t *lv) {
  unsigned long long *p = t.template arr<unsigned long long>(a);
  const uint32_t n = t.L.slots + 1;
#pragma unroll
  for (int l = 0; l < kLimbs; ++l)
    if (lv[l]) atomicAdd(p + (size_t)l * n + s, (unsigned long long)lv[l]);
}

// ---------------------------------------------------------------------------
// Global-memory accessors.  A pointer the compiler cannot place -- one read
// from memory (a query record, a table descriptor) or passed into a function
// it does not inline -- is accessed with generic ("flat") instructions, which
// count against the LDS counter as well as the vector-memory one: a wait for
// an LDS result then also waits for every global load in flight, and a
// prefetch stops overlapping anything.  Through gp() and the g_* atomics
// the accesses are global instructions (same relaxed, device-scope atomics
// as atomicAdd and its kin).
// ---------------------------------------------------------------------------
#if defined(__AMDGCN__)
#define TGA_GAS __attribute__((address_space(1)))
// constant memory: data no kernel writes while it runs (query records)
#define TGA_CAS __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ TGA_GAS T *gp(T *p) {
  return (TGA_GAS T *)p;
}
template <typename T, typename U>
__device__ __forceinline__ T g_add(T *p, U v) {
  return __hip_atomic_fetch_add(gp(p), (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, typename U>
__device__ __forceinline__ T g_min(T *p, U v) {
  return __hip_atomic_fetch_min(gp(p), (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, typename U>
__device__ __forceinline__ T g_max(T *p, U v) {
  return __hip_atomic_fetch_max(gp(p), (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T, typename U>
__device__ __forceinline__ T g_or(T *p, U v) {
  return __hip_atomic_fetch_or(gp(p), (T)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// compare-and-swap; returns the word found
template <typename T>
__device__ __forceinline__ T g_cas(T *p, T cmp, T val) {
  __hip_atomic_compare_exchange_strong(gp(p), &cmp, val, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
  return cmp;
}
#else  // host pass and host model of the kernels: plain pointers, the HIP atomics
#define TGA_GAS
#define TGA_CAS
template <typename T>
__device__ __forceinline__ T *gp(T *p) { return p; }
template <typename T, typename U>
__device__ __forceinline__ T g_add(T *p, U v) { return atomicAdd(p, (T)v); }
template <typename T, typename U>
__device__ __forceinline__ T g_min(T *p, U v) { return atomicMin(p, (T)v); }
template <typename T, typename U>
__device__ __forceinline__ T g_max(T *p, U v) { return atomicMax(p, (T)v); }
template <typename T, typename U>
__device__ __forceinline__ T g_or(T *p, U v) { return atomicOr(p, (T)v); }
template <typename T>
__device__ __forceinline__ T g_cas(T *p, T cmp, T val) { return atomicCAS(p, cmp, val); }
#endif

// ---------------------------------------------------------------------------
// global table in HBM: the same SoA shape with `S` regular slots
// ---------------------------------------------------------------------------

// No shared counter is touched: every insert incrementing (or every probe
// reading) one word serialises the whole GPU on that word's L2 channel
// (26.8M inserts took ~140 ms).  The table's load shows in its probe
// lengths instead: linear probing with this hash on 2^26 slots probes at
// most 51 slots at load 1/2 (the tables' sizing), 231 at 3/4, and past
// kMaxGlobalProbe = 256 dozens of times at 4/5, so a probe that long raises
// the overflow flag and the host runs the call again with a larger table.
// Once the flag is up nothing more is inserted; it is read every 32 probe
// steps, so only long probes read it.
template <typename K>
__device__ __forceinline__ int64_t g_find_or_insert(const GTable &g, K key, uint64_t h) {
  if (key == KeyTraits<K>::kEmpty) {
    g_or(g.flags + FLAG_SENTINEL, 1u);
    return (int64_t)g.S;
  }
  K *keys = (K *)g.keys;
  const uint64_t S = g.S;
  uint64_t s = hash_global_slot(h, S);
  const uint64_t limit = S < kMaxGlobalProbe ? S : kMaxGlobalProbe;
  for (uint64_t i = 0; i < limit; ++i) {
    if ((i & 31) == 31 && *(volatile TGA_GAS uint32_t *)gp(g.flags + FLAG_OVERFLOW)) return -1;
    const K cur = gp(keys)[s];
    if (cur == key) return (int64_t)s;
    if (cur == KeyTraits<K>::kEmpty) {
      const K prev = g_cas(keys + s, KeyTraits<K>::kEmpty, key);
      if (prev == KeyTraits<K>::kEmpty || prev == key) return (int64_t)s;
    }
    s = (s + 1 == S) ? 0 : s + 1;
  }
  g_or(g.flags + FLAG_OVERFLOW, 1u);
  return -1;
}

__device__ __forceinline__ void g_apply(const GTable &g, int a, int kind, int64_t s, uint64_t v) {
  switch (kind) {
    case ACC_COUNT: case ACC_SUM_I: g_add((unsigned long long *)g.acc[a] + s, (unsigned long long)v); break;
    case ACC_MIN_I32: g_min((int *)g.acc[a] + s, (int)(int32_t)v); break;
    case ACC_MAX_I32: g_max((int *)g.acc[a] + s, (int)(int32_t)v); break;
    case ACC_MIN_I64: g_min((long long *)g.acc[a] + s, (long long)v); break;
    case ACC_MAX_I64: g_max((long long *)g.acc[a] + s, (long long)v); break;
    default: break;
  }
}

__device__ __forceinline__ void g_apply_limbs(const GTable &g, int a, int64_t s, const int64_t *lv) {
  unsigned long long *p = (unsigned long long *)g.acc[a];
  const uint64_t n = g.S + 1;
#pragma unroll
  for (int l = 0; l < kLimbs; ++l)
    if (lv[l]) g_add(p + (uint64_t)l * n + s, (unsigned long long)lv[l]);
}

// Expand a double into a limb vector (zeros except at most three limbs).
__device__ __forceinline__ void fx_expand(uint64_t bits, int64_t *lv, uint32_t *flags) {
#pragma unroll
  for (int l = 0; l < kLimbs; ++l) lv[l] = 0;
  int64_t c[3];
  int range = 0;
  const int i0 = fx_split(bits, c, &range);
  if (range) g_or(flags + FLAG_RANGE, 1u);
  if (i0 < 0) return;
#pragma unroll
  for (int l = 0; l < kLimbs; ++l) {
    const int d = l - i0;
    if (d == 0) lv[l] = c[0];
    else if (d == 1) lv[l] = c[1];
    else if (d == 2) lv[l] = c[2];
  }
}

// ---------------------------------------------------------------------------
// decoupled look-back (Merrill & Garland 2016)
//
// desc[t] = flag (bits 63..62: 0 nothing yet, 1 tile aggregate, 2 inclusive
// prefix) | value (bits 61..0).  Tile ids come from an atomic ticket, so every
// lower tile belongs to a workgroup that started earlier and is running; the
// spin below therefore always ends.
// ---------------------------------------------------------------------------

constexpr uint64_t kLbAgg = 1ULL << 62;
constexpr uint64_t kLbInc = 2ULL << 62;
constexpr uint64_t kLbVal = (1ULL << 62) - 1;

__device__ __forceinline__ void lb_store(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lb_load(uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by all 64 lanes of one wave.  Returns the exclusive prefix of tile
// `tile` and publishes its inclusive prefix.
__device__ inline uint64_t lookback_exclusive(uint64_t *desc, uint64_t tile, uint64_t agg) {
  const int lane = (int)__lane_id();
  if (tile == 0) {
    if (lane == 0) lb_store(&desc[0], kLbInc | (agg & kLbVal));
    return 0;
  }
  if (lane == 0) lb_store(&desc[tile], kLbAgg | (agg & kLbVal));
  uint64_t excl = 0;
  int64_t end = (int64_t)tile;  // looking at tiles [end-64, end)
  for (;;) {
    const int64_t idx = end - 1 - lane;
    uint64_t d = idx >= 0 ? lb_load(&desc[idx]) : kLbInc;
    while (__ballot((d >> 62) == 0) != 0) {
      if ((d >> 62) == 0) {
        __builtin_amdgcn_s_sleep(1);
        d = lb_load(&desc[idx]);
      }
    }
    const uint64_t incmask = __ballot((d >> 62) == 2);
    if (incmask) {
      const int k = __builtin_ctzll(incmask);  // nearest tile with an inclusive prefix
      const uint64_t v = lane <= k ? (d & kLbVal) : 0;
      excl += wave_sum<uint64_t>(v);
      break;
    }
    excl += wave_sum<uint64_t>(d & kLbVal);
    end -= kWave;
  }
  if (lane == 0) lb_store(&desc[tile], kLbInc | ((excl + agg) & kLbVal));
  return excl;
}

// ---------------------------------------------------------------------------
// block-wide helpers and the row -> LDS table -> global table pipeline
// ---------------------------------------------------------------------------

// Workgroup barrier that orders LDS only.  __syncthreads() is also a
// workgroup-scope release fence, for which the compiler first waits for every
// outstanding global load and store of the wave (vmcnt(0)); a kernel that
// keeps loads in flight across its barriers (the next tile's rows) would wait
// for them at each one.  Here a wave waits for its own LDS accesses only
// (lgkmcnt(0) = s_waitcnt 0xc07f on gfx9-family targets), then meets the
// others at s_barrier; the signal fences keep the compiler from moving memory
// accesses across it.  Only for barriers that order LDS data.
__device__ __forceinline__ void lds_barrier() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_s_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Marks a register value as defined here: an empty asm statement that reads
// and writes it.  After an explicit s_waitcnt has completed the loads of a
// register set, pinning the set tells the compiler no load is outstanding on
// it any more, so it inserts no further waits for those registers (its own
// tracking loses that across loop back edges and then waits there, behind
// whatever was issued since).  A no-op outside the device compilation.
template <typename T>
__device__ __forceinline__ void pin_reg(T &x) {
#if defined(__AMDGCN__)
  asm volatile("" : "+v"(x));
#else
  (void)x;
#endif
}

struct SyncBarrier {
  __device__ __forceinline__ void operator()() const { __syncthreads(); }
};
struct LdsBarrier {
  __device__ __forceinline__ void operator()() const { lds_barrier(); }
};

// Exclusive scan of n u32 values; returns the total.  Element i is get(i);
// its exclusive prefix goes to put(i, prefix).  get(i) is called twice, both
// times before put(i).  scratch needs blockDim/64 + 1 words.  All threads
// must call it.  Bar: the barrier used between the phases (LdsBarrier when
// global accesses must stay in flight).
template <typename Bar, typename Get, typename Put>
__device__ __forceinline__ uint32_t block_exclusive_scan_fn(uint32_t n, uint32_t *scratch, Get get, Put put) {
  const Bar bar{};
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  const uint32_t per = (n + nt - 1) / nt;
  const uint32_t b = tid * per, e = min(n, b + per);
  uint32_t local = 0;
  for (uint32_t i = b; i < e; ++i) local += get(i);
  // inclusive wave scan of the per-thread sums
  const int lane = (int)__lane_id(), wave = (int)(tid >> 6);
  uint32_t x = local;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o, kWave);
    if (lane >= o) x += y;
  }
  if (lane == kWave - 1) scratch[wave] = x;
  bar();
  if (tid == 0) {
    uint32_t run = 0;
    const uint32_t nw = nt >> 6;
    for (uint32_t w = 0; w < nw; ++w) { const uint32_t v = scratch[w]; scratch[w] = run; run += v; }
    scratch[nw] = run;
  }
  bar();
  uint32_t run = scratch[wave] + x - local;
  for (uint32_t i = b; i < e; ++i) { const uint32_t v = get(i); put(i, run); run += v; }
  const uint32_t total = scratch[nt >> 6];
  bar();
  return total;
}

// Exclusive scan of n <= 64 * PER values by one wave, with no barrier;
// returns the total to every lane.  Lane l takes elements [l*per, l*per+per).
template <int PER, typename Get, typename Put>
__device__ __forceinline__ uint32_t wave_exclusive_scan_fn(uint32_t n, Get get, Put put) {
  const uint32_t lane = __lane_id();
  const uint32_t per = (n + kWave - 1) / kWave;
  const uint32_t b = lane * per;
  uint32_t v[PER], local = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    v[k] = (uint32_t)k < per && b + k < n ? get(b + k) : 0u;
    local += v[k];
  }
  uint32_t x = local;
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)x, o, kWave);
    if ((int)lane >= o) x += y;
  }
  const uint32_t total = (uint32_t)__shfl((int)x, kWave - 1, kWave);
  uint32_t run = x - local;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    if ((uint32_t)k < per && b + k < n) put(b + k, run);
    run += v[k];
  }
  return total;
}

// The same over n values held in LDS, in place.
template <typename Bar = SyncBarrier>
__device__ __forceinline__ uint32_t block_exclusive_scan_lds(uint32_t *arr, uint32_t n, uint32_t *scratch) {
  return block_exclusive_scan_fn<Bar>(n, scratch, [&](uint32_t i) { return arr[i]; },
                                      [&](uint32_t i, uint32_t v) { arr[i] = v; });
}

// ---------------------------------------------------------------------------
// rows of a work item
//
// Contiguous mode: rows [b, e), taken by the waves in blocks of kBucketRows
// rows (wave w takes blocks w, w+nw, ...).  Bucket mode (ents != null): b..e
// index a list of bucket entries, rows << 32 | id; bucket `id` holds `rows`
// (1..kBucketRows) rows at [id*kBucketRows, ...) of the partitioned copy.
// One entry is one wave block: the bucket of ents[k], followed by the bucket
// of ents2[k] when that is not 0 (two partly filled buckets of at most
// kBucketRows rows together).  An entry of 0 rows ends the list.
// ---------------------------------------------------------------------------

struct BucketList {
  const uint64_t *ents;   // null: contiguous mode
  const uint64_t *ents2;  // with ents: the second bucket of each block
};

struct ListEnt {
  uint64_t a, b;
};

// List entry of block i of wave `wave` (0 in contiguous mode).  Past the end
// of the list it is the last entry: the loads are unconditional, so the
// registers they fill are never also set by another path while in flight
// (which would make the compiler wait for every load in flight first).
__device__ __forceinline__ ListEnt list_entry(const BucketList &bl, uint64_t b, uint64_t e, uint64_t i,
                                              uint32_t wave, uint32_t nw) {
  if (!bl.ents || e <= b) return ListEnt{0, 0};
  const uint64_t k = b + wave + i * nw;
  const uint64_t kk = k < e ? k : e - 1;
  return ListEnt{bl.ents[kk], bl.ents2[kk]};
}

// Block i of wave `wave` (wave-uniform), given its list entry: its first
// row, row count (1..kBucketRows) and, for a block of two buckets, the first
// row of the second one and the row count of the first (ca == cnt: one
// bucket, or contiguous rows).  The row loop addresses a block from a scalar
// base with 32-bit lane offsets; false past the end.
__device__ __forceinline__ bool wave_block_n(const BucketList &bl, uint64_t b, uint64_t e, uint64_t i, uint32_t wave,
                                             uint32_t nw, ListEnt ent, uint64_t &bb, uint32_t &cnt, uint64_t &bb2,
                                             uint32_t &ca) {
  if (bl.ents) {
    const uint64_t x = wave_uniform64(ent.a), y = wave_uniform64(ent.b);
    if (b + wave + i * nw >= e || (x >> 32) == 0) return false;
    bb = (x & 0xffffffffULL) << kBucketShift;
    ca = (uint32_t)(x >> 32);
    bb2 = (y & 0xffffffffULL) << kBucketShift;
    cnt = ca + (uint32_t)(y >> 32);
    return true;
  }
  bb = b + (wave + i * nw) * (uint64_t)kBucketRows;
  if (bb >= e) return false;
  cnt = (uint32_t)min<uint64_t>(kBucketRows, e - bb);
  bb2 = bb;
  ca = cnt;
  return true;
}

// ---------------------------------------------------------------------------
// row loop for integer accumulators
// ---------------------------------------------------------------------------

// Raw 64-bit image of value column c at `row` (dtype packed in coldt, 4 bits
// per column).  VW = 4 or 8: every value column has that width, so the load
// is branch-free; VW = 0: the width is read from coldt.
// Value width codes of the row loaders for a partitioned copy stored as
// records: kVwRecBase + VD, where VD is the record's value dwords (the value
// columns in order, each at its own width; see RecW).  kVwRec4: one 4-byte
// value column, 8-byte records with 32-bit keys, 12-byte with 64-bit.
constexpr int kVwRecBase = 100;
constexpr int kVwRec4 = kVwRecBase + 1;
constexpr bool vw_is_rec(int vw) { return vw > kVwRecBase; }
constexpr int vw_rec_vd(int vw) { return vw > kVwRecBase ? vw - kVwRecBase : 0; }

// A record of n 4-byte words.  The compiler merges the word accesses of one
// record into wide loads and stores (dwordx3 for 12 bytes).  (A 3-element
// vector type would not do: clang widens its loads and stores to 16 bytes,
// which reaches into the next record.)
template <int N>
struct Dwords {
  uint32_t w[N];
};

// A record of the partitioned copy: the key (KD = 1 or 2 words, low word
// first) followed by VD value words, the value columns in order at their own
// widths (an 8-byte column low word first).  With one 4-byte value column
// these are the 8- and 12-byte (key, value) records.
template <typename K, int VD>
struct RecW {
  static constexpr int KD = (int)sizeof(K) / 4;
  static constexpr int RD = KD + VD;
  static constexpr int kBytes = RD * 4;
  __device__ __forceinline__ static void unpack(const uint32_t *w, K &key, uint32_t (&val)[VD]) {
    if constexpr (KD == 2) key = (K)(((uint64_t)w[1] << 32) | w[0]);
    else key = (K)w[0];
#pragma unroll
    for (int j = 0; j < VD; ++j) val[j] = w[KD + j];
  }
  // record at byte offset `off` (32-bit) of `base`
  __device__ __forceinline__ static void load_at(const void *base, uint32_t off, K &key, uint32_t (&val)[VD]) {
    const char *p = (const char *)base + off;
    if constexpr (RD == 2) {  // 8-byte aligned: one dwordx2, a global access
      const uint64_t v = *gp((const uint64_t *)p);
      const uint32_t w[2] = {(uint32_t)v, (uint32_t)(v >> 32)};
      unpack(w, key, val);
    } else if constexpr (RD == 4) {  // 16-byte aligned: one dwordx4
      const uint4 v = *(const uint4 *)p;
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
      unpack(w, key, val);
    } else {
      const Dwords<RD> v = *(const Dwords<RD> *)p;
      unpack(v.w, key, val);
    }
  }
  __device__ __forceinline__ static void load(const void *base, uint64_t r, K &key, uint32_t (&val)[VD]) {
    load_at((const char *)base + r * (uint64_t)kBytes, 0u, key, val);
  }
  __device__ __forceinline__ static void store(void *base, uint64_t r, K key, const uint32_t (&val)[VD]) {
    char *p = (char *)base + r * (uint64_t)kBytes;
    Dwords<RD> v;
    v.w[0] = (uint32_t)(uint64_t)key;
    if constexpr (KD == 2) v.w[1] = (uint32_t)((uint64_t)key >> 32);
#pragma unroll
    for (int j = 0; j < VD; ++j) v.w[KD + j] = val[j];
    if constexpr (RD == 2) *(uint2 *)p = make_uint2(v.w[0], v.w[1]);
    else *(Dwords<RD> *)p = v;
  }
};

// (key, 4-byte value) records
template <typename K>
struct Rec4 {
  static constexpr int kBytes = RecW<K, 1>::kBytes;
  __device__ __forceinline__ static void load_at(const void *base, uint32_t off, K &key, uint32_t &val) {
    uint32_t v[1];
    RecW<K, 1>::load_at(base, off, key, v);
    val = v[0];
  }
  __device__ __forceinline__ static void load(const void *base, uint64_t r, K &key, uint32_t &val) {
    uint32_t v[1];
    RecW<K, 1>::load(base, r, key, v);
    val = v[0];
  }
  __device__ __forceinline__ static void store(void *base, uint64_t r, K key, uint32_t val) {
    const uint32_t v[1] = {val};
    RecW<K, 1>::store(base, r, key, v);
  }
};

// Word `o` (uniform) of a record's VD value words, by selects (no indexed
// register access).  Past the end: word 0.
template <int VD>
__device__ __forceinline__ uint32_t rec_word(const uint32_t *w, uint32_t o) {
  uint32_t r = w[0];
#pragma unroll
  for (int d = 1; d < VD; ++d) r = o == (uint32_t)d ? w[d] : r;
  return r;
}

template <int VW>
__device__ __forceinline__ uint64_t load_col(const void *const *cols, uint32_t coldt, int c, uint64_t row) {
  const uint32_t dt = (coldt >> (4 * c)) & 15u;
  if constexpr (VW == 4) {
    const uint32_t v = gp((const uint32_t *)cols[c])[row];
    return dt == DT_I32 ? (uint64_t)(int64_t)(int32_t)v : (uint64_t)v;
  } else if constexpr (VW == 8) {
    return gp((const uint64_t *)cols[c])[row];
  } else {
    if (dt == DT_U32) return (uint64_t)gp((const uint32_t *)cols[c])[row];
    if (dt == DT_I32) return (uint64_t)(int64_t)gp((const int32_t *)cols[c])[row];
    return gp((const uint64_t *)cols[c])[row];
  }
}

template <int NC>
__device__ __forceinline__ uint64_t pick_col(const uint64_t *raw, uint32_t c) {
  if constexpr (NC == 1) return raw[0];
  else if constexpr (NC == 2) return c == 0 ? raw[0] : raw[1];
  else if constexpr (NC == 3) return c == 0 ? raw[0] : c == 1 ? raw[1] : raw[2];
  else return c == 0 ? raw[0] : c == 1 ? raw[1] : c == 2 ? raw[2] : raw[3];
}

// Operand of accumulator `kind` (operand map `map`) for one raw cell.
__device__ __forceinline__ uint64_t prog_operand(uint32_t kind, uint32_t map, uint64_t raw) {
  if (kind == ACC_COUNT) return 1;
  if (map == MAP_U64) return raw ^ (1ULL << 63);
  if (map == MAP_F64) return (uint64_t)f64_bits_to_ordered(raw);
  return raw;
}

// One row whose key found no LDS slot goes to the global table.  Out of line
// on purpose: the row loop then carries none of the global-table or schema
// descriptors (they are read here, from `g` and `s`).
template <typename K>
__device__ __noinline__ void spill_row(const GTable *g, const Schema *s, K key, uint64_t raw0,
                                       uint64_t raw1, uint64_t raw2, uint64_t raw3) {
  const int64_t gs = g_find_or_insert<K>(*g, key, key_hash((uint64_t)key));
  if (gs < 0) return;
  const uint64_t cell[kMaxCols] = {raw0, raw1, raw2, raw3};
  for (int a = 0; a < s->nacc; ++a) {
    if (s->acc[a].kind == ACC_SUM_F) {
      int64_t lv[kLimbs];
      fx_expand(cell[s->acc[a].col], lv, g->flags);
      g_apply_limbs(*g, a, gs, lv);
    } else {
      g_apply(*g, a, s->acc[a].kind, gs, acc_operand(*s, a, cell));
    }
  }
}

// Aggregates rows [b, e) into the LDS table for schemas without fp64 SUM.
//
// Per wave and step, U rows per lane: the keys' first buckets are read for
// all U rows before any is resolved (U LDS round trips in flight), keys not
// in their first bucket probe on / insert, rows with no slot go to the global
// table, and the accumulators are updated with return-less LDS atomics, one
// uniform dispatch per accumulator per step.
//
// FOLD: before the atomics, lanes of a wave holding the same slot are found
// with one ballot per slot bit and folded into their lowest lane, so each
// slot gets one atomic per wave.  That pays off only when one key fills a
// large share of the rows; measured on MI355X the LDS serialises the
// same-address lanes of one wave faster than the fold costs, so it is
// opt-in (tga_options.wave_fold).
//
// giveup (LDS, two words; null = spill mode): a work item whose groups do
// not fit the table is abandoned instead of spilled -- giveup[0] is set,
// every wave leaves the loop, and giveup[1] receives the rows the workgroup
// had consumed, from which the host sizes a finer partitioning.
// spilled (LDS word, optional, spill mode): set when a row went to the
// global table.
//
// The loop is software-pipelined: the next block's key and value loads are
// issued before the current block's LDS work, so every wave keeps its HBM
// reads in flight while it probes and updates the table.
// Register type of a prefetched value cell: 4-byte cells stay 32-bit until
// widen_block, so no 64-bit register pair is tied to a load in flight.
template <int VW>
using CellOf = typename std::conditional<(VW == 4 || vw_is_rec(VW)), uint32_t, uint64_t>::type;

// Rows [base, base + cnt) of a block (base and cnt wave-uniform, cnt >= 1):
// lane l takes rows u * 64 + l.
template <typename K, int NC, int VW, int U, bool PAIR = false>
__device__ __forceinline__ void load_block(const K *__restrict__ keys, const void *const *cols, uint32_t ncols,
                                           uint32_t coldt, uint64_t base, uint32_t cnt, int lane, K *kk,
                                           CellOf<VW> (*raw)[NC], uint64_t base2 = 0, uint32_t ca = 0) {
  // rows past the end re-read the block's last row (in bounds, no branch),
  // so all U key and value loads are in flight together
#pragma unroll
  for (int u = 0; u < U; ++u) {
    uint32_t j = min((uint32_t)(u * kWave + lane), cnt - 1);
    uint64_t blk = base;
    if constexpr (PAIR) {
      // rows ca.. of the block are the second bucket's (a per-lane base)
      const bool first = j < ca;
      blk = first ? base : base2;
      j = first ? j : j - ca;
    }
    const uint64_t r = blk + j;
    if constexpr (VW == kVwRec4) {  // records (key, 4-byte value) at `keys`
      static_assert(NC == 1, "records carry one value column");
      uint32_t v;
      // a scalar block base plus a 32-bit lane offset
      Rec4<K>::load_at((const char *)keys + blk * (uint64_t)Rec4<K>::kBytes, j * (uint32_t)Rec4<K>::kBytes, kk[u], v);
      raw[u][0] = v;
    } else if constexpr (vw_is_rec(VW)) {  // records of VD value words at `keys`
      constexpr int VD = vw_rec_vd(VW);
      static_assert(VD <= NC, "one cell per value word");
      using R = RecW<K, VD>;
      uint32_t v[VD];
      R::load_at((const char *)keys + blk * (uint64_t)R::kBytes, j * (uint32_t)R::kBytes, kk[u], v);
#pragma unroll
      for (int d = 0; d < NC; ++d) raw[u][d] = d < VD ? v[d < VD ? d : 0] : 0u;
    } else if constexpr (VW == 4) {
#pragma unroll
      for (int c = 0; c < NC; ++c) raw[u][c] = (uint32_t)c < ncols ? gp((const uint32_t *)cols[c])[r] : 0u;
      kk[u] = gp(keys)[r];
    } else {
      kk[u] = gp(keys)[r];
#pragma unroll
      for (int c = 0; c < NC; ++c) raw[u][c] = (uint32_t)c < ncols ? load_col<VW>(cols, coldt, c, r) : 0;
    }
  }
}

// The cells as 64-bit operands: 4-byte cells are widened (sign-extended for
// int32 columns) only here, where the row loop first uses them -- widening
// at the load would make the wave wait for the prefetched block as soon as
// it is issued.
template <int NC, int VW, int U>
__device__ __forceinline__ void widen_block(uint32_t coldt, const CellOf<VW> (*cell)[NC], uint64_t (*raw)[NC]) {
  if constexpr (vw_rec_vd(VW) > 1) {
    // records of several value words: column c starts at word o_c, the sum
    // of the widths before it (coldt is uniform, so are the o_c); absent
    // columns (dtype 0) read as 0
    constexpr int VD = vw_rec_vd(VW);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      uint32_t o = 0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const uint32_t dt = (coldt >> (4 * c)) & 15u;
        const uint32_t lo = rec_word<VD>(cell[u], o), hi = rec_word<VD>(cell[u], o + 1);
        raw[u][c] = dt == 0 ? 0 : dt >= DT_U64 ? ((uint64_t)hi << 32 | lo)
                  : dt == DT_I32 ? (uint64_t)(int64_t)(int32_t)lo : (uint64_t)lo;
        o += dt >= DT_U64 ? 2u : 1u;
      }
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if constexpr (VW == 4 || VW == kVwRec4)
        raw[u][c] = ((coldt >> (4 * c)) & 15u) == DT_I32 ? (uint64_t)(int64_t)(int32_t)cell[u][c] : (uint64_t)cell[u][c];
      else
        raw[u][c] = cell[u][c];
    }
}

// Volatile 4-byte read of an LDS word through a generic pointer to it, as an
// LDS access (a generic one would also wait for every global load in flight).
__device__ __forceinline__ uint32_t lds_read_volatile(const uint32_t *p) {
#if defined(__AMDGCN__)
  return *(const volatile __attribute__((address_space(3))) uint32_t *)p;
#else
  return *(const volatile uint32_t *)p;
#endif
}

// Accumulator program of the row loop.  SIG == 0: decoded from the AccProg
// at run time (any schema).  SIG != 0: fixed at compile time, so the per-row
// dispatch over accumulator kinds folds away -- accumulator a has kind
// ((SIG >> 4a) & 15) - 1, on value column 0 with the raw operand, for every
// nonzero nibble (the common one-column integer schemas).
constexpr uint32_t kSigCountSum = (ACC_COUNT + 1) | (ACC_SUM_I + 1) << 4;
constexpr uint32_t kSigCountSumMinMax = kSigCountSum | (ACC_MIN_I32 + 1) << 8 | (ACC_MAX_I32 + 1) << 12;

template <uint32_t SIG>
__device__ __forceinline__ uint32_t prog_nacc(const AccProg &ap) {
  if constexpr (SIG != 0) {
    uint32_t n = 0;
    while (n < 8 && ((SIG >> (4 * n)) & 15u)) ++n;
    return n;
  } else {
    return ap.nacc;
  }
}
template <uint32_t SIG>
__device__ __forceinline__ uint32_t prog_kind(const AccProg &ap, int a) {
  if constexpr (SIG != 0) return ((SIG >> (4 * a)) & 15u) - 1u;
  else return (ap.kinds >> (4 * a)) & 15u;
}
template <uint32_t SIG>
__device__ __forceinline__ uint32_t prog_col(const AccProg &ap, int a) {
  if constexpr (SIG != 0) return 0u;
  else return (ap.cols >> (2 * a)) & 3u;
}
template <uint32_t SIG>
__device__ __forceinline__ uint32_t prog_map(const AccProg &ap, int a) {
  if constexpr (SIG != 0) return MAP_RAW;
  else return (ap.maps >> (2 * a)) & 3u;
}

// FXA: the schema may hold exact fp64 SUM accumulators (never with FOLD):
// a row's double is split into at most three signed 32-bit chunks of its
// fixed-point image (fx_split) and each nonzero chunk is added to its limb
// with one return-less LDS atomic, as the integer accumulators are.
template <typename K, int NC, int VW, bool FOLD, uint32_t SIG = 0, bool FXA = false>
__device__ __forceinline__ void agg_rows_int(const K *__restrict__ keys, const void *const *cols, uint64_t b0,
                                             uint64_t e0, const LdsTable<K> &t, const AccProg &ap, int nbits,
                                             int max_probe, const GTable *gsp, const Schema *ssp,
                                             uint32_t *giveup = nullptr, uint32_t *spilled = nullptr,
                                             const BucketList bl = BucketList{nullptr}) {
  constexpr int U = 4;
  static_assert(kWave * U == (int)kBucketRows, "one bucket is one wave block");
  const int lane = (int)__lane_id();
  const uint32_t wave = wave_uniform(threadIdx.x >> 6);
  const uint32_t nw = blockDim.x >> 6;
  const uint32_t S = t.L.slots;
  const uint32_t nb = S >> 2;
  K *tk = t.keys();
  char *tb = t.base;
  const uint32_t ncols = ap.ncols, coldt = ap.coldt;
  uint32_t consumed = 0;
  // bucket mode: the list entry of a block is read two blocks ahead, so no
  // wave waits on it before issuing that block's row loads
  uint64_t base = 0, base2 = 0;
  uint32_t cnt = 0, ca = 0;
  bool have = wave_block_n(bl, b0, e0, 0, wave, nw, list_entry(bl, b0, e0, 0, wave, nw), base, cnt, base2, ca);
  ListEnt ent_next = list_entry(bl, b0, e0, 1, wave, nw);
  K kk[U];
  CellOf<VW> cel[U][NC];
  // Blocks of two buckets come only with 8-byte keys (the host pairs no
  // others): the kernels of 4-byte keys keep the one block load, whose
  // registers the second form would add to.
  constexpr bool kPairs = sizeof(K) == 8;
  if (have) {
    if (kPairs && ca != cnt) load_block<K, NC, VW, U, true>(keys, cols, ncols, coldt, base, cnt, lane, kk, cel, base2, ca);
    else load_block<K, NC, VW, U>(keys, cols, ncols, coldt, base, cnt, lane, kk, cel);
  }
  // The first block is waited for here, so the loop is entered with no load
  // in flight, as it is re-entered: otherwise the compiler, merging the two
  // states, waits inside every iteration for the loads it has just issued
  // (the counter retires loads in order).
  __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
  // the give-up word is read in one iteration and tested in the next, so the
  // wave never waits for the read
  uint32_t stop = 0;
  for (uint64_t it = 1; have; ++it) {
    if (giveup && wave_uniform(stop)) break;
    if (giveup) stop = lds_read_volatile(giveup);
    consumed += kWave * U;
    uint64_t nbase = 0, nbase2 = 0;
    uint32_t ncnt = 0, nca = 0;
    const bool next = wave_block_n(bl, b0, e0, it, wave, nw, ent_next, nbase, ncnt, nbase2, nca);
    K nkk[U];
    CellOf<VW> ncel[U][NC];
    if (next) {
      if (kPairs && nca != ncnt)
        load_block<K, NC, VW, U, true>(keys, cols, ncols, coldt, nbase, ncnt, lane, nkk, ncel, nbase2, nca);
      else
        load_block<K, NC, VW, U>(keys, cols, ncols, coldt, nbase, ncnt, lane, nkk, ncel);
    }
    ent_next = list_entry(bl, b0, e0, it + 1, wave, nw);
    uint32_t hh[U];
    int slot[U];
    // first buckets of all U rows, read before any is resolved
    uint32_t s0[U];
    K k4[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      hh[u] = key_slot_hash((uint64_t)kk[u]);
      s0[u] = hash_lds_slot(hh[u], nb) << 2;
      lds_bucket<K>(tk, s0[u], k4[u]);
    }
    bool sent = false;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool act = (uint32_t)(u * kWave + lane) < cnt;
      int m = -1;
#pragma unroll
      for (int j = 3; j >= 0; --j) m = k4[u][j] == kk[u] ? (int)(s0[u] + j) : m;
      const bool empty_key = kk[u] == KeyTraits<K>::kEmpty;
      sent |= act && empty_key;
      slot[u] = !act ? -1 : empty_key ? (int)S : m;
    }
    if (sent) *t.sentinel = 1;
    uint64_t raw[U][NC];
    widen_block<NC, VW, U>(coldt, cel, raw);
    // Rows whose key is not in its first bucket read the next one here.  A
    // key placed past its first bucket (about 1% of the keys at the fill the
    // tuner plans for) is nearly always in the next, and the probe loop below
    // -- which some lane would enter in almost every block otherwise -- is
    // left to first sightings and keys further on.  A key found is final; a
    // miss goes on to the probe loop, which starts over from the first
    // bucket.
    {
      bool far = false;
#pragma unroll
      for (int u = 0; u < U; ++u) far |= (uint32_t)(u * kWave + lane) < cnt && slot[u] < 0;
      if (__ballot(far)) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if ((uint32_t)(u * kWave + lane) < cnt && slot[u] < 0) {
            const uint32_t s1 = s0[u] + 4 == S ? 0u : s0[u] + 4;
            lds_bucket<K>(tk, s1, k4[u]);
            int m = -1;
#pragma unroll
            for (int j = 3; j >= 0; --j) m = k4[u][j] == kk[u] ? (int)(s1 + j) : m;
            slot[u] = m;
          }
        }
      }
    }
    // first sightings, and keys past their second bucket: probe on / insert
#pragma unroll
    for (int u = 0; u < U; ++u)
      if ((uint32_t)(u * kWave + lane) < cnt && slot[u] < 0) slot[u] = lds_find_or_insert(t, kk[u], hh[u], max_probe);
    // rows without an LDS slot (table full): the global table
    bool spill = false;
#pragma unroll
    for (int u = 0; u < U; ++u) spill |= (uint32_t)(u * kWave + lane) < cnt && slot[u] < 0;
    if (__ballot(spill)) {
      if (giveup) {
        if (spill) *giveup = 1;
      } else {
        if (spilled && spill) *spilled = 1;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if ((uint32_t)(u * kWave + lane) < cnt && slot[u] < 0)
            spill_row<K>(gsp, ssp, kk[u], raw[u][0], NC > 1 ? raw[u][NC > 1 ? 1 : 0] : 0,
                         NC > 2 ? raw[u][NC > 2 ? 2 : 0] : 0, NC > 3 ? raw[u][NC > 3 ? 3 : 0] : 0);
        }
      }
    }
    if constexpr (FOLD) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool on = slot[u] >= 0;
        if (!__ballot(on)) continue;
        const uint64_t m = match_mask((uint32_t)slot[u], nbits, on);
        const bool leader = on && ((m & lanemask_lt()) == 0);
        uint64_t rest = leader ? (m & (m - 1)) : 0ULL;
        uint64_t ops[kMaxAccs];
#pragma unroll
        for (int a = 0; a < kMaxAccs; ++a) {
          ops[a] = 0;
          if ((uint32_t)a < prog_nacc<SIG>(ap))
            ops[a] = prog_operand(prog_kind<SIG>(ap, a), prog_map<SIG>(ap, a),
                                  pick_col<NC>(raw[u], prog_col<SIG>(ap, a)));
        }
        while (__ballot(rest != 0)) {
          const int src = rest ? (int)__builtin_ctzll(rest) : lane;
#pragma unroll
          for (int a = 0; a < kMaxAccs; ++a) {
            if ((uint32_t)a >= prog_nacc<SIG>(ap)) break;
            const uint64_t o = shfl64(ops[a], src);
            if (rest) ops[a] = acc_fold((int)prog_kind<SIG>(ap, a), ops[a], o);
          }
          if (rest) rest &= rest - 1;
        }
        // leaders update their slot with the folded operands; every other
        // lane updates the discard entry
        {
          const int sl = leader ? slot[u] : (int)S + 1;
#pragma unroll
          for (int a = 0; a < kMaxAccs; ++a) {
            if ((uint32_t)a >= prog_nacc<SIG>(ap)) break;
            const uint32_t kind = prog_kind<SIG>(ap, a);
            char *arr = tb + ap.off[a];
            switch (kind) {
              case ACC_COUNT: atomicAdd((uint32_t *)arr + sl, (uint32_t)ops[a]); break;
              case ACC_SUM_I: atomicAdd((unsigned long long *)arr + sl, (unsigned long long)ops[a]); break;
              case ACC_MIN_I32: atomicMin((int *)arr + sl, (int)(int32_t)ops[a]); break;
              case ACC_MAX_I32: atomicMax((int *)arr + sl, (int)(int32_t)ops[a]); break;
              case ACC_MIN_I64: atomicMin((long long *)arr + sl, (long long)ops[a]); break;
              case ACC_MAX_I64: atomicMax((long long *)arr + sl, (long long)ops[a]); break;
              default: break;
            }
          }
        }
      }
    } else {
      // rows without a slot (past the end, or spilled) update the discard entry
      int sl[U];
#pragma unroll
      for (int u = 0; u < U; ++u) sl[u] = slot[u] >= 0 ? slot[u] : (int)S + 1;
#pragma unroll
      for (int a = 0; a < kMaxAccs; ++a) {
        if ((uint32_t)a >= prog_nacc<SIG>(ap)) break;
        const uint32_t kind = prog_kind<SIG>(ap, a);
        const uint32_t col = prog_col<SIG>(ap, a);
        const uint32_t map = prog_map<SIG>(ap, a);
        char *arr = tb + ap.off[a];
        switch (kind) {
          case ACC_COUNT:
#pragma unroll
            for (int u = 0; u < U; ++u) atomicAdd((uint32_t *)arr + sl[u], 1u);
            break;
          case ACC_SUM_I:
#pragma unroll
            for (int u = 0; u < U; ++u) atomicAdd((unsigned long long *)arr + sl[u], (unsigned long long)pick_col<NC>(raw[u], col));
            break;
          // MIN / MAX: a slot's value only moves one way, so a row that does
          // not improve the value read does not need its atomic (after k rows
          // of a group, the next improves it with probability ~1/k)
          case ACC_MIN_I32: {
            int cur[U];
#pragma unroll
            for (int u = 0; u < U; ++u) cur[u] = ((const int *)arr)[sl[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int v = (int)(int32_t)pick_col<NC>(raw[u], col);
              if (v < cur[u]) atomicMin((int *)arr + sl[u], v);
            }
            break;
          }
          case ACC_MAX_I32: {
            int cur[U];
#pragma unroll
            for (int u = 0; u < U; ++u) cur[u] = ((const int *)arr)[sl[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const int v = (int)(int32_t)pick_col<NC>(raw[u], col);
              if (v > cur[u]) atomicMax((int *)arr + sl[u], v);
            }
            break;
          }
          case ACC_MIN_I64: {
            long long cur[U];
#pragma unroll
            for (int u = 0; u < U; ++u) cur[u] = ((const long long *)arr)[sl[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const long long v = (long long)prog_operand(kind, map, pick_col<NC>(raw[u], col));
              if (v < cur[u]) atomicMin((long long *)arr + sl[u], v);
            }
            break;
          }
          case ACC_MAX_I64: {
            long long cur[U];
#pragma unroll
            for (int u = 0; u < U; ++u) cur[u] = ((const long long *)arr)[sl[u]];
#pragma unroll
            for (int u = 0; u < U; ++u) {
              const long long v = (long long)prog_operand(kind, map, pick_col<NC>(raw[u], col));
              if (v > cur[u]) atomicMax((long long *)arr + sl[u], v);
            }
            break;
          }
          case ACC_SUM_F:
            if constexpr (FXA) {
              // limb l of slot s at l * (slots + 1) + s; rows without a slot
              // add nothing (the discard entry has no limb rows)
              unsigned long long *lp = (unsigned long long *)arr;
#pragma unroll
              for (int u = 0; u < U; ++u) {
                if (slot[u] < 0) continue;
                int64_t c[3];
                int range = 0;
                const int i0 = fx_split(pick_col<NC>(raw[u], col), c, &range);
                if (range) g_or(gsp->flags + FLAG_RANGE, 1u);
                if (i0 < 0) continue;
#pragma unroll
                for (int j = 0; j < 3; ++j)
                  if (c[j] != 0 && i0 + j < kLimbs)
                    atomicAdd(lp + (uint32_t)(i0 + j) * (S + 1) + (uint32_t)slot[u], (unsigned long long)c[j]);
              }
            }
            break;
          default: break;
        }
      }
    }
    have = next;
    if (next) {
      base = nbase;
      cnt = ncnt;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        kk[u] = nkk[u];
#pragma unroll
        for (int c = 0; c < NC; ++c) cel[u][c] = ncel[u][c];
      }
    }
  }
  if (giveup && lane == 0) atomicAdd(giveup + 1, consumed);
}

// Adds LDS slot i's accumulators into global slot gs.
template <typename K>
__device__ __forceinline__ void flush_slot(const LdsTable<K> &t, const Schema &s, const GTable &g, uint32_t i,
                                           int64_t gs) {
  const uint32_t n = t.L.slots + 1;
#pragma unroll
  for (int a = 0; a < kMaxAccs; ++a) {
    if (a >= s.nacc) break;
    const int kind = s.acc[a].kind;
    if (kind == ACC_SUM_F) {
      int64_t lv[kLimbs];
      const int64_t *p = t.template arr<int64_t>(a);
#pragma unroll
      for (int l = 0; l < kLimbs; ++l) lv[l] = p[(size_t)l * n + i];
      g_apply_limbs(g, a, gs, lv);
    } else if (acc_lds_bytes(kind) == 4) {
      const uint32_t v = t.template arr<uint32_t>(a)[i];
      g_apply(g, a, kind, gs, kind == ACC_COUNT ? (uint64_t)v : (uint64_t)(int64_t)(int32_t)v);
    } else {
      g_apply(g, a, kind, gs, t.template arr<uint64_t>(a)[i]);
    }
  }
}

// Adds the LDS table into the global table.  Each lane takes R slots and
// issues their first global probes together, then their compare-and-swaps
// together: a group already in the global table costs one load, a new one
// one compare-and-swap, with R of either in flight per lane.  A key whose
// first slot holds another key takes the full probe.
template <typename K>
__device__ __forceinline__ void lds_flush(const LdsTable<K> &t, const Schema &s, const GTable &g) {
  constexpr int R = 4;
  const uint32_t S = t.L.slots;
  const K *keys = t.keys();
  K *gk = (K *)g.keys;
  const K kE = KeyTraits<K>::kEmpty;
  for (uint32_t i0 = threadIdx.x; i0 < S; i0 += R * blockDim.x) {
    K key[R], cur[R];
    uint64_t gs0[R];
    bool occ[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t i = i0 + (uint32_t)r * blockDim.x;
      key[r] = i < S ? keys[i] : kE;
      occ[r] = key[r] != kE;
      gs0[r] = occ[r] ? hash_global_slot(key_hash((uint64_t)key[r]), g.S) : 0;
      cur[r] = occ[r] ? gp(gk)[gs0[r]] : kE;
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (occ[r] && cur[r] == kE) cur[r] = g_cas(gk + gs0[r], kE, key[r]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!occ[r]) continue;
      // cur: the key read, or what the compare-and-swap found (empty: ours)
      const int64_t gs = (cur[r] == key[r] || cur[r] == kE) ? (int64_t)gs0[r]
                                                            : g_find_or_insert<K>(g, key[r], key_hash((uint64_t)key[r]));
      if (gs >= 0) flush_slot<K>(t, s, g, i0 + (uint32_t)r * blockDim.x, gs);
    }
  }
  // the sentinel key's slot
  if (threadIdx.x == 0 && *t.sentinel != 0) {
    const int64_t gs = g_find_or_insert<K>(g, kE, key_hash((uint64_t)kE));
    if (gs >= 0) flush_slot<K>(t, s, g, S, gs);
  }
}


// Writes one LDS table slot (key + every accumulator, converted to the output
// dtype) to output row `pos`.
template <typename K>
__device__ __forceinline__ void emit_row(const LdsTable<K> &t, const Schema &s, const OutDesc &od, uint32_t i, K key,
                                         uint64_t pos) {
  const uint32_t n = t.L.slots + 1;
  gp((K *)od.keys)[pos] = key;
  #pragma unroll
  for (int a = 0; a < kMaxAccs; ++a) {
    if (a >= s.nacc) break;
    const int kind = s.acc[a].kind;
    switch (kind) {
      case ACC_COUNT: gp((uint64_t *)od.cols[a])[pos] = t.template arr<uint32_t>(a)[i]; break;
      case ACC_SUM_I: gp((uint64_t *)od.cols[a])[pos] = t.template arr<uint64_t>(a)[i]; break;
      case ACC_MIN_I32: case ACC_MAX_I32: gp((uint32_t *)od.cols[a])[pos] = t.template arr<uint32_t>(a)[i]; break;
      case ACC_SUM_F: {
        int64_t L[kLimbs];
        const int64_t *p = t.template arr<int64_t>(a);
#pragma unroll
        for (int l = 0; l < kLimbs; ++l) L[l] = p[(size_t)l * n + i];
        gp((double *)od.cols[a])[pos] = fx_to_double(L);
        break;
      }
      default: {
        const int64_t v = t.template arr<int64_t>(a)[i];
        const int dt = s.acc[a].col_dtype;
        if (dt == DT_U64) gp((uint64_t *)od.cols[a])[pos] = (uint64_t)v ^ (1ULL << 63);
        else if (dt == DT_F64) gp((uint64_t *)od.cols[a])[pos] = ordered_to_f64_bits(v);
        else if (dt == DT_U32 || dt == DT_I32) gp((uint32_t *)od.cols[a])[pos] = (uint32_t)v;
        else gp((int64_t *)od.cols[a])[pos] = v;
        break;
      }
    }
  }
}

// Emits every occupied slot of the LDS table as final output rows: a
// partition processed whole by one workgroup holds final groups (its keys
// appear in no other partition).  Rows are reserved on *count with one
// atomic per workgroup.  scratch: blockDim/64 + 1 words; s_base: one word.
template <typename K>
__device__ __forceinline__ void lds_emit(const LdsTable<K> &t, const Schema &s, const OutDesc &od, uint64_t *count,
                         uint32_t *flags, uint32_t *scratch, uint64_t *s_base) {
  const uint32_t S = t.L.slots, n = S + 1;
  const K *keys = t.keys();
  const uint32_t per = (n + blockDim.x - 1) / blockDim.x;
  const uint32_t b = threadIdx.x * per, e = min(n, b + per);
  uint32_t local = 0;
  for (uint32_t i = b; i < e; ++i) local += (i < S) ? (keys[i] != KeyTraits<K>::kEmpty) : (*t.sentinel != 0);
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
    for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) { const uint32_t v = scratch[w]; scratch[w] = run; run += v; }
    *s_base = run ? g_add((unsigned long long *)count, (unsigned long long)run) : 0;
    if (*s_base + run > od.capacity) g_or(flags + FLAG_OVERFLOW, 1u);
  }
  __syncthreads();
  uint64_t pos = *s_base + scratch[wave] + (x - local);
  for (uint32_t i = b; i < e; ++i) {
    const bool occ = (i < S) ? (keys[i] != KeyTraits<K>::kEmpty) : (*t.sentinel != 0);
    if (!occ) continue;
    if (pos < od.capacity) emit_row<K>(t, s, od, i, i < S ? keys[i] : KeyTraits<K>::kEmpty, pos);
    ++pos;
  }
}

}  // namespace tga
