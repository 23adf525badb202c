// tga_cpu.cpp -- the host side of CPU + GPU co-processing: the compat
// phases' units of work (a phase-1 tile, a phase-2 merged partition) done on
// a host thread, with the same semantics as the GPU path (u32 keys, u32
// wrap-around SUM as the reference's atomicAdd on u32, T/src/kernel.cu:48,152,
// 334,340; the same partition function; the same chunk rule in phase 2).
#include "tga_host.hpp"

namespace tga {

void CpuAgg32::reset(uint64_t expect) {
  uint64_t cap = 1024;
  while (cap < 2 * expect) cap <<= 1;
  keys_.assign(cap, 0);
  sums_.assign(cap, 0);
  used_.assign(cap, 0);
  mask_ = cap - 1;
  groups_ = 0;
}

void CpuAgg32::grow() {
  std::vector<uint32_t> k(std::move(keys_)), s(std::move(sums_));
  std::vector<uint8_t> u(std::move(used_));
  const uint64_t cap = (mask_ + 1) * 2;
  keys_.assign(cap, 0);
  sums_.assign(cap, 0);
  used_.assign(cap, 0);
  mask_ = cap - 1;
  groups_ = 0;
  for (size_t i = 0; i < k.size(); ++i)
    if (u[i]) add_one(k[i], s[i]);
}

void CpuAgg32::add(const uint32_t *k, const uint32_t *v, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    if (2 * (groups_ + 1) > mask_ + 1) grow();
    add_one(k[i], v[i]);
  }
}

uint64_t CpuAgg32::emit(uint32_t *ok, uint32_t *ov) const {
  uint64_t w = 0;
  for (uint64_t i = 0; i <= mask_; ++i)
    if (used_[i]) {
      ok[w] = keys_[i];
      ov[w] = sums_[i];
      ++w;
    }
  return w;
}

uint32_t compat_partition_host(uint32_t key, uint32_t P) { return (uint32_t)(key_hash((uint64_t)key) >> 32) % P; }

uint64_t cpu_compat_tile(uint32_t *keys, uint32_t *vals, uint64_t m, uint32_t P, uint32_t *offs) {
  CpuAgg32 t;
  t.reset(std::min<uint64_t>(m, 1 << 16));
  t.add(keys, vals, m);
  const uint64_t G = t.groups();
  std::vector<uint32_t> gk(G), gv(G), part(G);
  t.emit(gk.data(), gv.data());
  std::vector<uint32_t> cnt(P + 1, 0);
  for (uint64_t i = 0; i < G; ++i) {
    part[i] = compat_partition_host(gk[i], P);
    ++cnt[part[i]];
  }
  uint32_t run = 0;
  for (uint32_t p = 0; p < P; ++p) {
    offs[p] = run;
    run += cnt[p];
  }
  offs[P] = run;
  std::vector<uint32_t> cur(offs, offs + P);
  for (uint64_t i = 0; i < G; ++i) {
    const uint32_t d = cur[part[i]]++;
    keys[d] = gk[i];
    vals[d] = gv[i];
  }
  return G;
}

tga_status cpu_compat_intra(uint32_t *k, uint32_t *v, uint64_t m, uint64_t min_load, uint64_t max_load,
                            uint64_t *groups) {
  CpuAgg32 t;
  t.reset(std::min<uint64_t>(max_load, std::max<uint64_t>(m, 1)));
  // the reference's chunk loop (intra:213-253): chunks from the back, each
  // as large as the table has room for; overflow once the groups exceed
  // min_load with rows left over
  uint64_t rest = m, already = 0, tail = m;
  while (rest > 0) {
    const uint64_t next = max_load - std::min(already, max_load);
    uint64_t take;
    if (rest <= next) take = rest;
    else {
      if (already > min_load || next == 0) return TGA_ERR_TABLE_OVERFLOW;
      take = next;
    }
    t.add(k + tail - take, v + tail - take, take);
    already = t.groups();
    tail -= take;
    rest -= take;
  }
  *groups = t.emit(k, v);
  return TGA_OK;
}

}  // namespace tga
