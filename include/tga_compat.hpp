// tga_compat.hpp -- the reference operator surface, rebuilt on the C ABI.
//
// Reproduces, by name and parameter list, the three symbols a caller of the
// reference library uses (its only caller is the reference's test/test.cu:134-161):
//   template<typename key_type, typename val_type>
//   void groupby_agg_partition(...)          T/include/group_then_partition.cuh:5-12
//   void groupby_agg_intra_partition(...)    T/include/groupby_agg_intra_partition.cuh:5-13
//   class par_result                          T/include/par_result.cuh:11-82
// (T/ = tiled_groupby_agg_partition/ of the reference repository), plus the
// typedefs of T/include/util.cuh:6-8.
//
// Documented deviations:
//  * errors throw tga::compat_error instead of printing and calling exit(2)
//    (the reference's "hash table overflow", intra:233-236, becomes a
//    compat_error whose status() is TGA_ERR_TABLE_OVERFLOW);
//  * the stream argument of pop2device is a hipStream_t;
//  * the partition of a key is a different hash than cuco's MurmurHash3_32,
//    so partition sizes and the merge grouping differ from the reference;
//    the final (key, sum) rows are the same after sorting by key;
//  * the key 0xffffffff is aggregated (the reference drops it, util.cuh:8).
//
// The library context is process-wide (the reference functions take none);
// the reference library is not reentrant either (gtp:9-11, intra:10-11).
#pragma once

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "tga_groupby.h"

typedef struct ihipStream_t *hipStream_t;

typedef uint32_t key_type;
typedef uint32_t val_type;
const key_type empty_k = 0xffffffff;

namespace tga {

class compat_error : public std::runtime_error {
 public:
  compat_error(tga_status st, const std::string &m) : std::runtime_error(m), st_(st) {}
  tga_status status() const { return st_; }

 private:
  tga_status st_;
};

inline tga_context *compat_context() {
  static tga_context *ctx = [] {
    tga_context *c = nullptr;
    tga_status st = tga_context_create(0, 0, &c);
    if (st != TGA_OK) throw compat_error(st, "tga_context_create failed");
    return c;
  }();
  return ctx;
}

inline void compat_check(tga_status st) {
  if (st != TGA_OK) throw compat_error(st, tga_last_error(compat_context()));
}

}  // namespace tga

// par_result: per-partition list of (keys, vals, length) segments pointing into
// host memory, as T/include/par_result.cuh:11-82.
class par_result {
 public:
  std::vector<key_type *> key;
  std::vector<val_type *> val;
  std::vector<uint32_t> end_indicator;
  uint32_t size;
  std::mutex par_lock;

  par_result() : size(0) {}
  par_result(const par_result &pr) : key(pr.key), val(pr.val), end_indicator(pr.end_indicator), size(pr.size) {}
  par_result &operator=(const par_result &pr) {
    key = pr.key;
    val = pr.val;
    end_indicator = pr.end_indicator;
    size = pr.size;
    return *this;
  }

  // thread-safe append of one segment
  void push(key_type *tile_par_keys, val_type *tile_par_vals, uint32_t n) {
    std::lock_guard<std::mutex> g(par_lock);
    key.push_back(tile_par_keys);
    val.push_back(tile_par_vals);
    end_indicator.push_back(n);
    size += n;
  }

  // Copies the last n rows (popping whole or partial segments from the back)
  // to device memory on `stream`.
  void pop2device(key_type *device_keys, val_type *device_vals, uint32_t n, hipStream_t stream) {
    tga_context *ctx = tga::compat_context();
    (void)stream;  // the copies run on the library stream and complete before return
    while (n != 0) {
      uint32_t slide = end_indicator.back();
      key_type *sk = key.back();
      val_type *sv = val.back();
      if (slide > n) {
        tga::compat_check(tga_memcpy_h2d(ctx, device_keys, sk + slide - n, (uint64_t)n * sizeof(key_type)));
        tga::compat_check(tga_memcpy_h2d(ctx, device_vals, sv + slide - n, (uint64_t)n * sizeof(val_type)));
        end_indicator.back() = slide - n;
        size -= n;
        n = 0;
      } else {
        tga::compat_check(tga_memcpy_h2d(ctx, device_keys, sk, (uint64_t)slide * sizeof(key_type)));
        tga::compat_check(tga_memcpy_h2d(ctx, device_vals, sv, (uint64_t)slide * sizeof(val_type)));
        key.pop_back();
        val.pop_back();
        end_indicator.pop_back();
        size -= slide;
        n -= slide;
        device_keys += slide;
        device_vals += slide;
      }
    }
  }

  void merge(const par_result &pr) {
    key.insert(key.end(), pr.key.begin(), pr.key.end());
    val.insert(val.end(), pr.val.begin(), pr.val.end());
    end_indicator.insert(end_indicator.end(), pr.end_indicator.begin(), pr.end_indicator.end());
    size += pr.size;
  }
};

// Phase 1: tile pre-aggregation + hash partitioning.  The caller's buffers are
// overwritten with the tile results and par_result_vec[p] receives the
// segments of partition p (one per tile), as T/src/group_then_partition.cu:427-528.
template <typename K, typename V>
void groupby_agg_partition(K *host_keys_buffer, V *host_vals_buffer, size_t kv_buffer_len, size_t tile_len, size_t P,
                           std::vector<par_result> &par_result_vec, size_t nstreams) {
  static_assert(sizeof(K) == 4 && sizeof(V) == 4, "the reference instantiates u32 keys and values only");
  tga_context *ctx = tga::compat_context();
  tga_partset *ps = nullptr;
  tga::compat_check(tga_compat_partition(ctx, (uint32_t *)host_keys_buffer, (uint32_t *)host_vals_buffer,
                                         kv_buffer_len, tile_len, P, nstreams, &ps));
  if (par_result_vec.size() < P) par_result_vec.resize(P);
  for (size_t p = 0; p < P; ++p) {
    size_t nseg = 0;
    tga_partset_info(ps, p, &nseg, nullptr);
    for (size_t j = 0; j < nseg; ++j) {
      uint32_t *k = nullptr, *v = nullptr, len = 0;
      tga_partset_segment(ps, p, j, &k, &v, &len);
      par_result_vec[p].push((key_type *)k, (val_type *)v, len);
    }
  }
  tga_partset_destroy(ps);
}

// Phase 2: merge_partition, per-partition aggregation, compaction.  Allocates
// pinned outputs (release with tga_host_free); merged partition i holds
// par_result_kv_num[i] groups from par_kv_begin[i], as
// T/src/groupby_agg_intra_partition.cu:283-411.
inline void groupby_agg_intra_partition(std::vector<par_result> &par_result_vec, key_type *&host_groupby_keys_result,
                                        val_type *&host_agg_vals_result, size_t Capacity, size_t min_load_num,
                                        size_t max_load_num, size_t nstreams, std::vector<size_t> &par_kv_begin,
                                        std::vector<size_t> &par_result_kv_num) {
  tga_context *ctx = tga::compat_context();
  const size_t P = par_result_vec.size();
  tga_partset *ps = nullptr;
  tga::compat_check(tga_partset_create(ctx, P ? P : 1, &ps));
  for (size_t p = 0; p < P; ++p) {
    const par_result &pr = par_result_vec[p];
    for (size_t j = 0; j < pr.key.size(); ++j)
      tga::compat_check(tga_partset_push(ps, p, pr.key[j], pr.val[j], pr.end_indicator[j]));
  }
  std::vector<size_t> begin(P ? P : 1), num(P ? P : 1);
  size_t nm = 0;
  uint32_t *ok = nullptr, *ov = nullptr;
  tga_status st = tga_compat_intra_partition(ctx, ps, Capacity, min_load_num, max_load_num, nstreams, &ok, &ov,
                                             begin.data(), num.data(), &nm);
  tga_partset_destroy(ps);
  tga::compat_check(st);
  host_groupby_keys_result = ok;
  host_agg_vals_result = ov;
  par_result_vec.assign(nm, par_result());
  par_kv_begin.assign(begin.begin(), begin.begin() + nm);
  par_result_kv_num.assign(num.begin(), num.begin() + nm);
}
