// compat_driver.cpp -- the reference driver's call sequence (test/test.cu:134-168
// of the reference repository) compiled against include/tga_compat.hpp.
//
// usage: tga_compat_driver <in.bin> <out.bin> <n> <tile_len> <P>
// in.bin holds n u32 keys then n u32 values (copied into pinned memory like
// test/test.cu:120-121); out.bin receives the G result keys then G sums.
// Prints "result kv num: G" as the reference does (test/test.cu:168); on the
// reference's overflow condition prints "hash table overflow" and exits 2.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tga_compat.hpp"

int main(int argc, char **argv) {
  if (argc != 6) { std::fprintf(stderr, "usage\n"); return 1; }
  size_t kv_num = std::strtoull(argv[3], nullptr, 10);
  size_t tile_len = std::strtoull(argv[4], nullptr, 10);
  size_t P = std::strtoull(argv[5], nullptr, 10);
  size_t nstreams = 4;
  auto Capacity = tile_len;
  size_t min_load_num = Capacity * 3 / 5;
  size_t max_load_num = Capacity * 4 / 5;
  key_type *host_keys = nullptr;
  val_type *host_vals = nullptr;
  if (hipHostMalloc((void **)&host_keys, sizeof(key_type) * kv_num) != hipSuccess ||
      hipHostMalloc((void **)&host_vals, sizeof(val_type) * kv_num) != hipSuccess)
    return 1;
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || std::fread(host_keys, 4, kv_num, f) != kv_num || std::fread(host_vals, 4, kv_num, f) != kv_num) return 1;
  std::fclose(f);
  std::vector<par_result> par_result_vec(P);
  key_type *host_groupby_keys_result = nullptr;
  val_type *host_agg_vals_result = nullptr;
  std::vector<size_t> par_kv_begin, par_result_kv_num;
  try {
    groupby_agg_partition<key_type, val_type>(host_keys, host_vals, kv_num, tile_len, P, par_result_vec, nstreams);
    groupby_agg_intra_partition(par_result_vec, host_groupby_keys_result, host_agg_vals_result, Capacity,
                                min_load_num, max_load_num, nstreams, par_kv_begin, par_result_kv_num);
  } catch (const tga::compat_error &e) {
    if (e.status() == TGA_ERR_TABLE_OVERFLOW) { std::printf("hash table overflow\n"); return 2; }
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  size_t result_kv_num = 0;
  for (size_t i = 0; i < par_result_kv_num.size(); i++) result_kv_num += par_result_kv_num[i];
  std::printf("result kv num: %zu\n", result_kv_num);
  FILE *o = std::fopen(argv[2], "wb");
  if (!o) return 1;
  for (size_t i = 0; i < par_kv_begin.size(); ++i)
    std::fwrite(host_groupby_keys_result + par_kv_begin[i], 4, par_result_kv_num[i], o);
  for (size_t i = 0; i < par_kv_begin.size(); ++i)
    std::fwrite(host_agg_vals_result + par_kv_begin[i], 4, par_result_kv_num[i], o);
  std::fclose(o);
  tga_host_free(host_groupby_keys_result);
  tga_host_free(host_agg_vals_result);
  hipHostFree(host_keys);
  hipHostFree(host_vals);
  return 0;
}
