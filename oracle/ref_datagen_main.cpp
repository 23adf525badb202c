// Driver around the reference's workload generator (test infrastructure).
//
// Reproduces the input set-up of the reference driver test/test.cu:92-132 --
// a default-constructed std::default_random_engine, skew 0.9, the reserved
// key 0xffffffff -- and calls generate_various_dist_kv_array from the
// reference's test/include/data_generator.hpp, which is compiled in place from
// the reference checkout (see oracle/Makefile, target `ref`).  Writes
// n u32 keys followed by n u32 values, little endian, to the output file.
//
// usage: ref_datagen <n> <cardinality_percentage> <dist_kind> <out.bin>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "data_generator.hpp"

int main(int argc, char **argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s n pct dist out.bin\n", argv[0]);
    return 2;
  }
  size_t kv_num = std::strtoull(argv[1], nullptr, 10);
  size_t pct = std::strtoull(argv[2], nullptr, 10);
  int dist = std::atoi(argv[3]);
  size_t cardinality = kv_num * pct / 100;  // test/test.cu:111
  std::vector<uint32_t> keys(kv_num), vals(kv_num);
  std::default_random_engine generator;     // test/test.cu:129
  generate_various_dist_kv_array<uint32_t, uint32_t>(keys.data(), vals.data(), cardinality, kv_num,
                                                     0.9, generator, dist, 0xffffffffu);
  FILE *f = std::fopen(argv[4], "wb");
  if (!f) return 1;
  std::fwrite(keys.data(), 4, kv_num, f);
  std::fwrite(vals.data(), 4, kv_num, f);
  std::fclose(f);
  return 0;
}
