// Host build of the voxel build's compact Morton key (csrc/morton.hpp), run by
// tests/test_morton.py: against the definition (level i holds bit i of each
// axis that still has bits, x, y, z order) for every width combination up to
// 21 bits per axis, on random and all-ones coordinates, and the 32-bit form
// (spread_morton32) for every combination of at most 32 bits in total.  Exit
// status 0 when every key matches.
#include <cstdint>
#include <cstdio>
#include <random>

#include "../../cuda_ndt_matcher_amd/csrc/morton.hpp"

static uint64_t by_definition(uint32_t rx, uint32_t ry, uint32_t rz, int bx, int by, int bz) {
  uint64_t out = 0;
  int pos = 0;
  for (int i = 0; i < 21; ++i) {
    if (i < bx) out |= (uint64_t)((rx >> i) & 1u) << pos++;
    if (i < by) out |= (uint64_t)((ry >> i) & 1u) << pos++;
    if (i < bz) out |= (uint64_t)((rz >> i) & 1u) << pos++;
  }
  return out;
}

int main() {
  std::mt19937_64 g(7);
  long n = 0, bad = 0;
  for (int bx = 0; bx <= 21; ++bx)
    for (int by = 0; by <= 21; ++by)
      for (int bz = 0; bz <= 21; ++bz) {
        const ndt::MortonLayout L = ndt::morton_layout(bx, by, bz);
        for (int t = 0; t < 64; ++t) {
          auto draw = [&](int b) { return b ? (uint32_t)(g() & ((1ull << b) - 1)) : 0u; };
          uint32_t rx = draw(bx), ry = draw(by), rz = draw(bz);
          if (t == 0) {
            rx = (uint32_t)((1ull << bx) - 1);
            ry = (uint32_t)((1ull << by) - 1);
            rz = (uint32_t)((1ull << bz) - 1);
          }
          ++n;
          const uint64_t want = by_definition(rx, ry, rz, bx, by, bz);
          if (want != ndt::spread_morton(rx, ry, rz, L)) {
            if (bad < 5) std::printf("mismatch at widths %d %d %d\n", bx, by, bz);
            ++bad;
          }
          if (bx + by + bz <= 32 && want != (uint64_t)ndt::spread_morton32(rx, ry, rz, L)) {
            if (bad < 5) std::printf("32-bit mismatch at widths %d %d %d\n", bx, by, bz);
            ++bad;
          }
        }
      }
  std::printf("%ld keys, %ld mismatches\n", n, bad);
  return bad != 0;
}
