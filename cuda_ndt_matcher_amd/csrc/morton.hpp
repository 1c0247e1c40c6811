// Compact Morton keys of the voxel build and the filter's downsampling.
//
// The key interleaves the significant bits of the three box-relative cell
// coordinates: level i holds bit i of each axis that still has bits, in
// x, y, z order, so levels below the smallest width take all three axes,
// then the two wider axes, then the widest alone.  The host derives that
// layout once per build from the widths (morton_layout); the device forms
// each key with the usual magic-number bit spreading (spread_morton), a
// fixed sequence of shifts and masks instead of a loop over the levels.
#pragma once

#include <cstdint>

namespace ndt {

struct MortonLayout {
  int m3;      // levels with all three axes
  int m2;      // end of the levels with two axes (m2 == m3: none)
  int a2, b2;  // the two axes of those levels, in x, y, z order
  int c1;      // the axis of the remaining levels, or -1
  int end1;    // end of those levels (that axis's width)
};

inline MortonLayout morton_layout(int bx, int by, int bz) {
  const int b[3] = {bx, by, bz};
  MortonLayout L{};
  L.m3 = bx < by ? (bx < bz ? bx : bz) : (by < bz ? by : bz);
  int wide[3], k = 0;
  for (int a = 0; a < 3; ++a)
    if (b[a] > L.m3) wide[k++] = a;
  L.m2 = L.m3;
  L.a2 = L.b2 = 0;
  L.c1 = -1;
  L.end1 = L.m3;
  if (k == 2) {
    L.a2 = wide[0];
    L.b2 = wide[1];
    L.m2 = b[wide[0]] < b[wide[1]] ? b[wide[0]] : b[wide[1]];
    if (b[wide[0]] != b[wide[1]]) {
      L.c1 = b[wide[0]] > b[wide[1]] ? wide[0] : wide[1];
      L.end1 = b[L.c1];
    }
  } else if (k == 1) {
    L.c1 = wide[0];
    L.end1 = b[wide[0]];
  }
  return L;
}

// bit i of v (i < 21) -> bit 3i
__host__ __device__ inline uint64_t spread_by3(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// bit i of v (i < 32) -> bit 2i
__host__ __device__ inline uint64_t spread_by2(uint64_t v) {
  v &= 0xFFFFFFFFull;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}

// bit i of v (i < 10) -> bit 3i, in 32-bit arithmetic
__host__ __device__ inline uint32_t spread_by3_32(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// bit i of v (i < 16) -> bit 2i, in 32-bit arithmetic
__host__ __device__ inline uint32_t spread_by2_32(uint32_t v) {
  v &= 0xFFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

__host__ __device__ inline uint32_t low_bits(uint32_t v, int n) { return n >= 32 ? v : (v & ((1u << n) - 1u)); }
__host__ __device__ inline uint32_t axis_of(int a, uint32_t x, uint32_t y, uint32_t z) {
  return a == 0 ? x : (a == 1 ? y : z);
}

// Widths up to 21 bits per axis (the build's limit), 63 bits in all.
__host__ __device__ inline uint64_t spread_morton(uint32_t rx, uint32_t ry, uint32_t rz, const MortonLayout& L) {
  uint64_t out = spread_by3(low_bits(rx, L.m3)) | (spread_by3(low_bits(ry, L.m3)) << 1) |
                 (spread_by3(low_bits(rz, L.m3)) << 2);
  int pos = 3 * L.m3;
  if (L.m2 > L.m3) {
    const int w = L.m2 - L.m3;
    const uint32_t a = low_bits(axis_of(L.a2, rx, ry, rz) >> L.m3, w);
    const uint32_t b = low_bits(axis_of(L.b2, rx, ry, rz) >> L.m3, w);
    out |= (spread_by2(a) | (spread_by2(b) << 1)) << pos;
    pos += 2 * w;
  }
  if (L.c1 >= 0) out |= (uint64_t)low_bits(axis_of(L.c1, rx, ry, rz) >> L.m2, L.end1 - L.m2) << pos;
  return out;
}

// spread_morton for layouts of at most 32 bits in total (then m3 <= 10, the
// two-axis levels end below bit 32 and the one-axis part starts below it), in
// 32-bit arithmetic.
__host__ __device__ inline uint32_t spread_morton32(uint32_t rx, uint32_t ry, uint32_t rz, const MortonLayout& L) {
  uint32_t out = spread_by3_32(low_bits(rx, L.m3)) | (spread_by3_32(low_bits(ry, L.m3)) << 1) |
                 (spread_by3_32(low_bits(rz, L.m3)) << 2);
  int pos = 3 * L.m3;
  if (L.m2 > L.m3) {
    const int w = L.m2 - L.m3;
    const uint32_t a = low_bits(axis_of(L.a2, rx, ry, rz) >> L.m3, w);
    const uint32_t b = low_bits(axis_of(L.b2, rx, ry, rz) >> L.m3, w);
    out |= (spread_by2_32(a) | (spread_by2_32(b) << 1)) << pos;
    pos += 2 * w;
  }
  if (L.c1 >= 0) out |= low_bits(axis_of(L.c1, rx, ry, rz) >> L.m2, L.end1 - L.m2) << pos;
  return out;
}

}  // namespace ndt
