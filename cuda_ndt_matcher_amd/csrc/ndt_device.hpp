// Device-side building blocks shared by the voxel build, the Newton kernel and
// the scoring kernels.  Written for gfx950 (CDNA4): 64-lane wavefronts, FP64
// VALU at full rate, 160 KiB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "eigen6.hpp"

namespace ndt {

constexpr int kWave = 64;
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int32_t kCoordBias = 1 << 20;  // voxel coordinates must lie in [-2^20, 2^20)
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint64_t kCellTablesMaxCells = 1ull << 26;  // 20 bytes per cell: 1.25 GiB at most

// ---------------------------------------------------------------------------
// Voxel grid as the kernels see it.  Voxel slots are in Morton-key order; a
// slot whose voxel was dropped (too few points, degenerate covariance) is
// never inserted in the hash and so never visited.
//
// Lookup structures:
//  - an open-addressing hash keyed by the packed voxel coordinate, 32-byte
//    entries {key, slot, mean} (linear probing, load <= 0.5), so one probe
//    returns the mean needed for the radius test;
//  - an occupancy bitmap over the voxel box padded by two cells on every
//    side (one bit per cell) when that box has fewer than 2^32 cells; it
//    answers "is this cell occupied" for the 3x3x3 block of a query with nine
//    independent word loads, so the hash is probed only for occupied cells;
//  - cell tables over the same padded box when it has at most 2^26 cells:
//    cell_nbr holds, per cell, the 27-bit occupancy of its 3x3x3 block and
//    cell_ent, per occupied cell, the voxel mean and slot in 16 bytes (cells
//    that cell_nbr marks empty are never read, so they are left unwritten).
//    A query costs one load for its block, one 16-byte load per occupied
//    candidate cell (enough for the radius test) and the record loads of
//    the accepted ones, with no probing.
struct __align__(32) HashEntry {
  unsigned long long key;
  uint32_t slot;
  float mx, my, mz;
  uint32_t pad[2];
};

// One voxel's record: {mx,my,mz,C00},{C01,C02,C10,C11},{C12,C20,C21,C22}
// (mean and inverse covariance, row-major).
struct VoxelRecord {
  float4 m0, m1, m2;
};

struct GridView {
  const float4* vox;        // 3 x float4 per slot: {mx,my,mz,C00},{C01,C02,C10,C11},{C12,C20,C21,C22}
  const float* axis;        // 3 floats per slot: principal axis (point-to-plane metric)
  const HashEntry* table;   // capacity is a power of two
  uint32_t table_mask;      // capacity - 1
  const uint32_t* bitmap;   // null when the padded box is too large
  const uint32_t* cell_nbr;   // null when the cell tables are not built
  const float4* cell_ent;     // per occupied cell {mx, my, mz, slot bits}
  int32_t borg[3];          // padded box origin (cmin - 2)
  uint32_t bdim[3];         // padded box extent (cmax - cmin + 5)
  int32_t cmin[3];          // a box holding every occupied voxel (all points' voxels, or the valid ones)
  int32_t cmax[3];
  float res;
};

__host__ __device__ inline uint64_t pack_coord(int32_t x, int32_t y, int32_t z) {
  return (uint64_t)(uint32_t)(x + kCoordBias) | ((uint64_t)(uint32_t)(y + kCoordBias) << 21) |
         ((uint64_t)(uint32_t)(z + kCoordBias) << 42);
}

// murmur3 finaliser: spreads the packed coordinate over the table.
__host__ __device__ inline uint64_t mix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__device__ inline bool in_box(const GridView& g, int32_t x, int32_t y, int32_t z) {
  return x >= g.cmin[0] && x <= g.cmax[0] && y >= g.cmin[1] && y <= g.cmax[1] && z >= g.cmin[2] && z <= g.cmax[2];
}

// Entry index of cell (x, y, z), or -1.
__device__ inline int32_t hash_find(const GridView& g, int32_t x, int32_t y, int32_t z) {
  if (!in_box(g, x, y, z)) return -1;
  const uint64_t key = pack_coord(x, y, z);
  uint32_t h = (uint32_t)mix64(key) & g.table_mask;
#pragma unroll 1
  for (;;) {
    const unsigned long long k = g.table[h].key;
    if (k == key) return (int32_t)h;
    if (k == kEmptyKey) return -1;
    h = (h + 1) & g.table_mask;
  }
}

// Cell index within the 3x3x3 block: c = (dx+1)*9 + (dy+1)*3 + (dz+1).
__device__ inline void cell_offset(int c, int& dx, int& dy, int& dz) {
  dx = c / 9 - 1;
  dy = (c / 3) % 3 - 1;
  dz = c % 3 - 1;
}

// Occupancy of the 3x3x3 block around (kx, ky, kz) as a 27-bit mask.  The
// centre must lie in [cmin-1, cmax+1] so the block is inside the padded box.
__device__ inline uint32_t occupancy27(const GridView& g, int32_t kx, int32_t ky, int32_t kz) {
  uint32_t m = 0;
  const uint64_t x0 = (uint64_t)(kx - 1 - g.borg[0]);
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int dy = r / 3 - 1, dz = r % 3 - 1;
    const uint64_t lin = x0 + (uint64_t)g.bdim[0] * ((uint64_t)(ky + dy - g.borg[1]) +
                                                      (uint64_t)g.bdim[1] * (uint64_t)(kz + dz - g.borg[2]));
    const uint64_t w = lin >> 5;
    const uint64_t two = (uint64_t)g.bitmap[w] | ((uint64_t)g.bitmap[w + 1] << 32);
    const uint32_t b3 = (uint32_t)(two >> (lin & 31)) & 7u;  // bit0: x-1, bit1: x, bit2: x+1
    m |= ((b3 & 1u) << r) | (((b3 >> 1) & 1u) << (9 + r)) | (((b3 >> 2) & 1u) << (18 + r));
  }
  return m;
}

// Visits the voxels of the cells set in `cand` (27-bit mask around k),
// probing two cells at a time so their table loads overlap.
// accept(mx, my, mz) decides from the voxel mean whether to visit;
// visit(slot, record).
template <class A, class F>
__device__ inline void probe_cells(const GridView& g, int32_t kx, int32_t ky, int32_t kz, uint32_t cand, A&& accept,
                                   F&& visit) {
#pragma unroll 1
  while (cand) {
    const int c0 = __ffs(cand) - 1;
    cand &= cand - 1;
    int c1 = -1;
    if (cand) {
      c1 = __ffs(cand) - 1;
      cand &= cand - 1;
    }
    int dx, dy, dz;
    cell_offset(c0, dx, dy, dz);
    const uint64_t k0 = pack_coord(kx + dx, ky + dy, kz + dz);
    uint32_t h0 = (uint32_t)mix64(k0) & g.table_mask;
    uint64_t k1 = 0;
    uint32_t h1 = 0;
    if (c1 >= 0) {
      cell_offset(c1, dx, dy, dz);
      k1 = pack_coord(kx + dx, ky + dy, kz + dz);
      h1 = (uint32_t)mix64(k1) & g.table_mask;
    }
    HashEntry e0 = g.table[h0];
    HashEntry e1;
    if (c1 >= 0) e1 = g.table[h1];
#pragma unroll 1
    while (e0.key != k0 && e0.key != kEmptyKey) {
      h0 = (h0 + 1) & g.table_mask;
      e0 = g.table[h0];
    }
    if (e0.key == k0 && accept(e0.mx, e0.my, e0.mz)) {
      const VoxelRecord r{g.vox[3 * e0.slot], g.vox[3 * e0.slot + 1], g.vox[3 * e0.slot + 2]};
      visit((int32_t)e0.slot, r);
    }
    if (c1 >= 0) {
#pragma unroll 1
      while (e1.key != k1 && e1.key != kEmptyKey) {
        h1 = (h1 + 1) & g.table_mask;
        e1 = g.table[h1];
      }
      if (e1.key == k1 && accept(e1.mx, e1.my, e1.mz)) {
        const VoxelRecord r{g.vox[3 * e1.slot], g.vox[3 * e1.slot + 1], g.vox[3 * e1.slot + 2]};
        visit((int32_t)e1.slot, r);
      }
    }
  }
}

// Index of cell (x, y, z) in the padded box (x fastest).
__device__ inline uint32_t box_index(const GridView& g, int32_t x, int32_t y, int32_t z) {
  return (uint32_t)(x - g.borg[0]) + g.bdim[0] * ((uint32_t)(y - g.borg[1]) + g.bdim[1] * (uint32_t)(z - g.borg[2]));
}

// Shape of one cell-table load stage: E candidate entries in flight, then
// the records of the accepted ones in groups of R.  The Newton kernel runs one
// wave per SIMD whatever the shape, so it takes the wide one (fewer dependent
// rounds); the scoring kernels keep two waves per SIMD with the narrow one.
template <int E, int R>
struct ProbeShape {
  static constexpr int kEnt = E, kRec = R;
};
using ProbeWide = ProbeShape<12, 6>;
using ProbeNarrow = ProbeShape<8, 8>;

// Cell-table visit of the candidate cells around the cell with box index
// `lin`, P::kEnt at a time: their 16-byte entries (mean, slot) are loaded
// together, then the records of the accepted ones, P::kRec at a time (a
// second group only when a lane accepted a candidate past its first kRec).
// Candidates are visited in increasing cell order, as in probe_cells.
template <class P = ProbeWide, class A, class F>
__device__ inline void probe_dense(const GridView& g, uint32_t lin, uint32_t cand, A&& accept, F&& visit) {
  constexpr int kEnt = P::kEnt, kRec = P::kRec;
  const int32_t sy = (int32_t)g.bdim[0], sz = (int32_t)(g.bdim[0] * g.bdim[1]);
#pragma unroll 1
  while (cand) {
    float4 ent[kEnt];
    bool ok[kEnt];
#pragma unroll
    for (int b = 0; b < kEnt; ++b) {
      ok[b] = false;
      if (cand) {
        const int c = __ffs(cand) - 1;
        cand &= cand - 1;
        const int32_t off = (c / 9 - 1) + (c / 3 % 3 - 1) * sy + (c % 3 - 1) * sz;
        ent[b] = g.cell_ent[(uint32_t)((int32_t)lin + off)];
        ok[b] = true;
      }
    }
#pragma unroll
    for (int b = 0; b < kEnt; ++b) ok[b] = ok[b] && accept(ent[b].x, ent[b].y, ent[b].z);
    if constexpr (kRec >= kEnt) {
      // One group: every accepted candidate's record in flight at once.
      float4 m0[kEnt], m1[kEnt], m2[kEnt];
#pragma unroll
      for (int b = 0; b < kEnt; ++b)
        if (ok[b]) {
          const uint32_t sl = __float_as_uint(ent[b].w);
          m0[b] = g.vox[3 * sl];
          m1[b] = g.vox[3 * sl + 1];
          m2[b] = g.vox[3 * sl + 2];
        }
#pragma unroll
      for (int b = 0; b < kEnt; ++b)
        if (ok[b]) visit((int32_t)__float_as_uint(ent[b].w), VoxelRecord{m0[b], m1[b], m2[b]});
    } else {
      uint32_t slot[kEnt];
#pragma unroll
      for (int b = 0; b < kEnt; ++b) slot[b] = __float_as_uint(ent[b].w);
      auto records = [&](auto h) {
        constexpr int o = decltype(h)::value * kRec;
        float4 m0[kRec], m1[kRec], m2[kRec];
#pragma unroll
        for (int b = 0; b < kRec; ++b)
          if (ok[o + b]) {
            m0[b] = g.vox[3 * slot[o + b]];
            m1[b] = g.vox[3 * slot[o + b] + 1];
            m2[b] = g.vox[3 * slot[o + b] + 2];
          }
#pragma unroll
        for (int b = 0; b < kRec; ++b)
          if (ok[o + b]) visit((int32_t)slot[o + b], VoxelRecord{m0[b], m1[b], m2[b]});
      };
      records(std::integral_constant<int, 0>{});
      bool late = false;
#pragma unroll
      for (int b = kRec; b < kEnt; ++b) late |= ok[b];
      if (__any(late)) records(std::integral_constant<int, 1>{});
    }
  }
}

// Visits the voxels of the cells in `mask` (27-bit, around k) that exist,
// through the cell tables when they are built and otherwise through the
// bitmap and the hash.  Nothing is visited when no occupied cell lies within
// one cell of k.
template <class P = ProbeWide, class A, class F>
__device__ inline void probe_block(const GridView& g, int32_t kx, int32_t ky, int32_t kz, uint32_t mask, A&& accept,
                                   F&& visit) {
  if (kx < g.cmin[0] - 1 || kx > g.cmax[0] + 1 || ky < g.cmin[1] - 1 || ky > g.cmax[1] + 1 ||
      kz < g.cmin[2] - 1 || kz > g.cmax[2] + 1)
    return;
  if (g.cell_ent) {
    const uint32_t lin = box_index(g, kx, ky, kz);
    probe_dense<P>(g, lin, mask & g.cell_nbr[lin], accept, visit);
    return;
  }
  if (g.bitmap) mask &= occupancy27(g, kx, ky, kz);
  if (mask) probe_cells(g, kx, ky, kz, mask, accept, visit);
}

// Rust `f32 as i32` / `f64 as i32`: saturating, NaN -> 0.
__host__ __device__ inline int32_t sat_i32(float v) {
  if (v != v) return 0;
  if (v >= 2147483647.0f) return 2147483647;
  if (v <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)v;
}
__host__ __device__ inline int32_t sat_i32d(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (int32_t)0x80000000u;
  return (int32_t)v;
}

// VoxelCoord::from_point (voxel_grid/types.rs:206-212): floor of the f32
// quotient.  HIP's f32 division is correctly rounded by default.
__host__ __device__ inline int32_t voxel_key(float p, float res) { return sat_i32(floorf(p / res)); }

// kiddo's SquaredEuclidean in f32: ((dx*dx) + dy*dy) + dz*dz, each operation
// rounded separately (no fused multiply-add).
__device__ inline float dist2_f32(float qx, float qy, float qz, float mx, float my, float mz) {
  const float dx = __fsub_rn(qx, mx);
  const float dy = __fsub_rn(qy, my);
  const float dz = __fsub_rn(qz, mz);
  float s = __fmul_rn(dx, dx);
  s = __fadd_rn(s, __fmul_rn(dy, dy));
  s = __fadd_rn(s, __fmul_rn(dz, dz));
  return s;
}

// ---------------------------------------------------------------------------
// Neighbour enumeration.  visit(slot, record) is called for each
// neighbouring voxel.
//
// Radius mode: every voxel whose f32 mean is within `res` of q (strict <).
// A voxel's mean lies in the closed box of its own cell, so a mean within
// res of q belongs to a cell whose index differs from floor(q/res) by at most
// one on each axis; floor is taken of the f64 quotient, which is exact for
// f32 inputs.  Cells whose box is farther than res from q are skipped.
template <class P = ProbeWide, class F>
__device__ inline void for_each_radius(const GridView& g, float qx, float qy, float qz, F&& visit) {
  const double res = (double)g.res;
  const int32_t kx = sat_i32d(floor((double)qx / res));
  const int32_t ky = sat_i32d(floor((double)qy / res));
  const int32_t kz = sat_i32d(floor((double)qz / res));
  // Squared distance from q to the low/high faces of its own cell, per axis
  // (index 0: the cell below, 1: own cell, 2: the cell above).
  double sx[3], sy[3], sz[3];
  {
    const double lx = (double)qx - kx * res, hx = (kx + 1) * res - (double)qx;
    const double ly = (double)qy - ky * res, hy = (ky + 1) * res - (double)qy;
    const double lz = (double)qz - kz * res, hz = (kz + 1) * res - (double)qz;
    sx[0] = lx * lx; sx[1] = 0.0; sx[2] = hx * hx;
    sy[0] = ly * ly; sy[1] = 0.0; sy[2] = hy * hy;
    sz[0] = lz * lz; sz[1] = 0.0; sz[2] = hz * hz;
  }
  const double lim = res * res * (1.0 + 1e-6);
  uint32_t mask = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double sxy = sx[a] + sy[b];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (sxy + sz[c] <= lim) mask |= 1u << (9 * a + 3 * b + c);
    }
  const float r2 = __fmul_rn(g.res, g.res);
  probe_block<P>(
      g, kx, ky, kz, mask, [&](float mx, float my, float mz) { return dist2_f32(qx, qy, qz, mx, my, mz) < r2; },
      visit);
}

// Autoware pclomp direct searches: cell = floor(q/res) in f32; DIRECT1 the
// cell, DIRECT7 the cell and its 6 faces, DIRECT26 the 26 cells around it
// (PCL getAllNeighborCellIndices, which leaves the centre out).
__host__ __device__ constexpr uint32_t direct_mask(int mode) {
  // bit c = (dx+1)*9 + (dy+1)*3 + (dz+1); centre 13, faces 4, 22, 10, 16, 12, 14
  return mode == 3 ? (1u << 13)
                   : (mode == 2 ? ((1u << 13) | (1u << 22) | (1u << 4) | (1u << 16) | (1u << 10) | (1u << 14) |
                                   (1u << 12))
                                : (((1u << 27) - 1) & ~(1u << 13)));
}

template <class P = ProbeWide, class F>
__device__ inline void for_each_direct(const GridView& g, int mode, float qx, float qy, float qz, F&& visit) {
  const int32_t kx = voxel_key(qx, g.res), ky = voxel_key(qy, g.res), kz = voxel_key(qz, g.res);
  probe_block<P>(g, kx, ky, kz, direct_mask(mode), [](float, float, float) { return true; }, visit);
}

template <class P = ProbeWide, class F>
__device__ inline void for_each_neighbor(const GridView& g, int mode, float qx, float qy, float qz, F&& visit) {
  if (mode == 0)
    for_each_radius<P>(g, qx, qy, qz, visit);
  else
    for_each_direct<P>(g, mode, qx, qy, qz, visit);
}

// compute_max_neighbor_score's cell set (scoring/nvtl.rs:160-190): 3x3x3
// cells around floor(x * (1/res)) in f64, no distance test.
template <class P = ProbeWide, class F>
__device__ inline void for_each_cell27(const GridView& g, double x, double y, double z, F&& visit) {
  const double inv = 1.0 / (double)g.res;
  const int32_t vx = sat_i32d(floor(x * inv)), vy = sat_i32d(floor(y * inv)), vz = sat_i32d(floor(z * inv));
  probe_block<P>(g, vx, vy, vz, (1u << 27) - 1, [](float, float, float) { return true; }, visit);
}

// ---------------------------------------------------------------------------
// Pose-dependent tables, computed once per pose (by one thread) and shared
// through LDS.
struct PoseTables {
  double t[3];
  double R[9];      // Rx(roll) Ry(pitch) Rz(yaw), exact sin/cos (derivatives/cpu.rs:339-372)
  double j[8][3];   // angular Jacobian with the |angle| < 1e-5 snap (angular.rs:37-99)
  double h[15][3];  // angular Hessian (angular.rs:104-181)
};

// The pose tables after t (R, j, h in PoseTables order) as signed products of
// the angles' sin/cos: entry = k0 f(a0) f(b0) f(c0) + k1 f(a1) f(b1) f(c1), with
// f(0..5) = sin r, cos r, sin p, cos p, sin y, cos y (exact: the rotation,
// derivatives/cpu.rs:339-372), f(6..11) the same with the snap of
// angular.rs:39-55 (|angle| < 1e-5: sin 0, cos 1; the angular Jacobian and
// Hessian, angular.rs:58-181) and f(12) = 1.
struct PoseTerm {
  int8_t k;
  uint8_t a, b, c;
};
struct PoseEntry {
  PoseTerm t[2];
};
constexpr int kPoseEntries = 9 + 24 + 45;
static __constant__ const PoseEntry kPoseTable[kPoseEntries] = {
    {{{1, 3, 5, 12}, {0, 12, 12, 12}}},  // R[0] = cp * cy
    {{{-1, 3, 4, 12}, {0, 12, 12, 12}}},  // R[1] = -cp * sy
    {{{1, 2, 12, 12}, {0, 12, 12, 12}}},  // R[2] = sp
    {{{1, 0, 2, 5}, {1, 1, 4, 12}}},  // R[3] = sr * sp * cy + cr * sy
    {{{1, 1, 5, 12}, {-1, 0, 2, 4}}},  // R[4] = cr * cy - sr * sp * sy
    {{{-1, 0, 3, 12}, {0, 12, 12, 12}}},  // R[5] = -sr * cp
    {{{1, 0, 4, 12}, {-1, 1, 2, 5}}},  // R[6] = sr * sy - cr * sp * cy
    {{{1, 1, 2, 4}, {1, 0, 5, 12}}},  // R[7] = cr * sp * sy + sr * cy
    {{{1, 1, 3, 12}, {0, 12, 12, 12}}},  // R[8] = cr * cp
    {{{-1, 6, 10, 12}, {1, 7, 8, 11}}},  // j[0][0] = -sx * sz + cx * sy * cz
    {{{-1, 6, 11, 12}, {-1, 7, 8, 10}}},  // j[0][1] = -sx * cz - cx * sy * sz
    {{{-1, 7, 9, 12}, {0, 12, 12, 12}}},  // j[0][2] = -cx * cy
    {{{1, 7, 10, 12}, {1, 6, 8, 11}}},  // j[1][0] = cx * sz + sx * sy * cz
    {{{1, 7, 11, 12}, {-1, 6, 8, 10}}},  // j[1][1] = cx * cz - sx * sy * sz
    {{{-1, 6, 9, 12}, {0, 12, 12, 12}}},  // j[1][2] = -sx * cy
    {{{-1, 8, 11, 12}, {0, 12, 12, 12}}},  // j[2][0] = -sy * cz
    {{{1, 8, 10, 12}, {0, 12, 12, 12}}},  // j[2][1] = sy * sz
    {{{1, 9, 12, 12}, {0, 12, 12, 12}}},  // j[2][2] = cy
    {{{1, 6, 9, 11}, {0, 12, 12, 12}}},  // j[3][0] = sx * cy * cz
    {{{-1, 6, 9, 10}, {0, 12, 12, 12}}},  // j[3][1] = -sx * cy * sz
    {{{1, 6, 8, 12}, {0, 12, 12, 12}}},  // j[3][2] = sx * sy
    {{{-1, 7, 9, 11}, {0, 12, 12, 12}}},  // j[4][0] = -cx * cy * cz
    {{{1, 7, 9, 10}, {0, 12, 12, 12}}},  // j[4][1] = cx * cy * sz
    {{{-1, 7, 8, 12}, {0, 12, 12, 12}}},  // j[4][2] = -cx * sy
    {{{-1, 9, 10, 12}, {0, 12, 12, 12}}},  // j[5][0] = -cy * sz
    {{{-1, 9, 11, 12}, {0, 12, 12, 12}}},  // j[5][1] = -cy * cz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // j[5][2] = 0.0
    {{{1, 7, 11, 12}, {-1, 6, 8, 10}}},  // j[6][0] = cx * cz - sx * sy * sz
    {{{-1, 7, 10, 12}, {-1, 6, 8, 11}}},  // j[6][1] = -cx * sz - sx * sy * cz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // j[6][2] = 0.0
    {{{1, 6, 11, 12}, {1, 7, 8, 10}}},  // j[7][0] = sx * cz + cx * sy * sz
    {{{1, 7, 8, 11}, {-1, 6, 10, 12}}},  // j[7][1] = cx * sy * cz - sx * sz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // j[7][2] = 0.0
    {{{-1, 7, 10, 12}, {-1, 6, 8, 11}}},  // h[0][0] = -cx * sz - sx * sy * cz
    {{{-1, 7, 11, 12}, {1, 6, 8, 10}}},  // h[0][1] = -cx * cz + sx * sy * sz
    {{{1, 6, 9, 12}, {0, 12, 12, 12}}},  // h[0][2] = sx * cy
    {{{-1, 6, 10, 12}, {1, 7, 8, 11}}},  // h[1][0] = -sx * sz + cx * sy * cz
    {{{-1, 7, 8, 10}, {-1, 6, 11, 12}}},  // h[1][1] = -cx * sy * sz - sx * cz
    {{{-1, 7, 9, 12}, {0, 12, 12, 12}}},  // h[1][2] = -cx * cy
    {{{1, 7, 9, 11}, {0, 12, 12, 12}}},  // h[2][0] = cx * cy * cz
    {{{-1, 7, 9, 10}, {0, 12, 12, 12}}},  // h[2][1] = -cx * cy * sz
    {{{1, 7, 8, 12}, {0, 12, 12, 12}}},  // h[2][2] = cx * sy
    {{{1, 6, 9, 11}, {0, 12, 12, 12}}},  // h[3][0] = sx * cy * cz
    {{{-1, 6, 9, 10}, {0, 12, 12, 12}}},  // h[3][1] = -sx * cy * sz
    {{{1, 6, 8, 12}, {0, 12, 12, 12}}},  // h[3][2] = sx * sy
    {{{-1, 6, 11, 12}, {-1, 7, 8, 10}}},  // h[4][0] = -sx * cz - cx * sy * sz
    {{{1, 6, 10, 12}, {-1, 7, 8, 11}}},  // h[4][1] = sx * sz - cx * sy * cz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[4][2] = 0.0
    {{{1, 7, 11, 12}, {-1, 6, 8, 10}}},  // h[5][0] = cx * cz - sx * sy * sz
    {{{-1, 6, 8, 11}, {-1, 7, 10, 12}}},  // h[5][1] = -sx * sy * cz - cx * sz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[5][2] = 0.0
    {{{-1, 9, 11, 12}, {0, 12, 12, 12}}},  // h[6][0] = -cy * cz
    {{{1, 9, 10, 12}, {0, 12, 12, 12}}},  // h[6][1] = cy * sz
    {{{1, 8, 12, 12}, {0, 12, 12, 12}}},  // h[6][2] = sy
    {{{-1, 6, 8, 11}, {0, 12, 12, 12}}},  // h[7][0] = -sx * sy * cz
    {{{1, 6, 8, 10}, {0, 12, 12, 12}}},  // h[7][1] = sx * sy * sz
    {{{1, 6, 9, 12}, {0, 12, 12, 12}}},  // h[7][2] = sx * cy
    {{{1, 7, 8, 11}, {0, 12, 12, 12}}},  // h[8][0] = cx * sy * cz
    {{{-1, 7, 8, 10}, {0, 12, 12, 12}}},  // h[8][1] = -cx * sy * sz
    {{{-1, 7, 9, 12}, {0, 12, 12, 12}}},  // h[8][2] = -cx * cy
    {{{1, 8, 10, 12}, {0, 12, 12, 12}}},  // h[9][0] = sy * sz
    {{{1, 8, 11, 12}, {0, 12, 12, 12}}},  // h[9][1] = sy * cz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[9][2] = 0.0
    {{{-1, 6, 9, 10}, {0, 12, 12, 12}}},  // h[10][0] = -sx * cy * sz
    {{{-1, 6, 9, 11}, {0, 12, 12, 12}}},  // h[10][1] = -sx * cy * cz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[10][2] = 0.0
    {{{1, 7, 9, 10}, {0, 12, 12, 12}}},  // h[11][0] = cx * cy * sz
    {{{1, 7, 9, 11}, {0, 12, 12, 12}}},  // h[11][1] = cx * cy * cz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[11][2] = 0.0
    {{{-1, 9, 11, 12}, {0, 12, 12, 12}}},  // h[12][0] = -cy * cz
    {{{1, 9, 10, 12}, {0, 12, 12, 12}}},  // h[12][1] = cy * sz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[12][2] = 0.0
    {{{-1, 7, 10, 12}, {-1, 6, 8, 11}}},  // h[13][0] = -cx * sz - sx * sy * cz
    {{{-1, 7, 11, 12}, {1, 6, 8, 10}}},  // h[13][1] = -cx * cz + sx * sy * sz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[13][2] = 0.0
    {{{-1, 6, 10, 12}, {1, 7, 8, 11}}},  // h[14][0] = -sx * sz + cx * sy * cz
    {{{-1, 7, 8, 10}, {-1, 6, 11, 12}}},  // h[14][1] = -cx * sy * sz - sx * cz
    {{{0, 12, 12, 12}, {0, 12, 12, 12}}},  // h[14][2] = 0.0
};

// nalgebra UnitQuaternion::from_euler_angles (w, i, j, k), used by the
// Isometry3 that scoring transforms points with (scoring/nvtl.rs:231).
__host__ __device__ inline void quat_from_euler(double roll, double pitch, double yaw, double q[4]) {
  const double sr = sin(roll * 0.5), cr = cos(roll * 0.5);
  const double sp = sin(pitch * 0.5), cp = cos(pitch * 0.5);
  const double sy = sin(yaw * 0.5), cy = cos(yaw * 0.5);
  q[0] = cr * cp * cy + sr * sp * sy;
  q[1] = sr * cp * cy - cr * sp * sy;
  q[2] = cr * sp * cy + sr * cp * sy;
  q[3] = cr * cp * sy - sr * sp * cy;
}

// Isometry3 * Point3: v + w t + u x t with t = 2 u x v, then + translation.
__device__ inline void iso_transform(const double q[4], const double t[3], double px, double py, double pz,
                                     double* out) {
  const double ux = q[1], uy = q[2], uz = q[3], w = q[0];
  const double tx = (uy * pz - uz * py) * 2.0;
  const double ty = (uz * px - ux * pz) * 2.0;
  const double tz = (ux * py - uy * px) * 2.0;
  out[0] = tx * w + (uy * tz - uz * ty) + px + t[0];
  out[1] = ty * w + (uz * tx - ux * tz) + py + t[1];
  out[2] = tz * w + (ux * ty - uy * tx) + pz + t[2];
}

// ---------------------------------------------------------------------------
// Accumulator of one derivative pass: score, gradient, lower triangle of the
// Hessian (21, row-major i >= j), pair count.
constexpr int kAccVals = 28;  // score + 6 + 21
constexpr int kAccSlots = 32; // + correspondences (as a double), padded to a power of two

__device__ inline int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

// ---------------------------------------------------------------------------
// Wavefront / workgroup reductions (f64).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Wave totals of N values per lane (N a power of two, at most 32).  Each
// halving step keeps half of a lane's values and trades the other half with
// the lane `o` away, so the wave spends N - 1 + log2(64 / N) shuffles instead
// of 6 N.  On return v[0] holds, in every lane, the wave total of value
// lane / (64 / N).
template <int N>
__device__ inline void wave_sums(double (&v)[N]) {
  static_assert(N >= 1 && N <= 32 && (N & (N - 1)) == 0, "N must be a power of two <= 32");
  constexpr int kSteps = N >= 32 ? 5 : N >= 16 ? 4 : N >= 8 ? 3 : N >= 4 ? 2 : N >= 2 ? 1 : 0;
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int k = 0; k < kSteps; ++k) {
    const int h = N >> (k + 1), o = (kWave / 2) >> k;
    const bool up = (lane & o) != 0;
#pragma unroll
    for (int i = 0; i < h; ++i) {
      const double send = up ? v[i] : v[i + h];
      const double keep = up ? v[i + h] : v[i];
      v[i] = keep + __shfl_xor(send, o, kWave);
    }
  }
#pragma unroll
  for (int o = (kWave / 2) >> kSteps; o >= 1; o >>= 1) v[0] += __shfl_xor(v[0], o, kWave);
}

// Workgroup totals of the first `count` of N values per thread: thread
// t < count returns total t (fixed summation order; other threads return 0).
// lds must hold (BLOCK / 64) * N doubles; the caller synchronises before
// reusing it.
template <int BLOCK, int N>
__device__ inline double block_sums(double (&v)[N], int count, double* lds) {
  wave_sums<N>(v);
  constexpr int kLanesPer = kWave / N;
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane % kLanesPer == 0 && lane / kLanesPer < count) lds[wave * N + lane / kLanesPer] = v[0];
  __syncthreads();
  double s = 0.0;
  if ((int)threadIdx.x < count)
#pragma unroll
    for (int w = 0; w < BLOCK / kWave; ++w) s += lds[w * N + threadIdx.x];
  return s;
}

// Workgroup totals of the first `count` values per thread through LDS, for
// wide accumulators (the shuffle tree above would hold every value twice in
// registers).  Each thread stores its values as column t of a value-major
// table; then 8 threads per value sum 32 table entries each (stride 8) and
// combine with three shuffles.  out[0..count) receives the totals (LDS,
// synchronised on return).  red must hold count * (BLOCK + 8) doubles; it may
// alias LDS that other threads were still reading, hence the barrier first.
template <int BLOCK>
__device__ inline void block_sums_lds(const double* v, int count, double* red, double* out) {
  static_assert(BLOCK == 256, "8 threads x 32 entries per value");
  constexpr int kRow = BLOCK + 8;  // padding spreads the value rows over the banks
  const int t = threadIdx.x;
  __syncthreads();
  for (int i = 0; i < count; ++i) red[i * kRow + t] = v[i];
  __syncthreads();
  const int val = t >> 3, part = t & 7;
  double s = 0.0;
  if (val < count) {
    const double* row = red + val * kRow + part;
#pragma unroll
    for (int k = 0; k < BLOCK / 8; ++k) s += row[8 * k];
  }
  s += __shfl_xor(s, 1, kWave);
  s += __shfl_xor(s, 2, kWave);
  s += __shfl_xor(s, 4, kWave);
  if (part == 0 && val < count) out[val] = s;
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Workgroup-cooperative 6x6 Newton step.
//
// newton_step_negative_definite (optimization/newton.rs:160-176): when every
// eigenvalue of H is negative the reference solves H d = -g with an SVD
// pseudo-inverse (tolerance 1e-10); otherwise it clamps eigenvalues above
// -1e-3 to -1e-3 and solves with the rebuilt matrix.  Two waves work at once:
//  - wave 1 (one lane) factors -H - 1e-10 I and -H (Cholesky).  Success means
//    every eigenvalue of H is below -1e-10, so the pseudo-inverse is the
//    plain inverse and a Cholesky solve of (-H) d = g gives the same step;
//  - wave 0 decomposes H by Householder tridiagonalisation and implicit QL
//    (eigen6_ql): every lane runs the same scalar recurrence, carries
//    y = Q^T (-g) along, and lane r < 6 accumulates row r of the eigenvector
//    matrix Q.  The step in the eigenbasis with the reference's clamp is then
//    d_r = sum_k Q[r][k] y_k / lambda'_k, formed by lane r without any
//    cross-lane reduction.
// The choice between the two depends on H alone, so every workgroup of a
// slot takes the same one.
struct SolveScratch {
  double dc[6];    // Cholesky step
  int sweeps;      // QL steps taken (profiling)
  int solves;      // calls that completed the eigendecomposition (profiling)
  int fast;        // set by wave 1 when the Cholesky step exists
  int ok;
};

// The device side of eigen6_ql's arithmetic policy.
struct DeviceEigenMath {
  __device__ static double rsqrt(double x) { return refine_rsqrt(x, __builtin_amdgcn_rsq(x)); }
  __device__ static double rcp(double x) { return refine_rcp(x, __builtin_amdgcn_rcp(x)); }
  __device__ static int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
  __device__ static double uni(double x) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  }
};

// Cholesky of -H - shift I (lower triangle of H read) in registers and, when
// it exists, the solve (-H - shift I) dc = g.  Returns false if -H - shift I
// is not positive definite.
__device__ inline bool neg_cholesky6_solve(const double* H, double shift, const double* g, double dc[6]) {
  double L[6][6], il[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double d = -H[j * 6 + j] - shift;
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
    if (!(d > 0.0)) return false;
    const double t = DeviceEigenMath::rsqrt(d);
    L[j][j] = d * t;
    il[j] = t;
#pragma unroll
    for (int i = j + 1; i < 6; ++i) {
      double s = -H[i * 6 + j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      L[i][j] = s * t;
    }
  }
  double y[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = g[i];
#pragma unroll
    for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
    y[i] = s * il[i];
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) s -= L[k][i] * dc[k];
    dc[i] = s * il[i];
  }
  return true;
}

// Orders LDS accesses of the lanes of one wave (no cross-wave barrier).
__device__ inline void wave_lds_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup"); }

// The pose tables by all lanes of one wave: lanes 0-2 evaluate the three
// sin/cos pairs at once (each is a long dependent chain), lanes 6-12 form the
// snapped values, then every lane evaluates table entries.  pose (6) and vals
// (13) are in LDS; pose must be written before the call by a lane of the same
// wave.
__device__ inline void make_pose_tables_wave(const double* pose, double* vals, PoseTables* T) {
  const int lane = threadIdx.x & (kWave - 1);
  wave_lds_fence();
  if (lane < 3) {
    double sn, cs;
    sincos(pose[3 + lane], &sn, &cs);
    vals[2 * lane] = sn;
    vals[2 * lane + 1] = cs;
  }
  wave_lds_fence();
  if (lane >= 6 && lane < 13) {
    double v = 1.0;
    if (lane < 12) {
      const int q = lane - 6;  // sin (even) or cos (odd) of angle q / 2
      v = fabs(pose[3 + q / 2]) < 1e-5 ? (double)(q & 1) : vals[q];
    }
    vals[lane] = v;
  }
  wave_lds_fence();
  double* out = reinterpret_cast<double*>(T);
  for (int e = lane; e < kPoseEntries; e += kWave) {
    const PoseEntry E = kPoseTable[e];
    const double p0 = vals[E.t[0].a] * vals[E.t[0].b] * vals[E.t[0].c];
    const double p1 = vals[E.t[1].a] * vals[E.t[1].b] * vals[E.t[1].c];
    out[3 + e] = (double)E.t[0].k * p0 + (double)E.t[1].k * p1;
  }
  if (lane < 3) out[lane] = pose[lane];
}

// All threads of the workgroup call this, after newton_step_reset and a
// barrier; waves 0 and 1 do the work.  H (36, lower triangle read) and g (6)
// are in LDS; on return W.ok says whether a step exists and d (LDS) holds it.
__device__ inline void newton_step_wg(const double* H, const double* g, double* d, SolveScratch& W) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  if (wave == 1) {
    // Lane 0 tests -H - 1e-10 I, lane 1 factors -H and solves, at once.
    double dc[6];
    const bool pd = lane < 2 && neg_cholesky6_solve(H, lane == 0 ? 1e-10 : 0.0, g, dc);
    if ((__ballot(pd) & 3ull) == 3ull && lane == 1) {
#pragma unroll
      for (int i = 0; i < 6; ++i) W.dc[i] = dc[i];
      __hip_atomic_store(&W.fast, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  } else if (wave == 0) {
    double mg[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) mg[k] = -g[k];
    Eigen6Row o;
    // Stops between QL blocks once wave 1 has the Cholesky step.
    eigen6_ql<DeviceEigenMath>(H, mg, lane, o, [&W] {
      return __hip_atomic_load(&W.fast, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
    });
    if (o.ok >= 0) {
      // The reference's clamp (newton.rs:119-144) and the SVD solve's
      // tolerance: d = sum_k Q_k (Q_k . -g) / lambda'_k over |lambda'_k| > 1e-10.
      bool neg_def = true;
#pragma unroll
      for (int k = 0; k < 6; ++k) neg_def &= o.lam[k] < 0.0;
      double max_sv = 0.0, dr = 0.0;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const double lk = (!neg_def && o.lam[k] > -1e-3) ? -1e-3 : o.lam[k];
        max_sv = fmax(max_sv, fabs(lk));
        dr += fabs(lk) > 1e-10 ? o.q[k] * (o.y[k] * DeviceEigenMath::rcp(lk)) : 0.0;
      }
      if (max_sv >= 1e-10) {
        if (lane < 6) d[lane] = dr;
        if (lane == 0) W.ok = 1;
      }
      if (lane == 0) ++W.solves;
    }
    if (lane == 0) W.sweeps += o.iters;
  }
  __syncthreads();
  if (W.fast) {
    if (threadIdx.x < 6) d[threadIdx.x] = W.dc[threadIdx.x];
    if (threadIdx.x == 0) W.ok = 1;
  }
  __syncthreads();
}

// Clears the per-call flags (one thread, before the barrier that precedes
// newton_step_wg).
__device__ inline void newton_step_reset(SolveScratch& W) {
  W.ok = 0;
  W.fast = 0;
}

}  // namespace ndt
