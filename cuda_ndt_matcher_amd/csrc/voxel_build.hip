// Voxel-grid build on gfx950: the GPU form of set_target's
// build_voxel_grid_cpu (src/ndt_cuda/src/voxel_grid/cpu.rs:51-75,
// voxel_grid/types.rs:54-212).
//
// Pipeline (all device work on the handle's stream):
//   1. key_bounds     : per-point voxel key floor(p/res) (f32 division, like
//                       VoxelCoord::from_point); min/max over the keys.
//   2. keygen_count   : compact Morton code of the key relative to the box
//                       (only the significant bits of each axis), written as
//                       u32 (<= 32 bits) or u64; first-pass digit counts per
//                       sort tile.
//   3. radix sort     : stable LSD, 8-bit digits; per pass sort_count (digit
//                       counts per 4096-key tile), sort_tile_scan (prefix
//                       over tiles), sort_scatter (wave64 ballot ranking,
//                       LDS staging, writes in runs).
//   4. segments       : head counts per tile, then segment starts (each tile
//                       sums the earlier tiles' counts itself).
//   5. voxel_stats    : one thread per segment: shifted f64 moments, mean,
//                       covariance, 3x3 Jacobi eigen-regularisation, inverse.
//   6. cell tables (occupancy word and mean/slot entry per cell of the box,
//      the entries written with the occupancy bytes),
//      or, when the box is too large for them, an open-addressing hash.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "eigen3.hpp"
#include "morton.hpp"
#include "ndt_device.hpp"
#include "ndt_internal.hpp"

namespace ndt {
namespace {

constexpr int kBuildBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kBuildBlock * kSortItems;  // 4096 keys per tile
constexpr int kRadix = 256;
static_assert(kRadix == kBuildBlock, "the sort passes give each thread one digit");
// Row length of the digit-major per-tile count table: the tile count rounded up
// to whole 8-entry (32-byte) groups, the unit sort_tile_scan loads.
constexpr uint32_t tile_stride(uint32_t tiles) { return (tiles + 7u) & ~7u; }

__global__ void fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// Per-workgroup min/max (and count) to part[8 * block .. 8 * block + 6]
// (min x, y, z, max x, y, z, count), no atomics: the host reads the parts back
// with the copy it waits for anyway and reduces them.
constexpr int kMaxBoundBlocks = 1024;
__device__ inline void block_minmax_store(int32_t mn[3], int32_t mx[3], uint32_t cnt, int32_t* part) {
  __shared__ int32_t s_mn[kBuildBlock / kWave][3], s_mx[kBuildBlock / kWave][3];
  __shared__ uint32_t s_c[kBuildBlock / kWave];
  for (int a = 0; a < 3; ++a)
    for (int off = 32; off >= 1; off >>= 1) {
      mn[a] = min(mn[a], __shfl_xor(mn[a], off, kWave));
      mx[a] = max(mx[a], __shfl_xor(mx[a], off, kWave));
    }
  for (int off = 32; off >= 1; off >>= 1) cnt += __shfl_xor(cnt, off, kWave);
  const int wave = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    for (int a = 0; a < 3; ++a) {
      s_mn[wave][a] = mn[a];
      s_mx[wave][a] = mx[a];
    }
    s_c[wave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBuildBlock / kWave; ++w) {
      for (int a = 0; a < 3; ++a) {
        mn[a] = min(mn[a], s_mn[w][a]);
        mx[a] = max(mx[a], s_mx[w][a]);
      }
      cnt += s_c[w];
    }
    int32_t* q = part + 8 * blockIdx.x;
    for (int a = 0; a < 3; ++a) {
      q[a] = mn[a];
      q[3 + a] = mx[a];
    }
    q[6] = (int32_t)cnt;
    q[7] = 0;
  }
}

// Per-block box of the voxel keys.  The key sat_i32(floor(p / res)) (or
// floor(p * inv)) never decreases as p grows (correctly rounded division or
// product, floor and saturation all keep order), so the smallest and largest
// keys are the keys of the smallest and largest coordinates; a NaN coordinate
// has key 0.  Threads keep the coordinate extremes and a NaN flag per axis and
// form six keys at the end.  With 16-byte aligned points a thread reads four
// points at a time as three 16-byte loads, four such groups per step.
__global__ __launch_bounds__(kBuildBlock) void key_bounds(const float* __restrict__ pts, size_t n, float res,
                                                          float inv, int mul, int aligned16, int32_t* part) {
  auto key = [&](float p) { return mul ? sat_i32(floorf(p * inv)) : voxel_key(p, res); };
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  bool nan[3] = {false, false, false};
  auto take = [&](int a, float p) {
    lo[a] = fminf(lo[a], p);  // fminf/fmaxf return the other operand for a NaN
    hi[a] = fmaxf(hi[a], p);
    nan[a] |= p != p;
  };
  const size_t S = (size_t)gridDim.x * blockDim.x;
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t tail = 0;
  if (aligned16) {
    const float4* q = reinterpret_cast<const float4*>(pts);
    const size_t quads = n / 4;
    size_t c = t;
    for (; c + 3 * S < quads; c += 4 * S) {
      float4 f[4][3];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k) f[u][k] = q[3 * (c + u * S) + k];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float* e = reinterpret_cast<const float*>(&f[u][0]);
#pragma unroll
        for (int k = 0; k < 12; ++k) take(k % 3, e[k]);
      }
    }
    for (; c < quads; c += S) {
      float4 f[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) f[k] = q[3 * c + k];
      const float* e = reinterpret_cast<const float*>(&f[0]);
#pragma unroll
      for (int k = 0; k < 12; ++k) take(k % 3, e[k]);
    }
    tail = 4 * quads;
  }
  for (size_t i = tail + t; i < n; i += S)
    for (int a = 0; a < 3; ++a) take(a, pts[3 * i + a]);
  // A thread without points keeps +inf / -inf, whose keys are the neutral
  // INT32_MAX / INT32_MIN.
  int32_t mn[3], mx[3];
  for (int a = 0; a < 3; ++a) {
    mn[a] = key(lo[a]);
    mx[a] = key(hi[a]);
    if (nan[a]) {
      mn[a] = min(mn[a], 0);
      mx[a] = max(mx[a], 0);
    }
  }
  block_minmax_store(mn, mx, 0u, part);
}

struct KeyParams {
  float res;
  float inv;  // 1 / res in f32 (mul keys)
  int mul;    // 0: floor(p / res) (map voxels); 1: floor(p * inv) (downsampling, and map
              // voxels at a power-of-two res, where the two agree)
  int32_t kmin[3];
  int bits[3];
  int passes;
  MortonLayout layout;  // from bits (morton_layout)
};

// voxel_downsample_cpu's cell (filtering/cpu.rs:79-81): floor of the f32
// product p * (1 / res), saturating cast.
__device__ inline int32_t mul_key(float p, float inv) { return sat_i32(floorf(p * inv)); }
__device__ inline int32_t key_of(float p, const KeyParams& kp) {
  return kp.mul ? mul_key(p, kp.inv) : voxel_key(p, kp.res);
}

// Keys in sort-tile order: item i of lane l in wave w of block t is point
// t*4096 + w*1024 + i*64 + l (the layout the sort passes load with), so the
// block also counts the first pass's digits of its tile:
// tile_hist[d * tstride + t].  Every load is issued before any key is formed;
// indices past n are clamped to n - 1 and their items not counted.
// Workgroups are handed to the eight XCDs in turn (workgroup b to XCD b % 8).
// xcd_tile gives the workgroups of one XCD consecutive tiles, so that the
// neighbouring entries of a tile-count row (a 64-byte line holds 16 tiles)
// are written through one L2 instead of eight.  A bijection on [0, tiles):
// XCD x holds tiles / 8 workgroups, plus one when x < tiles % 8, and takes
// the tiles from x * (tiles / 8) + min(x, tiles % 8) on.
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t tiles) {
  const uint32_t x = b & 7u, i = b >> 3, q = tiles >> 3, r = tiles & 7u;
  return x * q + (x < r ? x : r) + i;
}

template <class K>
__global__ __launch_bounds__(kBuildBlock) void keygen_count(const float* __restrict__ pts, uint32_t n, KeyParams kp,
                                                            K* __restrict__ keys, uint32_t* __restrict__ tile_hist,
                                                            uint32_t tstride) {
  __shared__ uint32_t s_hist[kRadix];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t first = tile * kSortTile + wave * (kSortItems * kWave) + lane;
  float v[kSortItems][3];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = first + i * kWave;
    const size_t c = 3 * (size_t)(idx < n ? idx : n - 1);
#pragma unroll
    for (int a = 0; a < 3; ++a) v[i][a] = pts[c + a];
  }
  s_hist[tid] = 0;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = first + i * kWave;
    const uint32_t rx = (uint32_t)(key_of(v[i][0], kp) - kp.kmin[0]);
    const uint32_t ry = (uint32_t)(key_of(v[i][1], kp) - kp.kmin[1]);
    const uint32_t rz = (uint32_t)(key_of(v[i][2], kp) - kp.kmin[2]);
    uint64_t m;
    if constexpr (sizeof(K) == 4)  // 32-bit keys: the layout has at most 32 bits
      m = spread_morton32(rx, ry, rz, kp.layout);
    else
      m = spread_morton(rx, ry, rz, kp.layout);
    if (idx < n) {
      keys[idx] = (K)m;
      atomicAdd(&s_hist[m & 0xFFu], 1u);
    }
  }
  __syncthreads();
  tile_hist[(size_t)tid * tstride + tile] = s_hist[tid];
}

// One stable LSD pass over the 8-bit digit at `shift` takes three kernels:
// sort_count (digit counts per tile of 4096 keys), sort_tile_scan (for each
// digit, the exclusive prefix over tiles plus the digit's global offset) and
// sort_scatter (ranks within the tile, scatter to the prefix).  No workgroup
// waits on another: tile ids are block ids and the only cross-tile data is
// the scanned table, written by the kernel before.
//
// Keys are loaded wave-striped: item i of lane l in wave w is element
// tile*4096 + w*1024 + i*64 + l, so (wave, item, lane) order is input order
// and the ranks are stable.  Within a wave the rank of an item is the wave's
// running count of its digit plus the number of lower lanes holding the same
// digit in this round (found with 8 ballots).  Every load is issued before the
// ranking starts (the ranking's LDS traffic would otherwise keep the compiler
// from hoisting them, one memory round trip per item); indices past n are
// clamped to n - 1 (n >= 1 whenever a tile exists) and the items marked
// invalid.

// Lanes of this wave whose item has the same digit, among the valid ones.
__device__ inline uint64_t digit_match(uint32_t d, bool valid) {
  uint64_t match = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t bal = __ballot((d >> b) & 1u);
    match &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return match;
}

// tile_hist[d * tstride + tile] = number of keys of the tile with digit d
// (digit-major, so that sort_tile_scan reads each digit's row contiguously).
template <class K>
__global__ __launch_bounds__(kBuildBlock) void sort_count(const K* __restrict__ keys_in, uint32_t n, int shift,
                                                          uint32_t* __restrict__ tile_hist, uint32_t tstride) {
  __shared__ uint32_t s_hist[kRadix];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t first = tile * kSortTile + wave * (kSortItems * kWave) + lane;
  K key[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = first + i * kWave;
    key[i] = keys_in[idx < n ? idx : n - 1];
  }
  s_hist[tid] = 0;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const bool valid = first + i * kWave < n;
    const uint32_t d = (uint32_t)(key[i] >> shift) & 0xFFu;
    if (valid) atomicAdd(&s_hist[d], 1u);
  }
  __syncthreads();
  tile_hist[(size_t)tid * tstride + tile] = s_hist[tid];
}

// In place on each digit's row: tile_hist[d * tstride + t] <- sum over tiles
// u < t of tile_hist[d * tstride + u]; digit_total[d] <- the sum over all
// tiles.  One workgroup per digit; each thread loads eight consecutive tiles'
// counts (two 16-byte loads; tstride is a multiple of 8 and the entries past
// `tiles` are ignored and overwritten).
__global__ __launch_bounds__(kBuildBlock) void sort_tile_scan(uint32_t* __restrict__ tile_hist, uint32_t tiles,
                                                              uint32_t tstride, uint32_t* __restrict__ digit_total) {
  constexpr int kWaves = kBuildBlock / kWave;
  constexpr int kPer = 8;
  __shared__ uint32_t s_wsum[kWaves];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t d = blockIdx.x;
  uint32_t* row = tile_hist + (size_t)d * tstride;
  uint32_t carry = 0;
  for (uint32_t t0 = 0; t0 < tstride; t0 += kPer * kBuildBlock) {
    const uint32_t tb = t0 + tid * kPer;
    const bool have = tb < tstride;
    uint32_t v[kPer];
    if (have) {
      const uint4 lo = *reinterpret_cast<const uint4*>(row + tb);
      const uint4 hi = *reinterpret_cast<const uint4*>(row + tb + 4);
      v[0] = lo.x, v[1] = lo.y, v[2] = lo.z, v[3] = lo.w;
      v[4] = hi.x, v[5] = hi.y, v[6] = hi.z, v[7] = hi.w;
    }
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      v[q] = have && tb + q < tiles ? v[q] : 0u;
      sum += v[q];
    }
    uint32_t x = sum;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (lane == kWave - 1) s_wsum[wave] = x;
    __syncthreads();
    uint32_t pre = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t s = s_wsum[w];
      pre += w < wave ? s : 0u;
      all += s;
    }
    if (have) {
      uint32_t e[kPer];
      uint32_t run = carry + pre + x - sum;
#pragma unroll
      for (int q = 0; q < kPer; ++q) {
        e[q] = run;
        run += v[q];
      }
      *reinterpret_cast<uint4*>(row + tb) = make_uint4(e[0], e[1], e[2], e[3]);
      *reinterpret_cast<uint4*>(row + tb + 4) = make_uint4(e[4], e[5], e[6], e[7]);
    }
    carry += all;
    __syncthreads();
  }
  if (tid == 0) digit_total[d] = carry;
}

// Ranks the tile's keys by digit, stages them in LDS in digit order and
// writes them out in runs starting at the digit's offset (exclusive scan of
// digit_total over digits) plus tile_prefix[d * tstride + tile].
template <class K>
__global__ __launch_bounds__(kBuildBlock) void sort_scatter(const K* __restrict__ keys_in,
                                                            const uint32_t* __restrict__ vals_in,
                                                            K* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                            uint32_t n, int shift,
                                                            const uint32_t* __restrict__ tile_prefix,
                                                            uint32_t tstride,
                                                            const uint32_t* __restrict__ digit_total) {
  constexpr int kWaves = kBuildBlock / kWave;
  __shared__ uint32_t s_wave_hist[kWaves][kRadix];
  __shared__ uint32_t s_local[kRadix];
  __shared__ uint32_t s_base[kRadix];
  __shared__ uint32_t s_wsum[2][kWaves];
  __shared__ K s_keys[kSortTile];
  __shared__ uint32_t s_vals[kSortTile];

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t base = tile * kSortTile;
  const uint32_t first = base + wave * (kSortItems * kWave) + lane;

  K key[kSortItems];
  uint32_t val[kSortItems];
  uint32_t rank[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = first + i * kWave;
    key[i] = keys_in[idx < n ? idx : n - 1];
  }
  if (vals_in) {
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
      const uint32_t idx = first + i * kWave;
      val[i] = vals_in[idx < n ? idx : n - 1];
    }
  } else {
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) val[i] = first + i * kWave;
  }
  const uint32_t my_prefix = tile_prefix[(size_t)tid * tstride + tile];
  const uint32_t my_total = digit_total[tid];
  for (int i = tid; i < kWaves * kRadix; i += kBuildBlock) (&s_wave_hist[0][0])[i] = 0;
  __syncthreads();

  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const bool valid = first + i * kWave < n;
    const uint32_t d = (uint32_t)(key[i] >> shift) & 0xFFu;
    const uint64_t match = digit_match(d, valid);
    const uint32_t lower = __popcll(match & lt_mask);
    uint32_t before = 0;
    if (valid) before = s_wave_hist[wave][d];
    __builtin_amdgcn_wave_barrier();
    if (valid && lower == 0) s_wave_hist[wave][d] = before + (uint32_t)__popcll(match);
    __builtin_amdgcn_wave_barrier();
    rank[i] = valid ? before + lower : 0xFFFFFFFFu;
  }
  __syncthreads();

  // Thread t owns digit t: per-wave exclusive offsets, the tile total, and
  // the exclusive scans over digits of the tile totals and of the global
  // totals (a shuffle scan within each wave of 64 digits, then the sums of the
  // lower waves).
  {
    const int d = tid;
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      const uint32_t c = s_wave_hist[w][d];
      s_wave_hist[w][d] = total;
      total += c;
    }
    uint32_t x = total, y = my_total;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t xs = __shfl_up(x, off);
      const uint32_t ys = __shfl_up(y, off);
      if (lane >= off) {
        x += xs;
        y += ys;
      }
    }
    if (lane == kWave - 1) {
      s_wsum[0][wave] = x;
      s_wsum[1][wave] = y;
    }
    __syncthreads();
    uint32_t xpre = 0, ypre = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      xpre += w < wave ? s_wsum[0][w] : 0u;
      ypre += w < wave ? s_wsum[1][w] : 0u;
    }
    s_local[d] = xpre + x - total;
    s_base[d] = ypre + y - my_total + my_prefix;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    if (rank[i] == 0xFFFFFFFFu) continue;
    const uint32_t d = (uint32_t)(key[i] >> shift) & 0xFFu;
    const uint32_t pos = s_local[d] + s_wave_hist[wave][d] + rank[i];
    s_keys[pos] = key[i];
    s_vals[pos] = val[i];
  }
  __syncthreads();
  const uint32_t count = n - base < (uint32_t)kSortTile ? n - base : (uint32_t)kSortTile;
  for (uint32_t j = tid; j < count; j += kBuildBlock) {
    const K k = s_keys[j];
    const uint32_t d = (uint32_t)(k >> shift) & 0xFFu;
    const uint32_t g = s_base[d] + (j - s_local[d]);
    keys_out[g] = k;
    vals_out[g] = s_vals[j];
  }
}

// Segment heads: flag[i] = (i == 0 || key[i] != key[i-1]).  Three phases:
// per-tile head counts, then the write of segment starts.
template <class K>
__global__ __launch_bounds__(kBuildBlock) void seg_count(const K* __restrict__ keys, uint32_t n,
                                                         uint32_t* __restrict__ tile_heads) {
  __shared__ uint32_t s[kBuildBlock / kWave];
  const uint32_t base = blockIdx.x * kSortTile;
  // Every load issued before any comparison (positions past n re-read the
  // last key and are not counted; n >= 1).
  K cur[kSortItems], prev[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = base + threadIdx.x + i * kBuildBlock;
    const uint32_t q = idx < n ? idx : n - 1;
    cur[i] = keys[q];
    prev[i] = keys[q > 0 ? q - 1 : 0];
  }
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = base + threadIdx.x + i * kBuildBlock;
    c += (idx < n && (idx == 0 || cur[i] != prev[i])) ? 1u : 0u;
  }
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, kWave);
  if ((threadIdx.x & (kWave - 1)) == 0) s[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kBuildBlock / kWave; ++w) t += s[w];
    tile_heads[blockIdx.x] = t;
  }
}

// The segment count goes to *total (device) and *host_total (pinned host memory).
__global__ __launch_bounds__(1024) void seg_scan_tiles(uint32_t* tile_heads, uint32_t tiles, uint32_t* total,
                                                       uint32_t* host_total) {
  __shared__ uint32_t s[1024];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < tiles; base += 1024) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < tiles ? tile_heads[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint32_t t = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < tiles) tile_heads[i] = carry + s[threadIdx.x] - v;
    const uint32_t block_total = s[1023];
    __syncthreads();
    carry += block_total;
  }
  if (threadIdx.x == 0) {
    *total = carry;
    *host_total = carry;
  }
}

// Segment starts of one tile, in position order.  Keys are loaded
// wave-striped (item i of lane l in wave w is position tile*4096 + w*1024 +
// i*64 + l, so (wave, item, lane) order is position order) with every load
// issued first; a head's segment id is the tile's offset plus the heads of
// the lower waves, the earlier items of its wave and its lower lanes.  The
// tile's offset is the sum of the earlier tiles' head counts (seg_count's
// table, a few KB that L2 serves to every workgroup), formed while the keys
// load; the last tile's workgroup writes the segment count to total and
// host_total (pinned, mapped host memory).
template <class K>
__global__ __launch_bounds__(kBuildBlock) void seg_write(const K* __restrict__ keys, uint32_t n,
                                                         const uint32_t* __restrict__ tile_count,
                                                         uint32_t* __restrict__ seg_start, uint32_t* total,
                                                         uint32_t* host_total) {
  constexpr int kWaves = kBuildBlock / kWave;
  __shared__ uint32_t s_wave[kWaves], s_pre[kWaves];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const uint32_t first = blockIdx.x * kSortTile + wave * (kSortItems * kWave) + lane;
  K cur[kSortItems], prev[kSortItems];
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = first + i * kWave;
    const uint32_t c = idx < n ? idx : n - 1;
    cur[i] = keys[c];
    prev[i] = keys[c > 0 ? c - 1 : 0];
  }
  uint32_t pre = 0;
  for (uint32_t t = tid; t < blockIdx.x; t += kBuildBlock) pre += tile_count[t];
#pragma unroll
  for (int off = kWave / 2; off >= 1; off >>= 1) pre += __shfl_xor(pre, off, kWave);
  uint64_t heads[kSortItems];
  uint32_t count = 0;
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    const uint32_t idx = first + i * kWave;
    heads[i] = __ballot(idx < n && (idx == 0 || cur[i] != prev[i]));
    count += (uint32_t)__popcll(heads[i]);
  }
  if (lane == 0) {
    s_wave[wave] = count;
    s_pre[wave] = pre;
  }
  __syncthreads();
  uint32_t seg = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) seg += s_pre[w];
  if (blockIdx.x + 1 == gridDim.x && tid == 0) {
    uint32_t all = seg;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) all += s_wave[w];
    *total = all;
    *host_total = all;
  }
#pragma unroll
  for (int w = 0; w < kWaves; ++w) seg += w < wave ? s_wave[w] : 0u;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int i = 0; i < kSortItems; ++i) {
    if ((heads[i] >> lane) & 1ull) seg_start[seg + (uint32_t)__popcll(heads[i] & lt_mask)] = first + i * kWave;
    seg += (uint32_t)__popcll(heads[i]);
  }
}

struct StatsParams {
  float res;
  uint32_t min_points;
  double eig_ratio;                // (double)0.01f
  const uint32_t* seg_count;       // device: number of segments
  uint32_t cap;                    // capacity of the per-voxel arrays (segments at or past it are skipped)
  uint32_t n;
};

// Voxel::from_statistics + regularize_covariance (voxel_grid/types.rs:54-190)
// for one segment.  Moments are accumulated in f64 about the segment's first
// point (shifted data), which is exact where the reference's uncentered sums
// lose digits; mean and covariance agree with the reference formula to
// rounding.
__global__ __launch_bounds__(kBuildBlock) void voxel_stats(const float* __restrict__ pts,
                                                           const uint32_t* __restrict__ order,
                                                           const uint32_t* __restrict__ seg_start, StatsParams sp,
                                                           float4* __restrict__ vox, float* __restrict__ cov_out,
                                                           float* __restrict__ axis_out, uint32_t* __restrict__ count_out,
                                                           int32_t* __restrict__ coord_out, uint32_t* __restrict__ valid) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t segments = *sp.seg_count;
  if (s >= segments || s >= sp.cap) return;
  const uint32_t b = seg_start[s];
  const uint32_t e = (s + 1 < segments) ? seg_start[s + 1] : sp.n;
  const uint32_t cnt = e - b;
  const uint32_t first = order[b];
  const float p0x = pts[3 * first], p0y = pts[3 * first + 1], p0z = pts[3 * first + 2];
  coord_out[3 * s + 0] = voxel_key(p0x, sp.res);
  coord_out[3 * s + 1] = voxel_key(p0y, sp.res);
  coord_out[3 * s + 2] = voxel_key(p0z, sp.res);
  count_out[s] = cnt;
  valid[s] = 0;
  if (cnt < sp.min_points) return;
  const double ox = p0x, oy = p0y, oz = p0z;
  double sx = 0, sy = 0, sz = 0, sxx = 0, sxy = 0, sxz = 0, syy = 0, syz = 0, szz = 0;
  auto add = [&](float px, float py, float pz) {
    const double x = (double)px - ox, y = (double)py - oy, z = (double)pz - oz;
    sx += x; sy += y; sz += z;
    sxx += x * x; sxy += x * y; sxz += x * z;
    syy += y * y; syz += y * z; szz += z * z;
  };
  uint32_t j = b;
  // Four points per step with every load issued first (the gathers dominate);
  // the sums still run in point order.
  for (; j + 4 <= e; j += 4) {
    uint32_t id[4];
    float v[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) id[u] = order[j + u];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int a = 0; a < 3; ++a) v[u][a] = pts[3 * id[u] + a];
#pragma unroll
    for (int u = 0; u < 4; ++u) add(v[u][0], v[u][1], v[u][2]);
  }
  for (; j < e; ++j) {
    const uint32_t i = order[j];
    add(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
  }
  const double n = (double)cnt;
  const double mx = sx / n, my = sy / n, mz = sz / n;
  const double inv_n1 = 1.0 / (n - 1.0);
  double c[9];
  c[0] = (sxx - sx * mx) * inv_n1;
  c[4] = (syy - sy * my) * inv_n1;
  c[8] = (szz - sz * mz) * inv_n1;
  c[1] = c[3] = (sxy - sx * my) * inv_n1;
  c[2] = c[6] = (sxz - sx * mz) * inv_n1;
  c[5] = c[7] = (syz - sy * mz) * inv_n1;
  double ev[3], V[9];
  sym_eigen3<DeviceEigenMath>(c, ev, V);
  const double max_ev = fmax(fmax(fmax(0.0, ev[0]), ev[1]), ev[2]);
  if (!(max_ev > 0.0)) return;  // types.rs:126-128 (NaN statistics are dropped too)
  int min_idx = 0;
  for (int i = 1; i < 3; ++i)
    if (ev[i] < ev[min_idx]) min_idx = i;
  const double thr = max_ev * sp.eig_ratio;
  bool needs = false;
  double cl[3];
  for (int i = 0; i < 3; ++i) {
    needs |= ev[i] < thr;
    cl[i] = ev[i] < thr ? thr : ev[i];
  }
  double reg[9], inv[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double r = 0.0, q = 0.0;
      for (int k = 0; k < 3; ++k) {
        r += V[i * 3 + k] * cl[k] * V[j * 3 + k];
        q += V[i * 3 + k] * (1.0 / cl[k]) * V[j * 3 + k];
      }
      reg[i * 3 + j] = needs ? r : c[i * 3 + j];
      inv[i * 3 + j] = q;
    }
  const float fmx = (float)(ox + mx), fmy = (float)(oy + my), fmz = (float)(oz + mz);
  vox[3 * s + 0] = make_float4(fmx, fmy, fmz, (float)inv[0]);
  vox[3 * s + 1] = make_float4((float)inv[1], (float)inv[2], (float)inv[3], (float)inv[4]);
  vox[3 * s + 2] = make_float4((float)inv[5], (float)inv[6], (float)inv[7], (float)inv[8]);
  for (int i = 0; i < 9; ++i) cov_out[9 * s + i] = (float)reg[i];
  for (int i = 0; i < 3; ++i) axis_out[3 * s + i] = (float)V[i * 3 + min_idx];
  valid[s] = 1;
}

__global__ __launch_bounds__(kBuildBlock) void hash_insert(const int32_t* __restrict__ coords,
                                                           const uint32_t* __restrict__ valid,
                                                           const float4* __restrict__ vox, uint32_t segments,
                                                           HashEntry* table, uint32_t mask) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= segments || !valid[s]) return;
  const uint64_t key = pack_coord(coords[3 * s], coords[3 * s + 1], coords[3 * s + 2]);
  uint32_t h = (uint32_t)mix64(key) & mask;
  for (;;) {
    unsigned long long* slot = &table[h].key;
    const unsigned long long prev = atomicCAS(slot, (unsigned long long)kEmptyKey, (unsigned long long)key);
    if (prev == kEmptyKey || prev == key) {
      const float4 m = vox[3 * s];
      table[h].slot = s;
      table[h].mx = m.x;
      table[h].my = m.y;
      table[h].mz = m.z;
      return;
    }
    h = (h + 1) & mask;
  }
}

// Occupancy bitmap over the padded box: bit (x - ox) + dx*((y - oy) + dy*(z - oz)).
// Occupied cells of the padded box as one byte per cell (occ zero-filled
// beforehand; every writer stores the same value, so no atomics are needed),
// then packed into the bitmap, 32 cells per word.
// With cell_ent set (the cell tables are built), also each valid voxel's
// entry {mean, slot} at the same cell index.  With part set, also the number
// of valid segments of every workgroup that holds segments, to
// part[workgroup] (pinned, mapped host memory; the host adds the first
// ceil(min(segments, cap) / 256) entries).
__global__ __launch_bounds__(kBuildBlock) void occ_set(const int32_t* __restrict__ coords,
                                                       const uint32_t* __restrict__ valid,
                                                       const uint32_t* __restrict__ seg_count, uint32_t cap,
                                                       int32_t ox, int32_t oy, int32_t oz, uint32_t dx, uint32_t dy,
                                                       uint8_t* __restrict__ occ, const float4* __restrict__ vox,
                                                       float4* __restrict__ cell_ent, uint32_t* part) {
  __shared__ uint32_t s_c[kBuildBlock / kWave];
  const uint32_t segments = *seg_count;
  const uint32_t limit = segments < cap ? segments : cap;
  const uint32_t s0 = blockIdx.x * kBuildBlock;
  if (s0 >= limit) return;  // the whole workgroup
  const uint32_t s = s0 + threadIdx.x;
  const bool ok = s < limit && valid[s] != 0u;
  if (ok) {
    const uint64_t lin = (uint64_t)(coords[3 * s] - ox) +
                         (uint64_t)dx * ((uint64_t)(coords[3 * s + 1] - oy) +
                                         (uint64_t)dy * (uint64_t)(coords[3 * s + 2] - oz));
    occ[lin] = 1;
    if (cell_ent) {
      const float4 m = vox[3 * s];
      cell_ent[lin] = make_float4(m.x, m.y, m.z, __uint_as_float(s));
    }
  }
  if (part) {
    const uint64_t bal = __ballot(ok);
    if ((threadIdx.x & (kWave - 1)) == 0) s_c[threadIdx.x / kWave] = (uint32_t)__popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t c = 0;
#pragma unroll
      for (int w = 0; w < kBuildBlock / kWave; ++w) c += s_c[w];
      part[blockIdx.x] = c;
    }
  }
}

// bitmap[w] bit b = occ[32 w + b] (bytes are 0 or 1).  occ holds 32 bytes per
// word, 16-byte aligned.
// The bytes are cleared once read, so occ is all zero again for the next
// build's occ_set.
__global__ __launch_bounds__(kBuildBlock) void bitmap_pack(uint8_t* __restrict__ occ, size_t words,
                                                           uint32_t* __restrict__ bitmap) {
  for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < words; w += (size_t)gridDim.x * blockDim.x) {
    uint4* q = reinterpret_cast<uint4*>(occ + 32 * w);
    const uint4 a = q[0], b = q[1];
    q[0] = make_uint4(0u, 0u, 0u, 0u);
    q[1] = make_uint4(0u, 0u, 0u, 0u);
    const uint32_t part[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t bits = 0;
    // Four 0/1 bytes to four bits: byte k (bit 8k) times 2^(21-7k) lands on bit
    // 21+k; the sixteen partial products fall on distinct bits, so none carries.
#pragma unroll
    for (int j = 0; j < 8; ++j)
      bits |= ((part[j] * 0x00204081u) >> 21 & 0xFu) << (4 * j);
    bitmap[w] = bits;
  }
}

// ... and the 27-bit block occupancy of every cell whose block lies inside
// the padded box (zero on the outer layer, where no query block is centred).
// floor(2^64 / d) + 1: for every 32-bit n, floor(n / d) is the high half of
// n * magic (the error term n (magic d - 2^64) / 2^64 stays below 1 / d).
// d >= 2.
uint64_t div_magic(uint32_t d) {
  uint64_t f = UINT64_MAX / d;
  if ((d & (d - 1)) == 0) ++f;  // d divides 2^64
  return f + 1;
}

// cell_nbr over the padded box.  Each thread takes four consecutive cells of
// a row (y, z), the quads of all rows numbered row by row: per neighbouring
// row it loads the six occupancy bits of x0-1 .. x0+4 once and forms the four
// cells' 3-bit windows from them.  Cells on the box's outer layer get 0
// (their block would leave the box).  The box has at most 2^26 cells, so the
// quad numbers fit 32 bits; the divisions by the quads per row and by the
// box's y extent are multiplications by div_magic's reciprocals.
__global__ __launch_bounds__(kBuildBlock) void cell_nbr_set(GridView v, uint32_t quads, uint64_t quads_magic,
                                                            uint64_t ydim_magic, uint32_t total,
                                                            uint32_t* cell_nbr) {
  for (uint32_t t = blockIdx.x * kBuildBlock + threadIdx.x; t < total; t += gridDim.x * kBuildBlock) {
    const uint32_t row = (uint32_t)__umul64hi((uint64_t)t, quads_magic);
    const uint32_t q = t - row * quads;
    const uint32_t z = (uint32_t)__umul64hi((uint64_t)row, ydim_magic);
    const uint32_t y = row - z * v.bdim[1];
    const uint32_t x0 = 4 * q;
    uint32_t m[4] = {0u, 0u, 0u, 0u};
    if (y >= 1 && y + 2 <= v.bdim[1] && z >= 1 && z + 2 <= v.bdim[2]) {
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const int dy = r / 3 - 1, dz = r % 3 - 1;
        const uint64_t lin = (uint64_t)x0 + (uint64_t)v.bdim[0] * ((uint64_t)((int32_t)y + dy) +
                                                                   (uint64_t)v.bdim[1] * (uint64_t)((int32_t)z + dz));
        // bit j of win: occupancy of cell x0 - 1 + j (the bit before lin = 0
        // does not exist and reads as 0; it only serves the outer-layer cell x0 = 0)
        const uint64_t s = lin > 0 ? lin - 1 : 0;
        const uint64_t w = s >> 5;
        const uint64_t two = (uint64_t)v.bitmap[w] | ((uint64_t)v.bitmap[w + 1] << 32);
        uint32_t win = (uint32_t)(two >> (s & 31));
        if (lin == 0) win <<= 1;
        // bit0 (x-1), bit1 (x), bit2 (x+1) of the window to bits r, 9 + r and
        // 18 + r: b3 * 0x10101 places copies of b3 at bits 0, 8 and 16, and
        // the mask keeps bit 0 of the first, bit 1 of the second and bit 2 of
        // the third (no copies overlap, so nothing carries).
#pragma unroll
        for (int c = 0; c < 4; ++c) m[c] |= ((((win >> c) & 7u) * 0x10101u) & 0x40201u) << r;
      }
    }
    uint32_t* out = cell_nbr + (size_t)row * v.bdim[0];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const uint32_t x = x0 + c;
      if (x < v.bdim[0]) out[x] = (x >= 1 && x + 2 <= v.bdim[0]) ? m[c] : 0u;
    }
  }
}

__global__ __launch_bounds__(kBuildBlock) void valid_bounds(const int32_t* __restrict__ coords,
                                                            const uint32_t* __restrict__ valid,
                                                            const uint32_t* __restrict__ seg_count, uint32_t cap,
                                                            int32_t* part) {
  int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  uint32_t c = 0;
  const uint32_t segments = min(*seg_count, cap);
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < segments; s += gridDim.x * blockDim.x) {
    const uint32_t v = valid[s];
    const int32_t k0 = coords[3 * s], k1 = coords[3 * s + 1], k2 = coords[3 * s + 2];  // loaded with valid
    if (!v) continue;
    ++c;
    mn[0] = min(mn[0], k0);
    mn[1] = min(mn[1], k1);
    mn[2] = min(mn[2], k2);
    mx[0] = max(mx[0], k0);
    mx[1] = max(mx[1], k1);
    mx[2] = max(mx[2], k2);
  }
  block_minmax_store(mn, mx, c, part);
}

int grid_for(size_t n, int block, int max_blocks) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (size_t)max_blocks) g = max_blocks;
  return (int)g;
}

template <class K>
hipError_t run_sort(DeviceGrid& G, const KeyParams& kp, size_t n, hipStream_t st) {
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  K* keys_a = reinterpret_cast<K*>(G.sort_keys_a.ptr);
  K* keys_b = reinterpret_cast<K*>(G.sort_keys_b.ptr);
  uint32_t* vals_a = reinterpret_cast<uint32_t*>(G.sort_vals_a.ptr);
  uint32_t* vals_b = reinterpret_cast<uint32_t*>(G.sort_vals_b.ptr);
  uint32_t* digit_total = reinterpret_cast<uint32_t*>(G.hist.ptr);    // 256, reused by every pass
  uint32_t* tile_hist = reinterpret_cast<uint32_t*>(G.tile_hist.ptr);  // 256 x tstride, reused by every pass
  const uint32_t tstride = tile_stride(tiles);
  hipLaunchKernelGGL(keygen_count<K>, dim3(tiles), dim3(kBuildBlock), 0, st, G.points, (uint32_t)n, kp, keys_a,
                     tile_hist, tstride);
  (void)hipEventRecord(G.ev_sort0, st);
  const K* kin = keys_a;
  const uint32_t* vin = nullptr;
  K* kout = keys_b;
  uint32_t* vout = vals_b;
  for (int p = 0; p < kp.passes; ++p) {
    if (p > 0)  // the first pass's counts come from keygen_count
      hipLaunchKernelGGL(sort_count<K>, dim3(tiles), dim3(kBuildBlock), 0, st, kin, (uint32_t)n, 8 * p, tile_hist,
                         tstride);
    hipLaunchKernelGGL(sort_tile_scan, dim3(kRadix), dim3(kBuildBlock), 0, st, tile_hist, tiles, tstride,
                       digit_total);
    hipLaunchKernelGGL(sort_scatter<K>, dim3(tiles), dim3(kBuildBlock), 0, st, kin, vin, kout, vout, (uint32_t)n,
                       8 * p, tile_hist, tstride, digit_total);
    // ping-pong
    const K* nk = kout;
    const uint32_t* nv = vout;
    kout = (kout == keys_b) ? keys_a : keys_b;
    vout = (vout == vals_b) ? vals_a : vals_b;
    kin = nk;
    vin = nv;
  }
  (void)hipEventRecord(G.ev_sort1, st);
  G.sorted_keys = const_cast<K*>(kin);
  G.sorted_vals = const_cast<uint32_t*>(vin);
  return hipGetLastError();
}

template <class K>
hipError_t run_segments(DeviceGrid& G, size_t n, hipStream_t st, uint32_t* segments_out) {
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  const K* keys = reinterpret_cast<const K*>(G.sorted_keys);
  uint32_t* tile_heads = reinterpret_cast<uint32_t*>(G.tile_heads.ptr);
  uint32_t* scalars = reinterpret_cast<uint32_t*>(G.scalars.ptr);
  hipLaunchKernelGGL(seg_count<K>, dim3(tiles), dim3(kBuildBlock), 0, st, keys, (uint32_t)n, tile_heads);
  hipLaunchKernelGGL(seg_write<K>, dim3(tiles), dim3(kBuildBlock), 0, st, keys, (uint32_t)n, tile_heads,
                     reinterpret_cast<uint32_t*>(G.seg_start.ptr), scalars, G.host_scalars);
  if (!segments_out) return hipGetLastError();  // read from host_scalars after a later synchronisation
  hipError_t e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  *segments_out = G.host_scalars[0];
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Sensor point filter (filtering/cpu.rs:7-58, voxel_downsample_cpu :67-100).

struct FilterDev {
  float min_dist_sq, max_dist_sq, min_z, max_z;  // squares formed in f32, as the reference does
};

// filter_points_cpu's tests in its order (cpu.rs:18-33): 0 kept, 1 removed by
// distance, 2 removed by z.  x*x + y*y + z*z rounded after every operation
// (no fused multiply-add), comparisons as written, so a NaN coordinate passes.
__device__ inline int filter_class(float x, float y, float z, const FilterDev& f) {
  float d = __fmul_rn(x, x);
  d = __fadd_rn(d, __fmul_rn(y, y));
  d = __fadd_rn(d, __fmul_rn(z, z));
  if (d < f.min_dist_sq || d > f.max_dist_sq) return 1;
  if (z < f.min_z || z > f.max_z) return 2;
  return 0;
}

// Per tile of kSortTile points: points kept (tile_kept[tile]); the removed
// counts are added to counts[1] (distance) and counts[2] (z).
__global__ __launch_bounds__(kBuildBlock) void filter_count(const float* __restrict__ pts, uint32_t n, FilterDev f,
                                                            uint32_t* __restrict__ tile_kept,
                                                            uint32_t* __restrict__ counts) {
  __shared__ uint32_t s[3][kBuildBlock / kWave];
  const uint32_t base = blockIdx.x * kSortTile;
  uint32_t k = 0, rd = 0, rz = 0;
  for (uint32_t j = threadIdx.x; j < kSortTile; j += kBuildBlock) {
    const uint32_t i = base + j;
    if (i < n) {
      const int c = filter_class(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], f);
      k += c == 0;
      rd += c == 1;
      rz += c == 2;
    }
  }
  for (int off = 32; off >= 1; off >>= 1) {
    k += __shfl_xor(k, off, kWave);
    rd += __shfl_xor(rd, off, kWave);
    rz += __shfl_xor(rz, off, kWave);
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    s[0][threadIdx.x / kWave] = k;
    s[1][threadIdx.x / kWave] = rd;
    s[2][threadIdx.x / kWave] = rz;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a)
      for (int w = 0; w < kBuildBlock / kWave; ++w) t[a] += s[a][w];
    tile_kept[blockIdx.x] = t[0];
    if (t[1]) atomicAdd(&counts[1], t[1]);
    if (t[2]) atomicAdd(&counts[2], t[2]);
  }
}

// Writes the kept points of each tile, in input order, from out + tile_off[tile].
// Each thread owns kSortItems consecutive points; a block scan of the per-thread
// kept counts places them.
__global__ __launch_bounds__(kBuildBlock) void filter_compact(const float* __restrict__ pts, uint32_t n, FilterDev f,
                                                              const uint32_t* __restrict__ tile_off,
                                                              float* __restrict__ out) {
  __shared__ uint32_t s[kBuildBlock];
  const uint32_t base = blockIdx.x * kSortTile + threadIdx.x * kSortItems;
  uint32_t keep = 0;  // bit k: point base + k is kept
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const uint32_t i = base + k;
    if (i < n && filter_class(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], f) == 0) keep |= 1u << k;
  }
  const uint32_t c = (uint32_t)__popc(keep);
  s[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < kBuildBlock; off <<= 1) {
    const uint32_t t = threadIdx.x >= (uint32_t)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t o = tile_off[blockIdx.x] + s[threadIdx.x] - c;
  for (int k = 0; k < kSortItems; ++k)
    if (keep & (1u << k)) {
      const uint32_t i = base + k;
      out[3 * o] = pts[3 * i];
      out[3 * o + 1] = pts[3 * i + 1];
      out[3 * o + 2] = pts[3 * i + 2];
      ++o;
    }
}

// One voxel per thread: f64 sums of its points in input order (the stable
// sort keeps it), centroid (sum / n) rounded to f32 (cpu.rs:82-98).
__global__ __launch_bounds__(kBuildBlock) void voxel_centroids(const float* __restrict__ pts,
                                                               const uint32_t* __restrict__ order,
                                                               const uint32_t* __restrict__ seg_start, uint32_t segs,
                                                               uint32_t n, float* __restrict__ out) {
  const uint32_t s = blockIdx.x * kBuildBlock + threadIdx.x;
  if (s >= segs) return;
  const uint32_t b = seg_start[s], e = s + 1 < segs ? seg_start[s + 1] : n;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  uint32_t i = b;
  // Sixteen points per step, all loads before the (in-order) additions: a
  // dense cell near the sensor holds hundreds of points, and one dependent
  // gather pair per point would serialise their latencies.
  constexpr int U = 16;
  for (; i + U <= e; i += U) {
    uint32_t j[U];
    float x[U], y[U], z[U];
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = order[i + u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      x[u] = pts[3 * j[u]];
      y[u] = pts[3 * j[u] + 1];
      z[u] = pts[3 * j[u] + 2];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      sx += (double)x[u];
      sy += (double)y[u];
      sz += (double)z[u];
    }
  }
  for (; i < e; ++i) {
    const uint32_t j = order[i];
    sx += (double)pts[3 * j];
    sy += (double)pts[3 * j + 1];
    sz += (double)pts[3 * j + 2];
  }
  const double cnt = (double)(e - b);
  out[3 * s] = (float)(sx / cnt);
  out[3 * s + 1] = (float)(sy / cnt);
  out[3 * s + 2] = (float)(sz / cnt);
}

uint32_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return (uint32_t)p;
}

}  // namespace

// Builds the grid from G.points (device, n points).  Returns a status code and
// fills G's voxel arrays, hash and views.
int build_grid(DeviceGrid& G, size_t n, float res, uint32_t min_points, bool cell_tables, hipStream_t st,
               std::string* err) {
  auto fail = [&](hipError_t e, const char* what) {
    *err = std::string(what) + ": " + hipGetErrorString(e);
    return -5;
  };
  if (n == 0) {
    *err = "Target point cloud is empty";
    return -2;
  }
  if (n >= (1ull << 30)) {
    *err = "target exceeds 2^30 points";
    return -7;
  }
  hipError_t e;
  (void)hipEventRecord(G.ev_build0, st);
  if ((e = G.scalars.reserve(64)) != hipSuccess) return fail(e, "alloc");  // [0]: segments (seg_write)
  if (!G.host_part &&
      (e = hipHostMalloc(reinterpret_cast<void**>(&G.host_part), (size_t)kMaxBoundBlocks * 8 * sizeof(int32_t),
                         hipHostMallocCoherent | hipHostMallocMapped)) != hipSuccess)
    return fail(e, "pinned alloc");
  // per-workgroup partials, written by the kernels into host_part and reduced on
  // the host (n > 0, so every block has an entry)
  auto reduce_parts = [&](int blocks, int32_t out[6], uint32_t* count) {
    for (int a = 0; a < 3; ++a) {
      out[a] = INT32_MAX;
      out[3 + a] = INT32_MIN;
    }
    uint64_t c = 0;
    for (int b = 0; b < blocks; ++b) {
      const int32_t* q = G.host_part + 8 * b;
      for (int a = 0; a < 3; ++a) {
        out[a] = std::min(out[a], q[a]);
        out[3 + a] = std::max(out[3 + a], q[3 + a]);
      }
      c += (uint32_t)q[6];
    }
    if (count) *count = (uint32_t)c;
  };
  const int kb_blocks = grid_for(n, kBuildBlock, kMaxBoundBlocks);
  // The per-workgroup partials go straight to pinned host memory (coherent and
  // mapped), so the stream carries no copy ahead of the synchronisation.
  // For a power-of-two resolution, p * (1 / res) is the same correctly rounded
  // value as p / res for every p (both round the one real value p * 2^-k), so
  // the keys take the multiplication instead of the division.
  int res_exp = 0;
  const float res_inv = 1.0f / res;
  const int key_mul = (std::frexp(res, &res_exp) == 0.5f && res_inv * res == 1.0f) ? 1 : 0;
  hipLaunchKernelGGL(key_bounds, dim3(kb_blocks), dim3(kBuildBlock), 0, st, G.points, n, res, res_inv, key_mul,
                     (reinterpret_cast<uintptr_t>(G.points) & 15u) == 0 ? 1 : 0, G.host_part);
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "bounds sync");
  int32_t kb[6];
  reduce_parts(kb_blocks, kb, nullptr);
  KeyParams kp;
  kp.res = res;
  kp.inv = res_inv;
  kp.mul = key_mul;
  int total_bits = 0;
  for (int a = 0; a < 3; ++a) {
    const int64_t span = (int64_t)kb[3 + a] - (int64_t)kb[a];
    if (kb[a] < -kCoordBias || kb[3 + a] >= kCoordBias || span >= (1ll << 21)) {
      *err = "map extent exceeds 2^20 voxels from the origin on some axis";
      return -7;
    }
    int bits = 0;
    while (((int64_t)1 << bits) <= span) ++bits;
    kp.kmin[a] = kb[a];
    kp.bits[a] = bits;
    total_bits += bits;
  }
  kp.passes = std::max(1, (total_bits + 7) / 8);
  kp.layout = morton_layout(kp.bits[0], kp.bits[1], kp.bits[2]);
  G.key_bits = total_bits;
  G.key64 = total_bits > 32;
  std::memcpy(G.key_min, kp.kmin, sizeof(kp.kmin));
  std::memcpy(G.key_bits_axis, kp.bits, sizeof(kp.bits));
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  const size_t ksz = G.key64 ? 8 : 4;
  if ((e = G.sort_keys_a.reserve(n * ksz)) != hipSuccess) return fail(e, "alloc keys");
  if ((e = G.sort_keys_b.reserve(n * ksz)) != hipSuccess) return fail(e, "alloc keys");
  if ((e = G.sort_vals_a.reserve(n * 4)) != hipSuccess) return fail(e, "alloc vals");
  if ((e = G.sort_vals_b.reserve(n * 4)) != hipSuccess) return fail(e, "alloc vals");
  if ((e = G.hist.reserve(kRadix * 4)) != hipSuccess) return fail(e, "alloc hist");
  if ((e = G.tile_hist.reserve((size_t)kRadix * tile_stride(tiles) * 4)) != hipSuccess) return fail(e, "alloc tile table");
  if ((e = G.tile_heads.reserve((size_t)tiles * 4)) != hipSuccess) return fail(e, "alloc tiles");
  if ((e = G.seg_start.reserve(n * 4)) != hipSuccess) return fail(e, "alloc segs");
  e = G.key64 ? run_sort<uint64_t>(G, kp, n, st) : run_sort<uint32_t>(G, kp, n, st);
  if (e != hipSuccess) return fail(e, "radix sort");
  // Lookup box and the per-voxel arrays' capacity.  When the padded
  // voxel-coordinate box of all points (kb, which holds every valid voxel) is
  // small enough for cell tables, it is used as is and nothing waits for the
  // host: the voxel kernels read the segment count from the device and the
  // arrays keep their capacity (at least n/4 + 1024 voxels on a first build,
  // at most n); the count and the valid count come back with the final
  // synchronisation, and a count above the capacity grows the arrays and
  // queues the per-voxel stage again.  Otherwise (a large map, far outlier
  // points whose voxels are too sparse to be valid, or no cell tables) the
  // build waits for the segment count and then for the valid voxels' box,
  // which may still fit the cell tables, as the hash capacity and the table
  // dimensions need them.
  auto padded_cells = [](const int32_t* box) {
    uint64_t c = 1;
    for (int a = 0; a < 3; ++a) c *= (uint64_t)((int64_t)box[3 + a] - box[a] + 5);
    return c;
  };
  const bool box_from_keys = cell_tables && padded_cells(kb) <= kCellTablesMaxCells;
  uint32_t segs = 0;
  e = G.key64 ? run_segments<uint64_t>(G, n, st, box_from_keys ? nullptr : &segs)
              : run_segments<uint32_t>(G, n, st, box_from_keys ? nullptr : &segs);
  if (e != hipSuccess) return fail(e, "segments");
  const uint32_t* seg_count = reinterpret_cast<const uint32_t*>(G.scalars.ptr);  // written by seg_write
  int32_t vb[6];
  int vb_blocks = 0;
  // Queues voxel_stats, the valid count and the lookup structures for up to
  // `cap` segments; 0 or a build error code.
  auto voxel_stage = [&](uint32_t cap) -> int {
    if ((e = G.vox.reserve((size_t)cap * 48)) != hipSuccess) return fail(e, "alloc voxels");
    if ((e = G.cov.reserve((size_t)cap * 36)) != hipSuccess) return fail(e, "alloc voxels");
    if ((e = G.axis.reserve((size_t)cap * 12)) != hipSuccess) return fail(e, "alloc voxels");
    if ((e = G.count.reserve((size_t)cap * 4)) != hipSuccess) return fail(e, "alloc voxels");
    if ((e = G.coord.reserve((size_t)cap * 12)) != hipSuccess) return fail(e, "alloc voxels");
    if ((e = G.valid.reserve((size_t)cap * 4)) != hipSuccess) return fail(e, "alloc voxels");
    G.seg_cap = std::max(G.seg_cap, cap);
    const uint32_t cap_blocks = (cap + kBuildBlock - 1) / kBuildBlock;
    StatsParams sp;
    sp.res = res;
    sp.min_points = min_points;
    sp.eig_ratio = (double)0.01f;  // ndt.rs:369 passes 0.01 as f32 (voxel_grid/types.rs:138)
    sp.seg_count = seg_count;
    sp.cap = cap;
    sp.n = (uint32_t)n;
    hipLaunchKernelGGL(voxel_stats, dim3(cap_blocks), dim3(kBuildBlock), 0, st, G.points, G.sorted_vals,
                       reinterpret_cast<const uint32_t*>(G.seg_start.ptr), sp, reinterpret_cast<float4*>(G.vox.ptr),
                       reinterpret_cast<float*>(G.cov.ptr), reinterpret_cast<float*>(G.axis.ptr),
                       reinterpret_cast<uint32_t*>(G.count.ptr), reinterpret_cast<int32_t*>(G.coord.ptr),
                       reinterpret_cast<uint32_t*>(G.valid.ptr));
    if (box_from_keys) {
      // The valid count comes from occ_set's per-workgroup counts.
      if (cap_blocks > G.host_valid_cap) {
        if (G.host_valid) (void)hipHostFree(G.host_valid);
        G.host_valid = nullptr;
        G.host_valid_cap = 0;
        if ((e = hipHostMalloc(reinterpret_cast<void**>(&G.host_valid), (size_t)cap_blocks * sizeof(uint32_t),
                               hipHostMallocCoherent | hipHostMallocMapped)) != hipSuccess)
          return fail(e, "pinned alloc");
        G.host_valid_cap = cap_blocks;
      }
    } else {
      // The box of the valid voxels, and their count, before the lookup structures.
      vb_blocks = grid_for(cap, kBuildBlock, kMaxBoundBlocks);
      hipLaunchKernelGGL(valid_bounds, dim3(vb_blocks), dim3(kBuildBlock), 0, st,
                         reinterpret_cast<const int32_t*>(G.coord.ptr),
                         reinterpret_cast<const uint32_t*>(G.valid.ptr), seg_count, cap, G.host_part);
      if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "build sync");
      reduce_parts(vb_blocks, vb, &G.num_valid);
    }
    const int32_t* box = box_from_keys ? kb : vb;
    GridView& v = G.view;
    v.vox = reinterpret_cast<const float4*>(G.vox.ptr);
    v.axis = reinterpret_cast<const float*>(G.axis.ptr);
    v.table = nullptr;
    v.table_mask = 0;
    for (int a = 0; a < 3; ++a) {
      v.cmin[a] = box[a];
      v.cmax[a] = box[3 + a];
    }
    v.res = res;
    v.bitmap = nullptr;
    v.cell_nbr = nullptr;
    v.cell_ent = nullptr;
    if (box_from_keys || G.num_valid > 0) {
      // Occupancy bitmap over the box padded by two cells, when it has < 2^32 cells.
      const uint64_t cells = padded_cells(box);
      uint32_t dims[3];
      for (int a = 0; a < 3; ++a) {
        dims[a] = (uint32_t)((int64_t)v.cmax[a] - v.cmin[a] + 5);
        v.borg[a] = v.cmin[a] - 2;
        v.bdim[a] = dims[a];
      }
      if (cells < (1ull << 32)) {
        const size_t words = (size_t)(cells / 32) + 2;
        if ((e = G.bitmap.reserve(words * 4)) != hipSuccess) return fail(e, "alloc bitmap");
        // The byte map stays all zero between builds (bitmap_pack clears it);
        // a new allocation, or a build that stopped between occ_set and
        // bitmap_pack, is filled here.
        const size_t had = G.occ.bytes;
        if ((e = G.occ.reserve(words * 32)) != hipSuccess) return fail(e, "alloc bitmap");
        if (G.occ.bytes != had) G.occ_zero = false;  // a new allocation
        const bool tables = cell_tables && cells <= kCellTablesMaxCells;
        if (tables) {
          const uint32_t nc = (uint32_t)cells;
          if ((e = G.cell_nbr.reserve((size_t)nc * 4)) != hipSuccess) return fail(e, "alloc cell tables");
          if ((e = G.cell_ent.reserve((size_t)nc * 16)) != hipSuccess) return fail(e, "alloc cell tables");
        }
        if (!G.occ_zero)
          hipLaunchKernelGGL(fill_u32, dim3(grid_for(G.occ.bytes / 4, 256, 4096)), dim3(256), 0, st,
                             reinterpret_cast<uint32_t*>(G.occ.ptr), G.occ.bytes / 4, 0u);
        G.occ_zero = false;
        hipLaunchKernelGGL(occ_set, dim3(cap_blocks), dim3(kBuildBlock), 0, st,
                           reinterpret_cast<const int32_t*>(G.coord.ptr),
                           reinterpret_cast<const uint32_t*>(G.valid.ptr), seg_count, cap, v.borg[0], v.borg[1],
                           v.borg[2], dims[0], dims[1], reinterpret_cast<uint8_t*>(G.occ.ptr),
                           reinterpret_cast<const float4*>(G.vox.ptr),
                           tables ? reinterpret_cast<float4*>(G.cell_ent.ptr) : nullptr,
                           box_from_keys ? G.host_valid : nullptr);
        hipLaunchKernelGGL(bitmap_pack, dim3(grid_for(words, kBuildBlock, 4096)), dim3(kBuildBlock), 0, st,
                           reinterpret_cast<uint8_t*>(G.occ.ptr), words, reinterpret_cast<uint32_t*>(G.bitmap.ptr));
        G.occ_zero = true;
        v.bitmap = reinterpret_cast<const uint32_t*>(G.bitmap.ptr);
        if (tables) {
          const uint32_t quads = (dims[0] + 3) / 4;
          const uint32_t total = quads * dims[1] * dims[2];  // < 2^32: at most 2^26 cells, rows of >= 5
          hipLaunchKernelGGL(cell_nbr_set, dim3(grid_for(total, kBuildBlock, 8192)), dim3(kBuildBlock), 0, st, v,
                             quads, div_magic(quads), div_magic(dims[1]), total,
                             reinterpret_cast<uint32_t*>(G.cell_nbr.ptr));
          v.cell_nbr = reinterpret_cast<const uint32_t*>(G.cell_nbr.ptr);
          v.cell_ent = reinterpret_cast<const float4*>(G.cell_ent.ptr);
        }
      }
    }
    // The open-addressing hash serves the lookups only when there are no cell
    // tables (box too large, or cell tables switched off); the segment count
    // is known on the host then.
    if (!v.cell_ent) {
      const uint32_t hcap = std::max<uint32_t>(1024, next_pow2(2ull * segs));
      if ((e = G.table.reserve((size_t)hcap * sizeof(HashEntry))) != hipSuccess) return fail(e, "alloc hash");
      hipLaunchKernelGGL(fill_u32, dim3(grid_for((size_t)hcap * 8, 256, 4096)), dim3(256), 0, st,
                         reinterpret_cast<uint32_t*>(G.table.ptr), (size_t)hcap * 8, 0xFFFFFFFFu);
      hipLaunchKernelGGL(hash_insert, dim3((segs + kBuildBlock - 1) / kBuildBlock), dim3(kBuildBlock), 0, st,
                         reinterpret_cast<const int32_t*>(G.coord.ptr),
                         reinterpret_cast<const uint32_t*>(G.valid.ptr), reinterpret_cast<const float4*>(G.vox.ptr),
                         segs, reinterpret_cast<HashEntry*>(G.table.ptr), hcap - 1);
      v.table = reinterpret_cast<const HashEntry*>(G.table.ptr);
      v.table_mask = hcap - 1;
    }
    (void)hipEventRecord(G.ev_build1, st);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "build sync");
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "build kernels");
    return 0;
  };
  const uint32_t cap0 =
      box_from_keys ? (uint32_t)std::min<size_t>(n, std::max<size_t>(G.seg_cap, n / 4 + 1024)) : segs;
  int rc = voxel_stage(cap0);
  if (rc) return rc;
  if (box_from_keys) {
    segs = G.host_scalars[0];
    if (segs > cap0 && (rc = voxel_stage(segs)) != 0) return rc;
    uint64_t c = 0;  // segs <= the last stage's capacity
    for (uint32_t b = 0; b < (segs + kBuildBlock - 1) / kBuildBlock; ++b) c += G.host_valid[b];
    G.num_valid = (uint32_t)c;
  }
  G.segments = segs;
  G.n_points = n;
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, G.ev_build0, G.ev_build1);
  G.build_ms = ms;
  (void)hipEventElapsedTime(&ms, G.ev_sort0, G.ev_sort1);
  G.sort_ms = ms;
  if (G.num_valid == 0) {
    *err = "No voxels created from target points (too sparse?)";
    return -4;
  }
  return 0;
}

// Sensor point filter on the device: G.points (n points) -> d_out.  The kept
// points (input order) or, with downsampling, one centroid per occupied cell
// (in the cells' key order).  Counts: [0] kept before downsampling, [1]
// removed by distance, [2] removed by z, [3] removed by downsampling.
// G is a grid used only for its scratch buffers (never a map).
int filter_points(DeviceGrid& G, size_t n, float min_distance, float max_distance, float min_z, float max_z,
                  int downsample, float resolution, float* d_out, size_t out_cap, size_t* out_n, uint64_t counts[4],
                  hipStream_t st, std::string* err) {
  auto fail = [&](hipError_t e, const char* what) {
    *err = std::string(what) + ": " + hipGetErrorString(e);
    return -5;
  };
  for (int i = 0; i < 4; ++i) counts[i] = 0;
  *out_n = 0;
  if (n == 0) return 0;
  if (n >= (1ull << 30)) {
    *err = "point cloud exceeds 2^30 points";
    return -7;
  }
  hipError_t e;
  if (!G.host_scalars &&
      (e = hipHostMalloc(reinterpret_cast<void**>(&G.host_scalars), 64 * sizeof(uint32_t),
                         hipHostMallocCoherent | hipHostMallocMapped)) !=
          hipSuccess)
    return fail(e, "pinned alloc");
  if (!G.host_part &&
      (e = hipHostMalloc(reinterpret_cast<void**>(&G.host_part), (size_t)kMaxBoundBlocks * 8 * sizeof(int32_t),
                         hipHostMallocCoherent | hipHostMallocMapped)) != hipSuccess)
    return fail(e, "pinned alloc");
  const uint32_t tiles = (uint32_t)((n + kSortTile - 1) / kSortTile);
  if ((e = G.scalars.reserve(64 * sizeof(uint32_t))) != hipSuccess) return fail(e, "alloc");
  if ((e = G.tile_heads.reserve((size_t)tiles * 4)) != hipSuccess) return fail(e, "alloc");
  uint32_t* scalars = reinterpret_cast<uint32_t*>(G.scalars.ptr);
  uint32_t* tile_kept = reinterpret_cast<uint32_t*>(G.tile_heads.ptr);
  FilterDev f;
  f.min_dist_sq = min_distance * min_distance;
  f.max_dist_sq = max_distance * max_distance;
  f.min_z = min_z;
  f.max_z = max_z;
  const bool ds = downsample && !(resolution <= 0.0f);  // cpu.rs:68: resolution <= 0 keeps the points
  hipLaunchKernelGGL(fill_u32, dim3(1), dim3(64), 0, st, scalars, (size_t)64, 0u);
  hipLaunchKernelGGL(filter_count, dim3(tiles), dim3(kBuildBlock), 0, st, G.points, (uint32_t)n, f, tile_kept,
                     scalars);
  hipLaunchKernelGGL(seg_scan_tiles, dim3(1), dim3(1024), 0, st, tile_kept, tiles, scalars, G.host_scalars);
  if ((e = hipMemcpyAsync(G.host_scalars, scalars, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return fail(e, "count copy");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "filter sync");
  const uint32_t kept = G.host_scalars[0];
  counts[0] = kept;
  counts[1] = G.host_scalars[1];
  counts[2] = G.host_scalars[2];
  if (!ds) {
    if (out_cap < kept) {
      *out_n = kept;
      *err = "output capacity below the kept point count";
      return -1;
    }
    if (kept) hipLaunchKernelGGL(filter_compact, dim3(tiles), dim3(kBuildBlock), 0, st, G.points, (uint32_t)n, f,
                                 tile_kept, d_out);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "filter sync");
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "filter kernels");
    *out_n = kept;
    return 0;
  }
  if (kept == 0) return 0;
  if ((e = G.owned_points.reserve((size_t)kept * 12)) != hipSuccess) return fail(e, "alloc");
  float* compact = reinterpret_cast<float*>(G.owned_points.ptr);
  hipLaunchKernelGGL(filter_compact, dim3(tiles), dim3(kBuildBlock), 0, st, G.points, (uint32_t)n, f, tile_kept,
                     compact);
  // cells of the kept points, as voxel_downsample_cpu forms them
  const float inv = 1.0f / resolution;
  const int kb_blocks = grid_for(kept, kBuildBlock, kMaxBoundBlocks);
  hipLaunchKernelGGL(key_bounds, dim3(kb_blocks), dim3(kBuildBlock), 0, st, compact, (size_t)kept, resolution, inv,
                     1, 1, G.host_part);  // compact is a device allocation (256-byte aligned)
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "bounds sync");
  int32_t kb[6];
  for (int a = 0; a < 3; ++a) {
    kb[a] = INT32_MAX;
    kb[3 + a] = INT32_MIN;
  }
  for (int b = 0; b < kb_blocks; ++b)
    for (int a = 0; a < 3; ++a) {
      kb[a] = std::min(kb[a], G.host_part[8 * b + a]);
      kb[3 + a] = std::max(kb[3 + a], G.host_part[8 * b + 3 + a]);
    }
  KeyParams kp;
  kp.res = resolution;
  kp.inv = inv;
  kp.mul = 1;
  int total_bits = 0;
  for (int a = 0; a < 3; ++a) {
    const int64_t span = (int64_t)kb[3 + a] - (int64_t)kb[a];
    if (kb[a] < -kCoordBias || kb[3 + a] >= kCoordBias || span >= (1ll << 21)) {
      *err = "point extent exceeds 2^20 cells from the origin on some axis at this resolution";
      return -7;
    }
    int bits = 0;
    while (((int64_t)1 << bits) <= span) ++bits;
    kp.kmin[a] = kb[a];
    kp.bits[a] = bits;
    total_bits += bits;
  }
  kp.passes = std::max(1, (total_bits + 7) / 8);
  kp.layout = morton_layout(kp.bits[0], kp.bits[1], kp.bits[2]);
  const bool key64 = total_bits > 32;
  const size_t ksz = key64 ? 8 : 4;
  if ((e = G.sort_keys_a.reserve((size_t)kept * ksz)) != hipSuccess) return fail(e, "alloc keys");
  if ((e = G.sort_keys_b.reserve((size_t)kept * ksz)) != hipSuccess) return fail(e, "alloc keys");
  if ((e = G.sort_vals_a.reserve((size_t)kept * 4)) != hipSuccess) return fail(e, "alloc vals");
  if ((e = G.sort_vals_b.reserve((size_t)kept * 4)) != hipSuccess) return fail(e, "alloc vals");
  if ((e = G.hist.reserve(kRadix * 4)) != hipSuccess) return fail(e, "alloc hist");
  const uint32_t ktiles = (uint32_t)((kept + kSortTile - 1) / kSortTile);
  if ((e = G.tile_hist.reserve((size_t)kRadix * tile_stride(ktiles) * 4)) != hipSuccess) return fail(e, "alloc tile table");
  if ((e = G.tile_heads.reserve((size_t)std::max(tiles, ktiles) * 4)) != hipSuccess) return fail(e, "alloc tiles");
  if ((e = G.seg_start.reserve((size_t)kept * 4)) != hipSuccess) return fail(e, "alloc segs");
  G.points = compact;
  e = key64 ? run_sort<uint64_t>(G, kp, kept, st) : run_sort<uint32_t>(G, kp, kept, st);
  if (e != hipSuccess) return fail(e, "radix sort");
  uint32_t segs = 0;
  e = key64 ? run_segments<uint64_t>(G, kept, st, &segs) : run_segments<uint32_t>(G, kept, st, &segs);
  if (e != hipSuccess) return fail(e, "segments");
  if (out_cap < segs) {
    *out_n = segs;
    *err = "output capacity below the downsampled point count";
    return -1;
  }
  hipLaunchKernelGGL(voxel_centroids, dim3((segs + kBuildBlock - 1) / kBuildBlock), dim3(kBuildBlock), 0, st, compact,
                     G.sorted_vals, reinterpret_cast<const uint32_t*>(G.seg_start.ptr), segs, kept, d_out);
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e, "filter sync");
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "filter kernels");
  counts[3] = kept - segs;
  *out_n = segs;
  return 0;
}

}  // namespace ndt
