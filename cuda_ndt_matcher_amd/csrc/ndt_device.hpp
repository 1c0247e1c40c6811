// Device-side building blocks shared by the voxel build, the Newton kernel and
// the scoring kernels.  Written for gfx950 (CDNA4): 64-lane wavefronts, FP64
// VALU at full rate, 160 KiB LDS per CU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ndt {

constexpr int kWave = 64;
constexpr uint64_t kEmptyKey = ~0ull;
constexpr int32_t kCoordBias = 1 << 20;  // voxel coordinates must lie in [-2^20, 2^20)

// ---------------------------------------------------------------------------
// Voxel grid as the kernels see it.  Voxel slots are in Morton-key order; a
// slot whose voxel was dropped (too few points, degenerate covariance) is
// never inserted in the hash and so never visited.
struct GridView {
  const float4* vox;     // 3 x float4 per slot: {mx,my,mz,C00},{C01,C02,C10,C11},{C12,C20,C21,C22}
  const float* axis;     // 3 floats per slot: principal axis (point-to-plane metric)
  const uint4* table;    // open-addressing hash: {key lo, key hi, slot, 0}
  uint32_t table_mask;   // capacity - 1 (capacity is a power of two, load <= 0.5)
  int32_t cmin[3];       // bounding box of the occupied voxel coordinates
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

__device__ inline int32_t hash_find(const GridView& g, int32_t x, int32_t y, int32_t z) {
  if (x < g.cmin[0] || x > g.cmax[0] || y < g.cmin[1] || y > g.cmax[1] || z < g.cmin[2] || z > g.cmax[2])
    return -1;
  const uint64_t key = pack_coord(x, y, z);
  uint32_t h = (uint32_t)mix64(key) & g.table_mask;
#pragma unroll 1
  for (;;) {
    const uint4 e = g.table[h];
    const uint64_t k = (uint64_t)e.x | ((uint64_t)e.y << 32);
    if (k == key) return (int32_t)e.z;
    if (k == kEmptyKey) return -1;
    h = (h + 1) & g.table_mask;
  }
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
// Neighbour enumeration.  `visit(slot)` is called for each neighbouring voxel.
//
// Radius mode: every voxel whose f32 mean is within `res` of q (strict <).
// A voxel's mean lies in the closed box of its own cell, so a mean within
// res of q belongs to a cell whose index differs from floor(q/res) by at most
// one on each axis; floor is taken of the f64 quotient, which is exact for
// f32 inputs.  Cells whose box is farther than res from q are skipped.
template <class F>
__device__ inline void for_each_radius(const GridView& g, float qx, float qy, float qz, F&& visit) {
  const double res = (double)g.res;
  const int32_t kx = sat_i32d(floor((double)qx / res));
  const int32_t ky = sat_i32d(floor((double)qy / res));
  const int32_t kz = sat_i32d(floor((double)qz / res));
  const float r2 = __fmul_rn(g.res, g.res);
  // Distance from q to the low/high faces of its own cell, per axis.
  const double lox = (double)qx - kx * res, hix = (kx + 1) * res - (double)qx;
  const double loy = (double)qy - ky * res, hiy = (ky + 1) * res - (double)qy;
  const double loz = (double)qz - kz * res, hiz = (kz + 1) * res - (double)qz;
  const double lim = res * res * (1.0 + 1e-6);
#pragma unroll 1
  for (int c = 0; c < 27; ++c) {
    const int dx = c / 9 - 1, dy = (c / 3) % 3 - 1, dz = c % 3 - 1;
    const double ex = dx < 0 ? lox : (dx > 0 ? hix : 0.0);
    const double ey = dy < 0 ? loy : (dy > 0 ? hiy : 0.0);
    const double ez = dz < 0 ? loz : (dz > 0 ? hiz : 0.0);
    if (ex * ex + ey * ey + ez * ez > lim) continue;
    const int32_t s = hash_find(g, kx + dx, ky + dy, kz + dz);
    if (s < 0) continue;
    const float4 m = g.vox[3 * s];
    if (dist2_f32(qx, qy, qz, m.x, m.y, m.z) < r2) visit(s, m);
  }
}

// Autoware pclomp direct searches: cell = floor(q/res) in f32; DIRECT1 the
// cell, DIRECT7 the cell and its 6 faces, DIRECT26 the 26 cells around it
// (PCL getAllNeighborCellIndices, which leaves the centre out).
template <class F>
__device__ inline void for_each_direct(const GridView& g, int mode, float qx, float qy, float qz, F&& visit) {
  const int32_t kx = voxel_key(qx, g.res), ky = voxel_key(qy, g.res), kz = voxel_key(qz, g.res);
  // DIRECT7 offsets in pclomp's order: centre, +x, -x, +y, -y, +z, -z.
  const int count = mode == 3 ? 1 : (mode == 2 ? 7 : 26);
#pragma unroll 1
  for (int c = 0; c < count; ++c) {
    int dx = 0, dy = 0, dz = 0;
    if (mode == 2) {
      const int axis = (c + 1) / 2, sign = (c & 1) ? 1 : -1;
      if (c > 0) {
        if (axis == 1) dx = sign;
        else if (axis == 2) dy = sign;
        else dz = sign;
      }
    } else if (mode == 1) {
      const int k = c < 13 ? c : c + 1;  // skip the centre (index 13 of 27)
      dx = k / 9 - 1;
      dy = (k / 3) % 3 - 1;
      dz = k % 3 - 1;
    }
    const int32_t s = hash_find(g, kx + dx, ky + dy, kz + dz);
    if (s >= 0) visit(s, g.vox[3 * s]);
  }
}

template <class F>
__device__ inline void for_each_neighbor(const GridView& g, int mode, float qx, float qy, float qz, F&& visit) {
  if (mode == 0)
    for_each_radius(g, qx, qy, qz, visit);
  else
    for_each_direct(g, mode, qx, qy, qz, visit);
}

// compute_max_neighbor_score's cell set (scoring/nvtl.rs:160-190): 3x3x3
// cells around floor(x * (1/res)) in f64, no distance test.
template <class F>
__device__ inline void for_each_cell27(const GridView& g, double x, double y, double z, F&& visit) {
  const double inv = 1.0 / (double)g.res;
  const int32_t vx = sat_i32d(floor(x * inv)), vy = sat_i32d(floor(y * inv)), vz = sat_i32d(floor(z * inv));
#pragma unroll 1
  for (int c = 0; c < 27; ++c) {
    const int32_t s = hash_find(g, vx + c / 9 - 1, vy + (c / 3) % 3 - 1, vz + c % 3 - 1);
    if (s >= 0) visit(s, g.vox[3 * s]);
  }
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

__device__ inline void snap_sincos(double a, double* s, double* c) {
  if (fabs(a) < 1e-5) {
    *s = 0.0;
    *c = 1.0;
  } else {
    sincos(a, s, c);
  }
}

__device__ inline void make_pose_tables(const double* pose, PoseTables* T) {
  T->t[0] = pose[0];
  T->t[1] = pose[1];
  T->t[2] = pose[2];
  double sr, cr, sp, cp, sy, cy;
  sincos(pose[3], &sr, &cr);
  sincos(pose[4], &sp, &cp);
  sincos(pose[5], &sy, &cy);
  T->R[0] = cp * cy;                T->R[1] = -cp * sy;               T->R[2] = sp;
  T->R[3] = sr * sp * cy + cr * sy; T->R[4] = cr * cy - sr * sp * sy; T->R[5] = -sr * cp;
  T->R[6] = sr * sy - cr * sp * cy; T->R[7] = cr * sp * sy + sr * cy; T->R[8] = cr * cp;
  double sx, cx, sz, cz;
  snap_sincos(pose[3], &sx, &cx);
  snap_sincos(pose[4], &sy, &cy);
  snap_sincos(pose[5], &sz, &cz);
  auto& J = T->j;
  J[0][0] = -sx * sz + cx * sy * cz; J[0][1] = -sx * cz - cx * sy * sz; J[0][2] = -cx * cy;
  J[1][0] = cx * sz + sx * sy * cz;  J[1][1] = cx * cz - sx * sy * sz;  J[1][2] = -sx * cy;
  J[2][0] = -sy * cz;                J[2][1] = sy * sz;                 J[2][2] = cy;
  J[3][0] = sx * cy * cz;            J[3][1] = -sx * cy * sz;           J[3][2] = sx * sy;
  J[4][0] = -cx * cy * cz;           J[4][1] = cx * cy * sz;            J[4][2] = -cx * sy;
  J[5][0] = -cy * sz;                J[5][1] = -cy * cz;                J[5][2] = 0.0;
  J[6][0] = cx * cz - sx * sy * sz;  J[6][1] = -cx * sz - sx * sy * cz; J[6][2] = 0.0;
  J[7][0] = sx * cz + cx * sy * sz;  J[7][1] = cx * sy * cz - sx * sz;  J[7][2] = 0.0;
  auto& H = T->h;
  H[0][0] = -cx * sz - sx * sy * cz;  H[0][1] = -cx * cz + sx * sy * sz;  H[0][2] = sx * cy;
  H[1][0] = -sx * sz + cx * sy * cz;  H[1][1] = -cx * sy * sz - sx * cz;  H[1][2] = -cx * cy;
  H[2][0] = cx * cy * cz;             H[2][1] = -cx * cy * sz;            H[2][2] = cx * sy;
  H[3][0] = sx * cy * cz;             H[3][1] = -sx * cy * sz;            H[3][2] = sx * sy;
  H[4][0] = -sx * cz - cx * sy * sz;  H[4][1] = sx * sz - cx * sy * cz;   H[4][2] = 0.0;
  H[5][0] = cx * cz - sx * sy * sz;   H[5][1] = -sx * sy * cz - cx * sz;  H[5][2] = 0.0;
  H[6][0] = -cy * cz;                 H[6][1] = cy * sz;                  H[6][2] = sy;
  H[7][0] = -sx * sy * cz;            H[7][1] = sx * sy * sz;             H[7][2] = sx * cy;
  H[8][0] = cx * sy * cz;             H[8][1] = -cx * sy * sz;            H[8][2] = -cx * cy;
  H[9][0] = sy * sz;                  H[9][1] = sy * cz;                  H[9][2] = 0.0;
  H[10][0] = -sx * cy * sz;           H[10][1] = -sx * cy * cz;           H[10][2] = 0.0;
  H[11][0] = cx * cy * sz;            H[11][1] = cx * cy * cz;            H[11][2] = 0.0;
  H[12][0] = -cy * cz;                H[12][1] = cy * sz;                 H[12][2] = 0.0;
  H[13][0] = -cx * sz - sx * sy * cz; H[13][1] = -cx * cz + sx * sy * sz; H[13][2] = 0.0;
  H[14][0] = -sx * sz + cx * sy * cz; H[14][1] = -cx * sy * sz - sx * cz; H[14][2] = 0.0;
}

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
constexpr int kAccSlots = 30; // + correspondences + points-with-neighbours/spare (as doubles)

__device__ inline int tri(int i, int j) { return i * (i + 1) / 2 + j; }  // i >= j

// One point-voxel pair, point-to-distribution metric: compute_derivative_single
// (derivatives/cpu.rs:139-229).  x = transformed point - mean, C = inverse
// covariance (row-major, as stored), J = the 3x6 point Jacobian (translation
// columns are the identity), ht = the 15 point-Hessian terms.
__device__ inline void pair_p2d(const double x[3], const double C[9], const double Jr[3][3], const double ht[15],
                                double d1, double d2, bool hess, double* acc) {
  double xc[3];
  for (int j = 0; j < 3; ++j) xc[j] = x[0] * C[j] + x[1] * C[3 + j] + x[2] * C[6 + j];
  const double xcx = xc[0] * x[0] + xc[1] * x[1] + xc[2] * x[2];
  const double e_raw = exp(-d2 * xcx * 0.5);
  double e = d2 * e_raw;
  if (!(e >= 0.0 && e <= 1.0) || !isfinite(e)) return;  // counted but contributes zero
  acc[0] += -d1 * e_raw;
  e *= d1;
  // C J: translation columns are C's columns; rotation columns C * Jr[:,c].
  double cJ[3][6];
  for (int r = 0; r < 3; ++r) {
    cJ[r][0] = C[3 * r + 0];
    cJ[r][1] = C[3 * r + 1];
    cJ[r][2] = C[3 * r + 2];
    for (int c = 0; c < 3; ++c) cJ[r][3 + c] = C[3 * r] * Jr[0][c] + C[3 * r + 1] * Jr[1][c] + C[3 * r + 2] * Jr[2][c];
  }
  double xcJ[6];
  for (int c = 0; c < 6; ++c) xcJ[c] = x[0] * cJ[0][c] + x[1] * cJ[1][c] + x[2] * cJ[2][c];
  for (int c = 0; c < 6; ++c) acc[1 + c] += e * xcJ[c];
  if (!hess) return;
  // (J^T C J)[j][i] = J[:,j] . (C J)[:,i]
  auto JcJ = [&](int jj, int ii) -> double {
    if (jj < 3) return cJ[jj][ii];
    return Jr[0][jj - 3] * cJ[0][ii] + Jr[1][jj - 3] * cJ[1][ii] + Jr[2][jj - 3] * cJ[2][ii];
  };
  // x^T C * d2T/dp_i dp_j for the rotation block (a..f with their zero rows).
  const double xa = xc[1] * ht[0] + xc[2] * ht[1];
  const double xb = xc[1] * ht[2] + xc[2] * ht[3];
  const double xcc = xc[1] * ht[4] + xc[2] * ht[5];
  const double xd = xc[0] * ht[6] + xc[1] * ht[7] + xc[2] * ht[8];
  const double xe = xc[0] * ht[9] + xc[1] * ht[10] + xc[2] * ht[11];
  const double xf = xc[0] * ht[12] + xc[1] * ht[13] + xc[2] * ht[14];
  const double xh[3][3] = {{xa, xb, xcc}, {xb, xd, xe}, {xcc, xe, xf}};
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      const double ph = (i >= 3 && j >= 3) ? xh[i - 3][j - 3] : 0.0;
      acc[7 + tri(i, j)] += e * (-d2 * xcJ[i] * xcJ[j] + ph + JcJ(j, i));
    }
}

// Point-to-plane metric: compute_derivative_point_to_plane (derivatives/cpu.rs:246-329).
__device__ inline void pair_p2plane(const double x[3], const double n[3], const double Jr[3][3], const double ht[15],
                                    double d1, double d2, bool hess, double* acc) {
  const double d = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
  const double dsq = d * d;
  const double e_raw = exp(-d2 * dsq * 0.5);
  const double e = d1 * d2 * e_raw;
  const double normalized = d2 * e_raw;
  if (!(normalized >= 0.0 && normalized <= 1.0) || !isfinite(normalized)) return;
  acc[0] += -d1 * e_raw;
  double dd[6];
  dd[0] = n[0];
  dd[1] = n[1];
  dd[2] = n[2];
  for (int c = 0; c < 3; ++c) dd[3 + c] = n[0] * Jr[0][c] + n[1] * Jr[1][c] + n[2] * Jr[2][c];
  for (int c = 0; c < 6; ++c) acc[1 + c] += e * d * dd[c];
  if (!hess) return;
  const double factor = -d2 * dsq + 1.0;
  const double na = n[1] * ht[0] + n[2] * ht[1];
  const double nb = n[1] * ht[2] + n[2] * ht[3];
  const double nc = n[1] * ht[4] + n[2] * ht[5];
  const double nd = n[0] * ht[6] + n[1] * ht[7] + n[2] * ht[8];
  const double ne = n[0] * ht[9] + n[1] * ht[10] + n[2] * ht[11];
  const double nf = n[0] * ht[12] + n[1] * ht[13] + n[2] * ht[14];
  const double nh[3][3] = {{na, nb, nc}, {nb, nd, ne}, {nc, ne, nf}};
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      const double ph = (i >= 3 && j >= 3) ? nh[i - 3][j - 3] : 0.0;
      acc[7 + tri(i, j)] += e * factor * (dd[i] * dd[j]) + e * d * ph;
    }
}

// ---------------------------------------------------------------------------
// Wavefront / workgroup reductions (f64).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Reduces `count` doubles held by every thread into out[] (valid in thread 0
// after the call; LDS scratch must hold (BLOCK/64) * count doubles).
template <int BLOCK>
__device__ inline void block_sum(const double* vals, int count, double* lds, double* out) {
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  constexpr int kWaves = BLOCK / kWave;
  for (int i = 0; i < count; ++i) {
    const double s = wave_sum(vals[i]);
    if (lane == 0) lds[wave * count + i] = s;
  }
  __syncthreads();
  if (threadIdx.x < count) {
    double s = 0.0;
    for (int w = 0; w < kWaves; ++w) s += lds[w * count + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Symmetric eigendecomposition, cyclic Jacobi, FP64 (device).  Reads the lower
// triangle.  v is row-major with eigenvectors as columns.
template <int N>
__device__ inline void sym_eigen(const double* a_in, double* ev, double* v) {
  double a[N][N];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j <= i; ++j) a[i][j] = a[j][i] = a_in[i * N + j];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) v[i * N + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < N; ++i) {
      diag += a[i][i] * a[i][i];
      for (int j = i + 1; j < N; ++j) off += a[i][j] * a[i][j];
    }
    if (off == 0.0 || off <= 1e-34 * diag) break;
    for (int p = 0; p < N - 1; ++p)
      for (int q = p + 1; q < N; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0);
        const double s = t * c;
        for (int k = 0; k < N; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < N; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        a[p][q] = a[q][p] = 0.0;
        for (int k = 0; k < N; ++k) {
          const double vkp = v[k * N + p], vkq = v[k * N + q];
          v[k * N + p] = c * vkp - s * vkq;
          v[k * N + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < N; ++i) ev[i] = a[i][i];
}

// Cholesky of a (6x6, lower triangle read); returns false if not positive definite.
__device__ inline bool cholesky6(const double* a, double L[6][6]) {
  for (int j = 0; j < 6; ++j) {
    double d = a[j * 6 + j];
    for (int k = 0; k < j; ++k) d -= L[j][k] * L[j][k];
    if (!(d > 0.0)) return false;
    const double ljj = sqrt(d);
    L[j][j] = ljj;
    for (int i = j + 1; i < 6; ++i) {
      double s = a[i * 6 + j];
      for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
      L[i][j] = s / ljj;
    }
  }
  return true;
}

// newton_step_negative_definite (optimization/newton.rs:160-176): Newton step
// for a maximisation with min eigenvalue -1e-3 and SVD tolerance 1e-10.  When
// -H - 1e-10 I is positive definite, every eigenvalue of H is below -1e-10,
// so the reference's SVD pseudo-inverse equals the plain inverse and a
// Cholesky solve of (-H) d = g gives the same step; otherwise the Jacobi
// eigendecomposition applies the reference's clamp and pseudo-inverse.
__device__ inline bool newton_step_device(const double g[6], const double H[36], double d[6]) {
  double L[6][6];
  double A[36];
  for (int i = 0; i < 36; ++i) A[i] = -H[i];
  for (int i = 0; i < 6; ++i) A[i * 6 + i] -= 1e-10;
  if (cholesky6(A, L)) {
    for (int i = 0; i < 36; ++i) A[i] = -H[i];
    if (cholesky6(A, L)) {
      double y[6];
      for (int i = 0; i < 6; ++i) {
        double s = g[i];
        for (int k = 0; k < i; ++k) s -= L[i][k] * y[k];
        y[i] = s / L[i][i];
      }
      for (int i = 5; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 6; ++k) s -= L[k][i] * d[k];
        d[i] = s / L[i][i];
      }
      return true;
    }
  }
  double ev[6], V[36];
  sym_eigen<6>(H, ev, V);
  bool neg_def = true;
  for (int i = 0; i < 6; ++i) neg_def &= (ev[i] < 0.0);
  double lam[6];
  if (neg_def) {
    for (int i = 0; i < 6; ++i) lam[i] = ev[i];
  } else {
    for (int i = 0; i < 6; ++i) lam[i] = ev[i] > -1e-3 ? -1e-3 : ev[i];
  }
  double max_sv = 0.0;
  for (int i = 0; i < 6; ++i) max_sv = fmax(max_sv, fabs(lam[i]));
  if (max_sv < 1e-10) return false;
  for (int i = 0; i < 6; ++i) d[i] = 0.0;
  for (int k = 0; k < 6; ++k) {
    if (!(fabs(lam[k]) > 1e-10)) continue;
    double proj = 0.0;
    for (int i = 0; i < 6; ++i) proj += V[i * 6 + k] * (-g[i]);
    const double c = proj / lam[k];
    for (int i = 0; i < 6; ++i) d[i] += c * V[i * 6 + k];
  }
  return true;
}

}  // namespace ndt
