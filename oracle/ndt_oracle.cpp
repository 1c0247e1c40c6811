// CPU oracle for the NDT scan-matching hot path.
//
// TEST INFRASTRUCTURE ONLY: a plain C++/FP64 restatement of the reference
// crate's CPU path (NDT_USE_GPU=0).  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg load it, always as the checker.  The product
// library never links it.
//
// Citations are to the reference repository (src/ndt_cuda/src/... unless
// stated otherwise).  The reference is Rust and cannot be compiled here (no
// Rust toolchain), so this file is written from reading it.  Third-party
// arithmetic it restates: nalgebra 0.33.2 symmetric_eigen / SVD solve /
// Euler conversions (replaced by a cyclic Jacobi eigensolver and the
// equivalent pseudo-inverse), kiddo 4.2.1 ImmutableKdTree::within (replaced
// by an exact cell-binned scan with the same f32 squared-distance test).
//
// Compile with -ffp-contract=off: Rust never fuses multiply-adds, and the f32
// radius test must round exactly as kiddo's does.

#include "ndt_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <utility>
#include <vector>
#include <tuple>
#include <map>

namespace {

struct Coord {
  int32_t x, y, z;
  bool operator==(const Coord& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator<(const Coord& o) const {
    if (x != o.x) return x < o.x;
    if (y != o.y) return y < o.y;
    return z < o.z;
  }
};

struct CoordHash {
  size_t operator()(const Coord& c) const {
    uint64_t h = (uint64_t)(uint32_t)c.x * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)c.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
    h ^= (uint64_t)(uint32_t)c.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

// Rust's `f32 as i32` saturates and maps NaN to 0.
int32_t sat_i32(float v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483647.0f) return std::numeric_limits<int32_t>::max();
  if (v <= -2147483648.0f) return std::numeric_limits<int32_t>::min();
  return (int32_t)v;
}
int32_t sat_i32d(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483647.0) return std::numeric_limits<int32_t>::max();
  if (v <= -2147483648.0) return std::numeric_limits<int32_t>::min();
  return (int32_t)v;
}

// VoxelCoord::from_point (voxel_grid/types.rs:206-212): f32 division, floor.
Coord key_of(const float* p, float res) {
  return Coord{sat_i32(std::floor(p[0] / res)), sat_i32(std::floor(p[1] / res)),
               sat_i32(std::floor(p[2] / res))};
}

// ---------------------------------------------------------------------------
// Symmetric eigendecomposition: cyclic Jacobi (Golub & Van Loan, Alg. 8.5.3).
// Stands in for nalgebra's SymmetricEigen, which reads only the lower
// triangle; so does this.  Eigenvalues are returned unordered (as nalgebra's
// are), eigenvectors as columns of v (row-major v[i*N+j] = component i of
// vector j).
template <int N>
void jacobi_eigen(const double* a_in, double* evals, double* v) {
  double a[N][N];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j <= i; ++j) a[i][j] = a[j][i] = a_in[i * N + j];
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) v[i * N + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < N; ++i) {
      diag += a[i][i] * a[i][i];
      for (int j = i + 1; j < N; ++j) off += a[i][j] * a[i][j];
    }
    if (off == 0.0 || off <= 1e-34 * diag) break;
    for (int p = 0; p < N - 1; ++p) {
      for (int q = p + 1; q < N; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0);
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
  }
  for (int i = 0; i < N; ++i) evals[i] = a[i][i];
}

// ---------------------------------------------------------------------------
struct Voxel {
  Coord coord;
  float mean[3];
  float cov[9];
  float inv[9];
  float axis[3];
  uint32_t count;
  double mean64[3];
  double inv64[9];
};

struct Acc {
  double sum[3] = {0, 0, 0};
  double sq[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t count = 0;
};

// Voxel::from_statistics + regularize_covariance (voxel_grid/types.rs:54-190).
bool from_statistics(const Acc& acc, uint32_t min_points, float ratio, Voxel* out) {
  if (acc.count < min_points) return false;  // types.rs:60-62
  const double n = (double)acc.count;
  double mean[3];
  for (int i = 0; i < 3; ++i) mean[i] = acc.sum[i] / n;  // types.rs:67
  double cov[9];
  for (int i = 0; i < 3; ++i)  // types.rs:78-79: (sum_sq - (mu mu^T) n) / (n - 1)
    for (int j = 0; j < 3; ++j) cov[i * 3 + j] = (acc.sq[i * 3 + j] - (mean[i] * mean[j]) * n) / (n - 1.0);

  double ev[3], V[9];
  jacobi_eigen<3>(cov, ev, V);  // types.rs:120
  double max_ev = 0.0;          // types.rs:124: fold(0.0, f64::max)
  for (double e : ev) max_ev = std::max(max_ev, e);
  if (max_ev <= 0.0) return false;  // types.rs:126-128
  int min_idx = 0;                  // types.rs:131-136
  for (int i = 1; i < 3; ++i)
    if (ev[i] < ev[min_idx]) min_idx = i;
  const double thr = max_ev * (double)ratio;  // types.rs:138
  bool needs = false;
  for (double e : ev) needs |= (e < thr);
  double clamped[3];
  for (int i = 0; i < 3; ++i) clamped[i] = ev[i] < thr ? thr : ev[i];
  double reg[9];
  if (needs) {  // types.rs:150-161: V diag V^T only when clamping happened
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += V[i * 3 + k] * clamped[k] * V[j * 3 + k];
        reg[i * 3 + j] = s;
      }
  } else {
    std::memcpy(reg, cov, sizeof(reg));
  }
  double inv[9];  // types.rs:164-180: inverse always from clamped eigenvalues
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += V[i * 3 + k] * (1.0 / clamped[k]) * V[j * 3 + k];
      inv[i * 3 + j] = s;
    }
  for (int i = 0; i < 3; ++i) {
    out->mean[i] = (float)mean[i];
    out->mean64[i] = mean[i];
    out->axis[i] = (float)V[i * 3 + min_idx];  // types.rs:183
  }
  for (int i = 0; i < 9; ++i) {
    out->cov[i] = (float)reg[i];
    out->inv[i] = (float)inv[i];
    out->inv64[i] = inv[i];
  }
  out->count = acc.count;
  return true;
}

}  // namespace

struct oracle_grid {
  float resolution;
  std::vector<Voxel> voxels;  // sorted by coordinate
  std::unordered_map<Coord, int32_t, CoordHash> by_coord;
  // Radius-search bins: voxel means binned by floor(mean / res) in f64.
  std::unordered_map<Coord, std::vector<int32_t>, CoordHash> bins;
};

namespace {

inline float dist2_f32(const float* q, const float* m) {
  // kiddo SquaredEuclidean: fold over axes of (a-b)^2 in f32, no fusing.
  const float dx = q[0] - m[0];
  const float dy = q[1] - m[1];
  const float dz = q[2] - m[2];
  float s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return s;
}

// VoxelSearch::within (voxel_grid/search.rs:90-99) with radius = resolution:
// every voxel mean whose f32 squared distance is < r*r (kiddo compares with a
// strict '<'; the boundary tie is not pinned by any reference test), nearest
// first.  Completeness: a mean within r of q lies in a bin between
// floor((q-r)/res) and floor((q+r)/res); the 1e-7 widening covers rounding.
void radius_search(const oracle_grid* g, const float* q, std::vector<int32_t>& out) {
  out.clear();
  const float r = g->resolution;
  const float r2 = r * r;
  const double res = (double)r;
  int32_t lo[3], hi[3];
  for (int i = 0; i < 3; ++i) {
    lo[i] = sat_i32d(std::floor(((double)q[i] - res) / res - 1e-7));
    hi[i] = sat_i32d(std::floor(((double)q[i] + res) / res + 1e-7));
  }
  std::vector<std::pair<float, int32_t>> hits;
  for (int32_t x = lo[0]; x <= hi[0]; ++x)
    for (int32_t y = lo[1]; y <= hi[1]; ++y)
      for (int32_t z = lo[2]; z <= hi[2]; ++z) {
        auto it = g->bins.find(Coord{x, y, z});
        if (it == g->bins.end()) continue;
        for (int32_t vi : it->second) {
          const float d = dist2_f32(q, g->voxels[vi].mean);
          if (d < r2) hits.emplace_back(d, vi);
        }
      }
  std::sort(hits.begin(), hits.end());
  for (auto& h : hits) out.push_back(h.second);
}

// Direct cell lookups following Autoware's pclomp VoxelGridCovariance
// (getNeighborhoodAtPoint / ...7 / ...1): cell = floor(q / res) in f32, the
// listed relative cells, occupied cells only.  DIRECT26 uses PCL's
// getAllNeighborCellIndices, i.e. the 26 cells around the point's own cell.
void direct_search(const oracle_grid* g, const float* q, int mode, std::vector<int32_t>& out) {
  out.clear();
  const Coord c = key_of(q, g->resolution);
  auto probe = [&](int dx, int dy, int dz) {
    auto it = g->by_coord.find(Coord{c.x + dx, c.y + dy, c.z + dz});
    if (it != g->by_coord.end()) out.push_back(it->second);
  };
  if (mode == ORACLE_SEARCH_DIRECT1) {
    probe(0, 0, 0);
  } else if (mode == ORACLE_SEARCH_DIRECT7) {
    probe(0, 0, 0);
    probe(1, 0, 0);
    probe(-1, 0, 0);
    probe(0, 1, 0);
    probe(0, -1, 0);
    probe(0, 0, 1);
    probe(0, 0, -1);
  } else {  // DIRECT26
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz)
          if (dx || dy || dz) probe(dx, dy, dz);
  }
}

// compute_max_neighbor_score's cell scan (scoring/nvtl.rs:160-190):
// floor(x * (1/res)) in f64, 3x3x3 cells, no distance test.
void cell27_search(const oracle_grid* g, const double* x, std::vector<int32_t>& out) {
  out.clear();
  const double inv = 1.0 / (double)g->resolution;
  const int32_t vx = sat_i32d(std::floor(x[0] * inv));
  const int32_t vy = sat_i32d(std::floor(x[1] * inv));
  const int32_t vz = sat_i32d(std::floor(x[2] * inv));
  for (int di = -1; di <= 1; ++di)
    for (int dj = -1; dj <= 1; ++dj)
      for (int dk = -1; dk <= 1; ++dk) {
        auto it = g->by_coord.find(Coord{vx + di, vy + dj, vz + dk});
        if (it != g->by_coord.end()) out.push_back(it->second);
      }
}

void search(const oracle_grid* g, int mode, const float* q, std::vector<int32_t>& out) {
  if (mode == ORACLE_SEARCH_KDTREE)
    radius_search(g, q, out);
  else
    direct_search(g, q, mode, out);
}

struct Gauss {
  double d1, d2;
};

// GaussianParams::new (derivatives/types.rs:61-75).
Gauss make_gauss(double res, double outlier) {
  const double c1 = 10.0 * (1.0 - outlier);
  const double c2 = outlier / (res * res * res);
  const double d3 = -std::log(c2);
  const double d1 = -std::log(c1 + c2) - d3;
  const double d2_nom = -std::log(c1 * std::exp(-0.5) + c2) - d3;
  const double d2 = -2.0 * std::log(d2_nom / d1);
  return Gauss{d1, d2};
}

// AngularDerivatives::new (derivatives/angular.rs:37-183).
struct Angular {
  double j[8][3];
  double h[15][3];
};

void sin_cos_snap(double a, double* s, double* c) {
  if (std::fabs(a) < 1e-5) {  // angular.rs:12,39-55
    *s = 0.0;
    *c = 1.0;
  } else {
    *s = std::sin(a);
    *c = std::cos(a);
  }
}

Angular make_angular(double roll, double pitch, double yaw, bool hess) {
  double sx, cx, sy, cy, sz, cz;
  sin_cos_snap(roll, &sx, &cx);
  sin_cos_snap(pitch, &sy, &cy);
  sin_cos_snap(yaw, &sz, &cz);
  Angular A;
  std::memset(&A, 0, sizeof(A));
  // Eq. 6.19 (angular.rs:58-99)
  A.j[0][0] = -sx * sz + cx * sy * cz; A.j[0][1] = -sx * cz - cx * sy * sz; A.j[0][2] = -cx * cy;
  A.j[1][0] = cx * sz + sx * sy * cz;  A.j[1][1] = cx * cz - sx * sy * sz;  A.j[1][2] = -sx * cy;
  A.j[2][0] = -sy * cz;                A.j[2][1] = sy * sz;                 A.j[2][2] = cy;
  A.j[3][0] = sx * cy * cz;            A.j[3][1] = -sx * cy * sz;           A.j[3][2] = sx * sy;
  A.j[4][0] = -cx * cy * cz;           A.j[4][1] = cx * cy * sz;            A.j[4][2] = -cx * sy;
  A.j[5][0] = -cy * sz;                A.j[5][1] = -cy * cz;
  A.j[6][0] = cx * cz - sx * sy * sz;  A.j[6][1] = -cx * sz - sx * sy * cz;
  A.j[7][0] = sx * cz + cx * sy * sz;  A.j[7][1] = cx * sy * cz - sx * sz;
  if (hess) {  // Eq. 6.21 (angular.rs:104-181)
    A.h[0][0] = -cx * sz - sx * sy * cz;  A.h[0][1] = -cx * cz + sx * sy * sz;  A.h[0][2] = sx * cy;
    A.h[1][0] = -sx * sz + cx * sy * cz;  A.h[1][1] = -cx * sy * sz - sx * cz;  A.h[1][2] = -cx * cy;
    A.h[2][0] = cx * cy * cz;             A.h[2][1] = -cx * cy * sz;            A.h[2][2] = cx * sy;
    A.h[3][0] = sx * cy * cz;             A.h[3][1] = -sx * cy * sz;            A.h[3][2] = sx * sy;
    A.h[4][0] = -sx * cz - cx * sy * sz;  A.h[4][1] = sx * sz - cx * sy * cz;
    A.h[5][0] = cx * cz - sx * sy * sz;   A.h[5][1] = -sx * sy * cz - cx * sz;
    A.h[6][0] = -cy * cz;                 A.h[6][1] = cy * sz;                  A.h[6][2] = sy;
    A.h[7][0] = -sx * sy * cz;            A.h[7][1] = sx * sy * sz;             A.h[7][2] = sx * cy;
    A.h[8][0] = cx * sy * cz;             A.h[8][1] = -cx * sy * sz;            A.h[8][2] = -cx * cy;
    A.h[9][0] = sy * sz;                  A.h[9][1] = sy * cz;
    A.h[10][0] = -sx * cy * sz;           A.h[10][1] = -sx * cy * cz;
    A.h[11][0] = cx * cy * sz;            A.h[11][1] = cx * cy * cz;
    A.h[12][0] = -cy * cz;                A.h[12][1] = cy * sz;
    A.h[13][0] = -cx * sz - sx * sy * cz; A.h[13][1] = -cx * cz + sx * sy * sz;
    A.h[14][0] = -sx * sz + cx * sy * cz; A.h[14][1] = -cx * sy * sz - sx * cz;
  }
  return A;
}

// transform_point (derivatives/cpu.rs:339-372): R = Rx(roll) Ry(pitch) Rz(yaw).
void transform_point(const double* p, const double* pose, double* out) {
  const double sr = std::sin(pose[3]), cr = std::cos(pose[3]);
  const double sp = std::sin(pose[4]), cp = std::cos(pose[4]);
  const double sy = std::sin(pose[5]), cy = std::cos(pose[5]);
  const double r00 = cp * cy, r01 = -cp * sy, r02 = sp;
  const double r10 = sr * sp * cy + cr * sy, r11 = cr * cy - sr * sp * sy, r12 = -sr * cp;
  const double r20 = sr * sy - cr * sp * cy, r21 = cr * sp * sy + sr * cy, r22 = cr * cp;
  out[0] = r00 * p[0] + r01 * p[1] + r02 * p[2] + pose[0];
  out[1] = r10 * p[0] + r11 * p[1] + r12 * p[2] + pose[1];
  out[2] = r20 * p[0] + r21 * p[1] + r22 * p[2] + pose[2];
}

// nalgebra UnitQuaternion::from_euler_angles(roll, pitch, yaw) (w, i, j, k).
void quat_from_euler(double roll, double pitch, double yaw, double q[4]) {
  const double sr = std::sin(roll * 0.5), cr = std::cos(roll * 0.5);
  const double sp = std::sin(pitch * 0.5), cp = std::cos(pitch * 0.5);
  const double sy = std::sin(yaw * 0.5), cy = std::cos(yaw * 0.5);
  q[0] = cr * cp * cy + sr * sp * sy;
  q[1] = sr * cp * cy - cr * sp * sy;
  q[2] = cr * sp * cy + sr * cp * sy;
  q[3] = cr * cp * sy - sr * sp * cy;
}

// Isometry3 * Point3 with the quaternion rotation v + w t + u x t, t = 2 u x v.
void iso_transform(const double q[4], const double* t, const double* p, double* out) {
  const double ux = q[1], uy = q[2], uz = q[3], w = q[0];
  const double tx = (uy * p[2] - uz * p[1]) * 2.0;
  const double ty = (uz * p[0] - ux * p[2]) * 2.0;
  const double tz = (ux * p[1] - uy * p[0]) * 2.0;
  const double cx = uy * tz - uz * ty;
  const double cy = uz * tx - ux * tz;
  const double cz = ux * ty - uy * tx;
  out[0] = tx * w + cx + p[0] + t[0];
  out[1] = ty * w + cy + p[1] + t[1];
  out[2] = tz * w + cz + p[2] + t[2];
}

// compute_point_derivatives (derivatives/cpu.rs:27-123): J is 4x6, pH is 24x6.
void point_derivatives(const double* p, const Angular& A, bool hess, double J[4][6], double pH[24][6]) {
  std::memset(J, 0, sizeof(double) * 24);
  J[0][0] = 1.0;
  J[1][1] = 1.0;
  J[2][2] = 1.0;
  double gt[8];
  for (int r = 0; r < 8; ++r) gt[r] = A.j[r][0] * p[0] + A.j[r][1] * p[1] + A.j[r][2] * p[2];
  J[1][3] = gt[0]; J[2][3] = gt[1];
  J[0][4] = gt[2]; J[1][4] = gt[3]; J[2][4] = gt[4];
  J[0][5] = gt[5]; J[1][5] = gt[6]; J[2][5] = gt[7];
  std::memset(pH, 0, sizeof(double) * 144);
  if (!hess) return;
  double ht[15];
  for (int r = 0; r < 15; ++r) ht[r] = A.h[r][0] * p[0] + A.h[r][1] * p[1] + A.h[r][2] * p[2];
  pH[13][3] = ht[0];  pH[14][3] = ht[1];                       // a
  pH[17][3] = ht[2];  pH[18][3] = ht[3];                       // b
  pH[21][3] = ht[4];  pH[22][3] = ht[5];                       // c
  pH[13][4] = ht[2];  pH[14][4] = ht[3];                       // b
  pH[16][4] = ht[6];  pH[17][4] = ht[7];  pH[18][4] = ht[8];   // d
  pH[20][4] = ht[9];  pH[21][4] = ht[10]; pH[22][4] = ht[11];  // e
  pH[13][5] = ht[4];  pH[14][5] = ht[5];                       // c
  pH[16][5] = ht[9];  pH[17][5] = ht[10]; pH[18][5] = ht[11];  // e
  pH[20][5] = ht[12]; pH[21][5] = ht[13]; pH[22][5] = ht[14];  // f
}

struct Deriv {
  double score = 0.0;
  double g[6] = {0, 0, 0, 0, 0, 0};
  double H[6][6] = {};
  uint32_t corr = 0;
};

// compute_derivative_single (derivatives/cpu.rs:139-229).  Returns false for
// the guarded "zero contribution" case; the caller still counts the pair
// (types.rs:188-193).
bool derivative_single(const double x[3], const double Ci[9], const double J[4][6], const double pH[24][6],
                       const Gauss& gs, bool hess, double* score, double g[6], double H[6][6]) {
  double c4[4][4] = {};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c4[i][j] = Ci[i * 3 + j];
  const double x4[4] = {x[0], x[1], x[2], 0.0};
  double xc[4];  // x4^T * c4
  for (int j = 0; j < 4; ++j) {
    double s = 0.0;
    for (int k = 0; k < 4; ++k) s += x4[k] * c4[k][j];
    xc[j] = s;
  }
  double xcx = 0.0;
  for (int k = 0; k < 4; ++k) xcx += xc[k] * x4[k];
  const double e_raw = std::exp(-gs.d2 * xcx * 0.5);
  const double sc = -gs.d1 * e_raw;
  double e = gs.d2 * e_raw;
  if (!(e >= 0.0 && e <= 1.0) || !std::isfinite(e)) return false;  // cpu.rs:170-172
  e *= gs.d1;
  double cJ[4][6];  // c4 * J
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 6; ++j) {
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += c4[i][k] * J[k][j];
      cJ[i][j] = s;
    }
  double xcJ[6];
  for (int j = 0; j < 6; ++j) {
    double s = 0.0;
    for (int k = 0; k < 4; ++k) s += x4[k] * cJ[k][j];
    xcJ[j] = s;
  }
  *score = sc;
  for (int j = 0; j < 6; ++j) g[j] = e * xcJ[j];
  if (hess) {
    double JcJ[6][6];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += J[k][i] * cJ[k][j];
        JcJ[i][j] = s;
      }
    for (int i = 0; i < 6; ++i) {
      double xcH[6];
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < 4; ++k) s += xc[k] * pH[i * 4 + k][j];
        xcH[j] = s;
      }
      for (int j = 0; j < 6; ++j)
        H[i][j] = e * (-gs.d2 * xcJ[i] * xcJ[j] + xcH[j] + JcJ[j][i]);  // Eq. 6.13
    }
  } else {
    std::memset(H, 0, sizeof(double) * 36);
  }
  return true;
}

// compute_derivative_point_to_plane (derivatives/cpu.rs:246-329).
bool derivative_p2plane(const double x[3], const double n[3], const double J[4][6], const double pH[24][6],
                        const Gauss& gs, bool hess, double* score, double g[6], double H[6][6]) {
  const double d = x[0] * n[0] + x[1] * n[1] + x[2] * n[2];
  const double d2v = d * d;
  const double e_raw = std::exp(-gs.d2 * d2v * 0.5);
  const double sc = -gs.d1 * e_raw;
  const double e = gs.d1 * gs.d2 * e_raw;
  const double normalized = gs.d2 * e_raw;
  if (!(normalized >= 0.0 && normalized <= 1.0) || !std::isfinite(normalized)) return false;
  double dd[6];
  for (int j = 0; j < 6; ++j) dd[j] = n[0] * J[0][j] + n[1] * J[1][j] + n[2] * J[2][j];
  *score = sc;
  for (int j = 0; j < 6; ++j) g[j] = e * d * dd[j];
  std::memset(H, 0, sizeof(double) * 36);
  if (hess) {
    const double factor = -gs.d2 * d2v + 1.0;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) H[i][j] += e * factor * (dd[i] * dd[j]);
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        const double nph = n[0] * pH[i * 4 + 0][j] + n[1] * pH[i * 4 + 1][j] + n[2] * pH[i * 4 + 2][j];
        H[i][j] += e * d * nph;
      }
  }
  return true;
}

Deriv derivatives(const oracle_grid* g, const float* xyz, size_t n, const double* pose, const Gauss& gs,
                  bool hess, int mode, int metric, bool fp64) {
  Deriv R;
  const Angular A = make_angular(pose[3], pose[4], pose[5], hess);
  std::vector<int32_t> nb;
  double J[4][6], pH[24][6];
  for (size_t i = 0; i < n; ++i) {
    const double p[3] = {(double)xyz[3 * i], (double)xyz[3 * i + 1], (double)xyz[3 * i + 2]};
    double t[3];
    transform_point(p, pose, t);
    const float tf[3] = {(float)t[0], (float)t[1], (float)t[2]};  // cpu.rs:452-456
    search(g, mode, tf, nb);
    if (nb.empty()) continue;
    point_derivatives(p, A, hess, J, pH);
    for (int32_t vi : nb) {
      const Voxel& v = g->voxels[vi];
      double x[3], Ci[9];
      for (int k = 0; k < 3; ++k) x[k] = t[k] - (fp64 ? v.mean64[k] : (double)v.mean[k]);
      for (int k = 0; k < 9; ++k) Ci[k] = fp64 ? v.inv64[k] : (double)v.inv[k];
      double s, gg[6], HH[6][6];
      bool ok;
      if (metric == 1) {
        const double nn[3] = {(double)v.axis[0], (double)v.axis[1], (double)v.axis[2]};
        ok = derivative_p2plane(x, nn, J, pH, gs, hess, &s, gg, HH);
      } else {
        ok = derivative_single(x, Ci, J, pH, gs, hess, &s, gg, HH);
      }
      R.corr += 1;  // AggregatedDerivatives::add counts every pair (types.rs:188-193)
      if (!ok) continue;
      R.score += s;
      for (int a = 0; a < 6; ++a) R.g[a] += gg[a];
      for (int a = 0; a < 6; ++a)
        for (int b = 0; b < 6; ++b) R.H[a][b] += HH[a][b];
    }
  }
  return R;
}

// newton_step (optimization/newton.rs:21-38): SVD pseudo-inverse solve of
// H d = -g with absolute singular-value tolerance.  For a symmetric H the
// singular values are |lambda| and the SVD solve equals
// sum_i v_i (v_i . -g) / lambda_i over |lambda_i| > tol.
bool newton_step_svd(const double g[6], const double H[36], double tol, double d[6]) {
  double ev[6], V[36];
  jacobi_eigen<6>(H, ev, V);
  double max_sv = 0.0;
  for (double e : ev) max_sv = std::max(max_sv, std::fabs(e));
  if (max_sv < tol) return false;
  for (int i = 0; i < 6; ++i) d[i] = 0.0;
  for (int k = 0; k < 6; ++k) {
    if (!(std::fabs(ev[k]) > tol)) continue;
    double proj = 0.0;
    for (int i = 0; i < 6; ++i) proj += V[i * 6 + k] * (-g[i]);
    const double c = proj / ev[k];
    for (int i = 0; i < 6; ++i) d[i] += c * V[i * 6 + k];
  }
  return true;
}

// newton_step_negative_definite (newton.rs:160-176) with
// regularize_hessian_negative_definite (newton.rs:119-144).
bool newton_step_nd(const double g[6], const double H[36], double min_ev, double tol, double d[6]) {
  double ev[6], V[36];
  jacobi_eigen<6>(H, ev, V);
  bool neg_def = true;
  for (double e : ev) neg_def &= (e < 0.0);
  if (neg_def) return newton_step_svd(g, H, tol, d);
  double cl[6];
  bool modified = false;
  for (int i = 0; i < 6; ++i) {
    cl[i] = ev[i];
    if (cl[i] > min_ev) {
      cl[i] = min_ev;
      modified = true;
    }
  }
  double Hr[36];
  if (!modified) {
    std::memcpy(Hr, H, sizeof(Hr));
  } else {
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
        for (int k = 0; k < 6; ++k) s += V[i * 6 + k] * cl[k] * V[j * 6 + k];
        Hr[i * 6 + j] = s;
      }
  }
  return newton_step_svd(g, Hr, tol, d);
}

// More-Thuente (optimization/more_thuente.rs:90-420).
double trial_value_selection(double a_l, double f_l, double g_l, double a_u, double f_u, double g_u, double a_t,
                             double f_t, double g_t) {
  if (f_t > f_l) {
    const double z = 3.0 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
    const double w_sq = z * z - g_t * g_l;
    if (w_sq < 0.0) return a_l - 0.5 * (a_l - a_t) * g_l / (g_l - (f_l - f_t) / (a_l - a_t));
    const double w = std::sqrt(w_sq);
    const double a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2.0 * w);
    const double denom = g_l - (f_l - f_t) / (a_l - a_t);
    if (std::fabs(denom) < 1e-15) return a_c;
    const double a_q = a_l - 0.5 * (a_l - a_t) * g_l / denom;
    if (std::fabs(a_c - a_l) < std::fabs(a_q - a_l)) return a_c;
    return 0.5 * (a_q + a_c);
  }
  if (g_t * g_l < 0.0) {
    const double z = 3.0 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
    const double w_sq = z * z - g_t * g_l;
    if (w_sq < 0.0) return a_l - (a_l - a_t) / (g_l - g_t) * g_l;
    const double w = std::sqrt(w_sq);
    const double a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2.0 * w);
    const double a_s = a_l - (a_l - a_t) / (g_l - g_t) * g_l;
    if (std::fabs(a_c - a_t) >= std::fabs(a_s - a_t)) return a_c;
    return a_s;
  }
  if (std::fabs(g_t) <= std::fabs(g_l)) {
    const double z = 3.0 * (f_t - f_l) / (a_t - a_l) - g_t - g_l;
    const double w_sq = z * z - g_t * g_l;
    double a_c;
    if (w_sq >= 0.0) {
      const double w = std::sqrt(w_sq);
      a_c = a_l + (a_t - a_l) * (w - g_l - z) / (g_t - g_l + 2.0 * w);
    } else {
      a_c = (a_t > a_l) ? a_t + 0.66 * (a_u - a_t) : a_t - 0.66 * (a_t - a_u);
    }
    const double a_s = (std::fabs(g_l - g_t) > 1e-15) ? a_l - (a_l - a_t) / (g_l - g_t) * g_l : a_c;
    const double a_t_next = (std::fabs(a_c - a_t) < std::fabs(a_s - a_t)) ? a_c : a_s;
    if (a_t > a_l) return std::min(a_t + 0.66 * (a_u - a_t), a_t_next);
    return std::max(a_t + 0.66 * (a_u - a_t), a_t_next);
  }
  const double z = 3.0 * (f_t - f_u) / (a_t - a_u) - g_t - g_u;
  const double w_sq = z * z - g_t * g_u;
  if (w_sq < 0.0) return 0.5 * (a_l + a_u);
  const double w = std::sqrt(w_sq);
  return a_u + (a_t - a_u) * (w - g_u - z) / (g_t - g_u + 2.0 * w);
}

bool update_interval(double& a_l, double& f_l, double& g_l, double& a_u, double& f_u, double& g_u, double a_t,
                     double f_t, double g_t) {
  if (f_t > f_l) {
    a_u = a_t; f_u = f_t; g_u = g_t;
    return false;
  }
  if (g_t * (a_l - a_t) > 0.0) {
    a_l = a_t; f_l = f_t; g_l = g_t;
    return false;
  }
  if (g_t * (a_l - a_t) < 0.0) {
    a_u = a_l; f_u = f_l; g_u = g_l;
    a_l = a_t; f_l = f_t; g_l = g_t;
    return false;
  }
  return true;
}

template <class F>
void more_thuente(F&& fn, double init_score, double init_deriv, double init_step, double step_max, double step_min,
                  double mu, double nu, uint32_t max_it, double* step, int32_t* conv, uint32_t* iters, double* fval) {
  auto clampd = [](double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); };
  if (init_deriv <= 0.0) {
    *step = 0.0; *conv = 0; *iters = 0; *fval = init_score;
    return;
  }
  const double phi_0 = -init_score, dphi_0 = -init_deriv;
  double a_t = clampd(init_step, step_min, step_max);
  double s_t, d_t;
  fn(a_t, &s_t, &d_t);
  double phi_t = -s_t, dphi_t = -d_t;
  if (step_max - step_min < 0.0) {
    *step = a_t; *conv = 1; *iters = 0; *fval = s_t;
    return;
  }
  auto psi = [&](double a, double phi_a) { return phi_a - phi_0 - mu * dphi_0 * a; };
  auto dpsi = [&](double dphi_a) { return dphi_a - mu * dphi_0; };
  double a_l = 0.0, a_u = 0.0;
  double f_l = psi(a_l, phi_0), g_l = dpsi(dphi_0);
  double f_u = psi(a_u, phi_0), g_u = dpsi(dphi_0);
  bool open = true;
  uint32_t it = 0;
  double psi_t = psi(a_t, phi_t), dpsi_t = dpsi(dphi_t);
  while (it < max_it) {
    if (psi_t <= 0.0 && std::fabs(dphi_t) <= nu * std::fabs(dphi_0)) {
      *step = a_t; *conv = 1; *iters = it; *fval = -phi_t;
      return;
    }
    if (open)
      a_t = trial_value_selection(a_l, f_l, g_l, a_u, f_u, g_u, a_t, psi_t, dpsi_t);
    else
      a_t = trial_value_selection(a_l, f_l, g_l, a_u, f_u, g_u, a_t, phi_t, dphi_t);
    a_t = clampd(a_t, step_min, step_max);
    fn(a_t, &s_t, &d_t);
    phi_t = -s_t;
    dphi_t = -d_t;
    psi_t = psi(a_t, phi_t);
    dpsi_t = dpsi(dphi_t);
    if (open && psi_t <= 0.0 && dpsi_t >= 0.0) {
      open = false;
      f_l = f_l + phi_0 - mu * dphi_0 * a_l;
      g_l += mu * dphi_0;
      f_u = f_u + phi_0 - mu * dphi_0 * a_u;
      g_u += mu * dphi_0;
    }
    bool done = open ? update_interval(a_l, f_l, g_l, a_u, f_u, g_u, a_t, psi_t, dpsi_t)
                     : update_interval(a_l, f_l, g_l, a_u, f_u, g_u, a_t, phi_t, dphi_t);
    if (done) break;
    ++it;
  }
  *step = a_t; *conv = 0; *iters = it; *fval = -phi_t;
}

// RegularizationTerm::compute_derivatives (optimization/regularization.rs:102-159).
void regularization(const oracle_params* p, const double* pose, uint32_t corr, double* s, double g[6],
                    double H[36]) {
  *s = 0.0;
  for (int i = 0; i < 6; ++i) g[i] = 0.0;
  for (int i = 0; i < 36; ++i) H[i] = 0.0;
  if (!p->regularization_enabled || !p->has_regularization_pose) return;
  const double scale = p->regularization_scale_factor, w = (double)corr;
  const double dx = p->regularization_translation[0] - pose[0];
  const double dy = p->regularization_translation[1] - pose[1];
  const double sy = std::sin(pose[5]), cy = std::cos(pose[5]);
  const double lon = dy * sy + dx * cy;
  *s = -scale * w * lon * lon;
  g[0] = scale * w * 2.0 * cy * lon;
  g[1] = scale * w * 2.0 * sy * lon;
  H[0] = -scale * w * 2.0 * cy * cy;
  H[1] = -scale * w * 2.0 * cy * sy;
  H[6] = H[1];
  H[7] = -scale * w * 2.0 * sy * sy;
}

// count_oscillation (optimization/oscillation.rs:50-99).
uint32_t oscillation(const double* t, size_t k) {
  if (k < 3) return 0;
  uint32_t cnt = 0, mx = 0;
  for (size_t i = 2; i < k; ++i) {
    double cur[3], prev[3];
    for (int a = 0; a < 3; ++a) {
      cur[a] = t[3 * i + a] - t[3 * (i - 1) + a];
      prev[a] = t[3 * (i - 1) + a] - t[3 * (i - 2) + a];
    }
    const double cn = std::sqrt(cur[0] * cur[0] + cur[1] * cur[1] + cur[2] * cur[2]);
    const double pn = std::sqrt(prev[0] * prev[0] + prev[1] * prev[1] + prev[2] * prev[2]);
    if (cn < 1e-10 || pn < 1e-10) {
      cnt = 0;
      continue;
    }
    double cosv = 0.0;
    for (int a = 0; a < 3; ++a) cosv += (cur[a] / cn) * (prev[a] / pn);
    if (cosv < -0.9)
      ++cnt;
    else
      cnt = 0;
    mx = std::max(mx, cnt);
  }
  return mx;
}

double point_score(const double* x, const Voxel& v, const Gauss& gs, bool fp64) {
  // compute_point_voxel_score (scoring/nvtl.rs:196-206), compute_point_score (metrics.rs:189-203)
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = x[k] - (fp64 ? v.mean64[k] : (double)v.mean[k]);
  double Cd[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += (fp64 ? v.inv64[i * 3 + k] : (double)v.inv[i * 3 + k]) * d[k];
    Cd[i] = s;
  }
  const double m = d[0] * Cd[0] + d[1] * Cd[1] + d[2] * Cd[2];
  return -gs.d1 * std::exp(-gs.d2 * m / 2.0);
}

// compute_nvtl_simple (scoring/nvtl.rs:211-268).
double nvtl_impl(const oracle_grid* g, const float* xyz, size_t n, const double* pose, const oracle_params* p) {
  if (n == 0) return 0.0;
  const Gauss gs = make_gauss(p->resolution, p->outlier_ratio);
  double q[4];
  quat_from_euler(pose[3], pose[4], pose[5], q);
  std::vector<int32_t> nb;
  double total = 0.0;
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) {
    const double pt[3] = {(double)xyz[3 * i], (double)xyz[3 * i + 1], (double)xyz[3 * i + 2]};
    double x[3];
    iso_transform(q, pose, pt, x);
    const float xf[3] = {(float)x[0], (float)x[1], (float)x[2]};
    radius_search(g, xf, nb);  // scoring always uses the radius search (nvtl.rs:241, metrics.rs:94)
    if (nb.empty()) continue;
    double mx = -std::numeric_limits<double>::infinity();
    for (int32_t vi : nb) mx = std::max(mx, point_score(x, g->voxels[vi], gs, p->fp64_voxels != 0));
    total += mx;
    ++count;
  }
  return count ? total / (double)count : 0.0;
}

}  // namespace

extern "C" {

void oracle_gauss(double resolution, double outlier_ratio, double* d1, double* d2) {
  const Gauss g = make_gauss(resolution, outlier_ratio);
  *d1 = g.d1;
  *d2 = g.d2;
}

void oracle_point_keys(const float* xyz, size_t n, float resolution, int32_t* keys) {
  for (size_t i = 0; i < n; ++i) {
    const Coord c = key_of(xyz + 3 * i, resolution);
    keys[3 * i] = c.x;
    keys[3 * i + 1] = c.y;
    keys[3 * i + 2] = c.z;
  }
}

oracle_grid* oracle_grid_build(const float* xyz, size_t n, float resolution, uint32_t min_points,
                               float eigen_ratio) {
  // build_voxel_grid_cpu (voxel_grid/cpu.rs:51-75): serial f64 accumulation in
  // point order (cpu.rs:27-32), then per-voxel statistics.
  std::unordered_map<Coord, Acc, CoordHash> acc;
  acc.reserve(n / 8 + 16);
  for (size_t i = 0; i < n; ++i) {
    const float* p = xyz + 3 * i;
    Acc& a = acc[key_of(p, resolution)];
    const double v[3] = {(double)p[0], (double)p[1], (double)p[2]};
    for (int k = 0; k < 3; ++k) a.sum[k] += v[k];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) a.sq[r * 3 + c] += v[r] * v[c];
    a.count += 1;
  }
  auto* g = new oracle_grid();
  g->resolution = resolution;
  g->voxels.reserve(acc.size());
  for (auto& kv : acc) {
    Voxel v;
    if (from_statistics(kv.second, min_points, eigen_ratio, &v)) {
      v.coord = kv.first;
      g->voxels.push_back(v);
    }
  }
  std::sort(g->voxels.begin(), g->voxels.end(),
            [](const Voxel& a, const Voxel& b) { return a.coord < b.coord; });
  const double res = (double)resolution;
  for (size_t i = 0; i < g->voxels.size(); ++i) {
    const Voxel& v = g->voxels[i];
    g->by_coord[v.coord] = (int32_t)i;
    const Coord b{sat_i32d(std::floor((double)v.mean[0] / res)), sat_i32d(std::floor((double)v.mean[1] / res)),
                  sat_i32d(std::floor((double)v.mean[2] / res))};
    g->bins[b].push_back((int32_t)i);
  }
  return g;
}

void oracle_grid_free(oracle_grid* g) { delete g; }

size_t oracle_grid_size(const oracle_grid* g) { return g->voxels.size(); }

void oracle_grid_export(const oracle_grid* g, int32_t* coords, float* means, float* covs, float* invs,
                        float* axes, uint32_t* counts, double* means64, double* invs64) {
  for (size_t i = 0; i < g->voxels.size(); ++i) {
    const Voxel& v = g->voxels[i];
    if (coords) {
      coords[3 * i] = v.coord.x;
      coords[3 * i + 1] = v.coord.y;
      coords[3 * i + 2] = v.coord.z;
    }
    if (means) std::memcpy(means + 3 * i, v.mean, sizeof(v.mean));
    if (covs) std::memcpy(covs + 9 * i, v.cov, sizeof(v.cov));
    if (invs) std::memcpy(invs + 9 * i, v.inv, sizeof(v.inv));
    if (axes) std::memcpy(axes + 3 * i, v.axis, sizeof(v.axis));
    if (counts) counts[i] = v.count;
    if (means64) std::memcpy(means64 + 3 * i, v.mean64, sizeof(v.mean64));
    if (invs64) std::memcpy(invs64 + 9 * i, v.inv64, sizeof(v.inv64));
  }
}

void oracle_search(const oracle_grid* g, int32_t mode, const float* q, size_t n, uint32_t* out_count,
                   int32_t* out_idx, uint32_t max_k) {
  std::vector<int32_t> nb;
  for (size_t i = 0; i < n; ++i) {
    if (mode == ORACLE_SEARCH_CELL27) {
      const double x[3] = {(double)q[3 * i], (double)q[3 * i + 1], (double)q[3 * i + 2]};
      cell27_search(g, x, nb);
    } else {
      search(g, mode, q + 3 * i, nb);
    }
    out_count[i] = (uint32_t)nb.size();
    for (uint32_t k = 0; k < max_k; ++k) out_idx[i * max_k + k] = k < nb.size() ? nb[k] : -1;
  }
}

void oracle_derivatives(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                        const oracle_params* p, int32_t compute_hessian, double* score, double* grad6,
                        double* hess36, uint32_t* num_corr) {
  const Gauss gs = make_gauss(p->resolution, p->outlier_ratio);
  const Deriv d = derivatives(g, xyz, n, pose, gs, compute_hessian != 0, p->search_mode, p->distance_metric,
                              p->fp64_voxels != 0);
  *score = d.score;
  for (int i = 0; i < 6; ++i) grad6[i] = d.g[i];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) hess36[i * 6 + j] = d.H[i][j];
  *num_corr = d.corr;
}

int32_t oracle_newton_step(const double grad6[6], const double hess36[36], double min_eigenvalue, double tolerance,
                           double delta6[6]) {
  return newton_step_nd(grad6, hess36, min_eigenvalue, tolerance, delta6) ? 1 : 0;
}

void oracle_sym_eigen3(const double a[9], double evals[3], double evecs[9]) { jacobi_eigen<3>(a, evals, evecs); }
void oracle_sym_eigen6(const double a[36], double evals[6], double evecs[36]) { jacobi_eigen<6>(a, evals, evecs); }

void oracle_align(const oracle_grid* g, const float* xyz, size_t n, const double pose_in[6], const oracle_params* p,
                  oracle_result* out) {
  // NdtOptimizer::align (optimization/solver.rs:135-305).
  const Gauss gs = make_gauss(p->resolution, p->outlier_ratio);
  double pose[6];
  std::memcpy(pose, pose_in, sizeof(pose));
  double best_score = -std::numeric_limits<double>::infinity();
  double best_pose[6];
  std::memcpy(best_pose, pose, sizeof(pose));
  double last_H[36] = {};
  uint32_t last_corr = 0;
  std::vector<double> hist(pose, pose + 3);  // pose_history translations (solver.rs:150-152)
  auto finish = [&](const double* fpose, int status, double score, int iters, const double* H, uint32_t corr) {
    std::memcpy(out->pose, fpose, sizeof(double) * 6);
    out->status = status;
    out->iterations = iters;
    out->score = score;
    out->transform_probability = n > 0 ? score / (double)n : 0.0;  // solver.rs:707-713
    out->nvtl = nvtl_impl(g, xyz, n, fpose, p);                      // solver.rs:566-574
    std::memcpy(out->hessian, H, sizeof(double) * 36);
    out->num_correspondences = corr;
    out->oscillation_count = oscillation(hist.data(), hist.size() / 3);
  };
  for (uint32_t it = 0; it < p->max_iterations; ++it) {
    const Deriv d = derivatives(g, xyz, n, pose, gs, true, p->search_mode, p->distance_metric, p->fp64_voxels != 0);
    if (d.corr < 10) {  // min_correspondences (solver.rs:58,168-174)
      if (it == 0) {
        std::memset(out, 0, sizeof(*out));
        std::memcpy(out->pose, pose_in, sizeof(double) * 6);
        out->status = 2;  // NdtResult::no_correspondences (types.rs:133-145)
        return;
      }
      break;
    }
    last_corr = d.corr;
    double rs, rg[6], rH[36];
    regularization(p, pose, d.corr, &rs, rg, rH);
    double score = d.score + rs;
    double grad[6], H[36];
    for (int i = 0; i < 6; ++i) grad[i] = d.g[i] + rg[i];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) H[i * 6 + j] = d.H[i][j] + rH[i * 6 + j];
    if (p->noise > 0.0) {
      auto u = [&](uint64_t k) {
        uint64_t x = ((uint64_t)p->noise_seed << 40) ^ ((uint64_t)it << 20) ^ k;
        x += 0x9E3779B97F4A7C15ull;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
      };
      // Rounding in a sum is relative to the size of its terms, not of the
      // result: scale by the largest entry so that small (cancelled) entries
      // move as much as rounding would move them.
      double gmax = 0.0, hmax = 0.0;
      for (int i = 0; i < 6; ++i) gmax = std::max(gmax, std::fabs(grad[i]));
      for (int i = 0; i < 36; ++i) hmax = std::max(hmax, std::fabs(H[i]));
      for (int i = 0; i < 6; ++i) grad[i] += p->noise * gmax * u(i);
      for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {  // keep the Hessian symmetric
          const double e = p->noise * hmax * u(64 + i * 6 + j);
          H[i * 6 + j] += e;
          if (j != i) H[j * 6 + i] += e;
        }
      score *= 1.0 + p->noise * u(128);  // the best-pose choice at max iterations compares scores
    }
    std::memcpy(last_H, H, sizeof(H));
    if (score > best_score) {  // solver.rs:189-193
      best_score = score;
      std::memcpy(best_pose, pose, sizeof(pose));
    }
    double delta[6];
    if (!newton_step_nd(grad, H, -1e-3, 1e-10, delta)) {  // solver.rs:196-227
      finish(best_pose, 3, best_score, (int)it, last_H, last_corr);
      return;
    }
    double norm = 0.0;
    for (double v : delta) norm += v * v;
    norm = std::sqrt(norm);
    if (norm < p->trans_epsilon) {  // solver.rs:229-253: before stepping
      finish(pose, 0, score, (int)it + 1, H, d.corr);
      return;
    }
    double dir[6];
    for (int i = 0; i < 6; ++i) dir[i] = delta[i] / norm;
    double gd = 0.0;
    for (int i = 0; i < 6; ++i) gd += grad[i] * dir[i];
    if (gd <= 0.0)
      for (double& v : dir) v = -v;  // solver.rs:259-264
    double step;
    if (p->use_line_search) {
      // line_search_with_regularization (solver.rs:640-699)
      double init_d = 0.0;
      for (int i = 0; i < 6; ++i) init_d += grad[i] * dir[i];
      if (init_d <= 0.0) {
        step = p->step_size;
      } else {
        auto phi = [&](double a, double* s, double* dd) {
          double tp[6];
          for (int i = 0; i < 6; ++i) tp[i] = pose[i] + dir[i] * a;
          const Deriv r = derivatives(g, xyz, n, tp, gs, false, p->search_mode, p->distance_metric,
                                      p->fp64_voxels != 0);
          double s2, g2[6], H2[36];
          regularization(p, tp, r.corr, &s2, g2, H2);
          *s = r.score + s2;
          double v = 0.0;
          for (int i = 0; i < 6; ++i) v += (r.g[i] + g2[i]) * dir[i];
          *dd = v;
        };
        double st, fv;
        int32_t conv;
        uint32_t its;
        more_thuente(phi, score, init_d, p->step_size, 0.1, 1e-9, 1e-4, 0.9, 10, &st, &conv, &its, &fv);
        step = conv ? st : 0.1;  // MoreThuenteConfig::default step_max (more_thuente.rs:35-45)
      }
    } else {
      step = std::min(norm, p->step_size);  // solver.rs:278-280
    }
    for (int i = 0; i < 6; ++i) pose[i] = pose[i] + dir[i] * step;  // apply_pose_delta (types.rs:183-197)
    hist.insert(hist.end(), pose, pose + 3);
  }
  finish(best_pose, 1, best_score, (int)p->max_iterations, last_H, last_corr);  // solver.rs:286-304
}

double oracle_nvtl(const oracle_grid* g, const float* xyz, size_t n, const double pose[6], const oracle_params* p) {
  return nvtl_impl(g, xyz, n, pose, p);
}

double oracle_transform_probability(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                                    const oracle_params* p) {
  // compute_transform_probability (scoring/metrics.rs:58-135): sum / #pairs.
  if (n == 0) return 0.0;
  const Gauss gs = make_gauss(p->resolution, p->outlier_ratio);
  double q[4];
  quat_from_euler(pose[3], pose[4], pose[5], q);
  std::vector<int32_t> nb;
  double total = 0.0;
  size_t corr = 0;
  for (size_t i = 0; i < n; ++i) {
    const double pt[3] = {(double)xyz[3 * i], (double)xyz[3 * i + 1], (double)xyz[3 * i + 2]};
    double x[3];
    iso_transform(q, pose, pt, x);
    const float xf[3] = {(float)x[0], (float)x[1], (float)x[2]};
    radius_search(g, xf, nb);  // scoring always uses the radius search (nvtl.rs:241, metrics.rs:94)
    for (int32_t vi : nb) {
      total += point_score(x, g->voxels[vi], gs, p->fp64_voxels != 0);
      ++corr;
    }
  }
  return corr ? total / (double)corr : 0.0;
}

void oracle_per_point_scores(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                             const oracle_params* p, double* out) {
  // compute_per_point_scores (scoring/metrics.rs:137-178): per-point sum.
  const Gauss gs = make_gauss(p->resolution, p->outlier_ratio);
  double q[4];
  quat_from_euler(pose[3], pose[4], pose[5], q);
  std::vector<int32_t> nb;
  for (size_t i = 0; i < n; ++i) {
    const double pt[3] = {(double)xyz[3 * i], (double)xyz[3 * i + 1], (double)xyz[3 * i + 2]};
    double x[3];
    iso_transform(q, pose, pt, x);
    const float xf[3] = {(float)x[0], (float)x[1], (float)x[2]};
    radius_search(g, xf, nb);  // scoring always uses the radius search (nvtl.rs:241, metrics.rs:94)
    double s = 0.0;
    for (int32_t vi : nb) s += point_score(x, g->voxels[vi], gs, p->fp64_voxels != 0);
    out[i] = s;
  }
}

double oracle_nvtl_cell27(const oracle_grid* g, const float* xyz, size_t n, const double pose[6],
                          const oracle_params* p) {
  // compute_nvtl (scoring/nvtl.rs:83-150), the evaluate_nvtl_batch CPU path (ndt.rs:879-888).
  if (n == 0) return 0.0;
  const Gauss gs = make_gauss(p->resolution, p->outlier_ratio);
  double q[4];
  quat_from_euler(pose[3], pose[4], pose[5], q);
  std::vector<int32_t> nb;
  double total = 0.0;
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) {
    const double pt[3] = {(double)xyz[3 * i], (double)xyz[3 * i + 1], (double)xyz[3 * i + 2]};
    double x[3];
    iso_transform(q, pose, pt, x);
    cell27_search(g, x, nb);
    if (nb.empty()) continue;
    double mx = -std::numeric_limits<double>::infinity();
    for (int32_t vi : nb) mx = std::max(mx, point_score(x, g->voxels[vi], gs, p->fp64_voxels != 0));
    total += mx;
    ++count;
  }
  return count ? total / (double)count : 0.0;
}

void oracle_more_thuente(oracle_score_fn fn, void* user, double initial_score, double initial_derivative,
                         double initial_step, double step_max, double step_min, double mu, double nu,
                         uint32_t max_iterations, double* step_length, int32_t* converged, uint32_t* iterations,
                         double* final_value) {
  auto f = [&](double a, double* s, double* d) { fn(a, s, d, user); };
  more_thuente(f, initial_score, initial_derivative, initial_step, step_max, step_min, mu, nu, max_iterations,
               step_length, converged, iterations, final_value);
}

uint32_t oracle_oscillation(const double* translations, size_t k) { return oscillation(translations, k); }

void oracle_regularization(const oracle_params* p, const double pose[6], uint32_t corr, double* score,
                           double grad6[6], double hess36[36]) {
  regularization(p, pose, corr, score, grad6, hess36);
}

void oracle_laplace_xy(const double hess36[36], double cov4[4]) {
  // Matrix2::try_inverse fails only for a zero determinant.
  const double a = hess36[0], b = hess36[1], c = hess36[6], d = hess36[7];
  const double det = a * d - b * c;
  if (det == 0.0) {
    cov4[0] = 1.0; cov4[1] = 0.0; cov4[2] = 0.0; cov4[3] = 1.0;
    return;
  }
  cov4[0] = -(d / det);
  cov4[1] = -(-b / det);
  cov4[2] = -(-c / det);
  cov4[3] = -(a / det);
}

void oracle_pose_to_matrix(const double pose[6], double m12[12]) {
  // pose_vector_to_isometry (optimization/types.rs:156-161) then
  // UnitQuaternion::to_rotation_matrix.
  double q[4];
  quat_from_euler(pose[3], pose[4], pose[5], q);
  const double w = q[0], i = q[1], j = q[2], k = q[3];
  const double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
  const double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0, ik = i * k * 2.0, jk = j * k * 2.0,
               wi = w * i * 2.0;
  const double R[9] = {ww + ii - jj - kk, ij - wk, wj + ik, wk + ij, ww - ii + jj - kk, jk - wi,
                       ik - wj, wi + jk, ww - ii - jj + kk};
  for (int r = 0; r < 3; ++r) {
    for (int c = 0; c < 3; ++c) m12[r * 4 + c] = R[r * 3 + c];
    m12[r * 4 + 3] = pose[r];
  }
}

void oracle_matrix_to_pose(const double m12[12], double pose[6]) {
  // isometry_to_pose_vector (optimization/types.rs:164-174) with
  // Rotation3::euler_angles (Slabaugh's method as nalgebra implements it).
  const double r00 = m12[0], r01 = m12[1], r02 = m12[2];
  const double r10 = m12[4], r20 = m12[8], r21 = m12[9], r22 = m12[10];
  double roll, pitch, yaw;
  if (std::fabs(r20) < 1.0) {
    pitch = -std::asin(r20);
    const double c = std::cos(pitch);
    roll = std::atan2(r21 / c, r22 / c);
    yaw = std::atan2(r10 / c, r00 / c);
  } else if (r20 <= -1.0) {
    roll = std::atan2(r01, r02);
    pitch = M_PI / 2.0;
    yaw = 0.0;
  } else {
    roll = -std::atan2(-r01, -r02);
    pitch = -M_PI / 2.0;
    yaw = 0.0;
  }
  pose[0] = m12[3];
  pose[1] = m12[7];
  pose[2] = m12[11];
  pose[3] = roll;
  pose[4] = pitch;
  pose[5] = yaw;
}

// filter_points_cpu (src/ndt_cuda/src/filtering/cpu.rs:7-58): distance test
// on x*x + y*y + z*z in f32 (every operation rounded), then the z test; the
// kept points in input order.  With downsample != 0 and resolution > 0,
// voxel_downsample_cpu (:67-100): cell floor(p * (1 / res)) in f32 with a
// saturating cast, f64 sums in input order, centroid (sum / n) as f32.  The
// reference emits the cells in HashMap order; here they come out in
// (x, y, z) lexicographic order of the cell.  counts: kept (before
// downsampling), removed by distance, by z, by downsampling.
size_t oracle_filter_points(const float* xyz, size_t n, float min_distance, float max_distance, float min_z,
                            float max_z, int32_t downsample, float resolution, float* out, uint64_t counts[4]) {
  const float min_sq = min_distance * min_distance, max_sq = max_distance * max_distance;
  std::vector<float> kept;
  kept.reserve(3 * n);
  uint64_t rd = 0, rz = 0;
  for (size_t i = 0; i < n; ++i) {
    const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    const float d = x * x + y * y + z * z;  // built with -ffp-contract=off
    if (d < min_sq || d > max_sq) {
      ++rd;
      continue;
    }
    if (z < min_z || z > max_z) {
      ++rz;
      continue;
    }
    kept.push_back(x);
    kept.push_back(y);
    kept.push_back(z);
  }
  const size_t k = kept.size() / 3;
  counts[0] = k;
  counts[1] = rd;
  counts[2] = rz;
  counts[3] = 0;
  if (!downsample || resolution <= 0.0f || k == 0) {
    std::copy(kept.begin(), kept.end(), out);
    return k;
  }
  const float inv = 1.0f / resolution;
  struct Acc {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    size_t c = 0;
  };
  std::map<std::tuple<int32_t, int32_t, int32_t>, Acc> cells;
  for (size_t i = 0; i < k; ++i) {
    const float* q = &kept[3 * i];
    Acc& a = cells[std::make_tuple(sat_i32(std::floor(q[0] * inv)), sat_i32(std::floor(q[1] * inv)),
                                   sat_i32(std::floor(q[2] * inv)))];
    a.sx += (double)q[0];
    a.sy += (double)q[1];
    a.sz += (double)q[2];
    ++a.c;
  }
  size_t o = 0;
  for (const auto& kv : cells) {
    const double c = (double)kv.second.c;
    out[3 * o] = (float)(kv.second.sx / c);
    out[3 * o + 1] = (float)(kv.second.sy / c);
    out[3 * o + 2] = (float)(kv.second.sz / c);
    ++o;
  }
  counts[3] = k - o;
  return o;
}

}  // extern "C"
